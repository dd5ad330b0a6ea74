// Batched KNN: one workgroup runs four queries through every level in one launch.
//
// Per level the workgroup walks the level's points in position order (each
// point's reconstructed coordinates are read once for the four queries) and,
// for every point whose parent survived the previous level, computes
// lb = sqrt_rd(|q - c|^2) - r exactly as the per-query path does (KNN.cu:9-24).
// STEP (KNN.cu:75-96) needs only the clusters that sort first by (lb, id) until
// the running child count reaches K: they are kept in an LDS buffer of the best
// M = K + (empty clusters) candidates, filtered by the current M-th key and
// re-sorted by a bitonic network when the buffer fills.  Survivors (lb <= ub)
// of a middle level are flagged per query for the next level.
//
// The last level is fused with the exact top-K refine.  Its pass 1 keeps at
// least `seedMin` clusters; the surviving prefix of that sorted list (never
// shorter than phase 1, whose clusters hold at least K points) is expanded
// first, and the K-th smallest squared distance among those children gives a
// certified radius rho.  Every child x of a cluster c satisfies
// |q-x| >= |q-c| - r_c, so in pass 2 a surviving cluster is expanded only if
// its lower bound is within rho (plus a relative margin of 1e-9, far above the
// rounding error of the bounds).  The top-K is then the same set the
// per-query path computes over all survivors' children, ordered by
// (squared distance, id).  Clusters queued for expansion are processed as one
// flat (cluster, child) index space so that every thread of a step scores a
// point, whatever the cluster sizes.
#include <algorithm>

#include "../../../include/tblaeq_learn.h"
#include "../batch_internal.hpp"

namespace tblaeq {

constexpr int kKnnQB = 4;        // queries per workgroup
constexpr int kKnnCap = 512;     // LDS candidates per query
constexpr int kKnnMaxM = kKnnCap - 256;
constexpr int kKnnThreads = 256;
constexpr int kKnnList = 1280;   // clusters queued for expansion (flush at >= 256)
constexpr int kKnnFlush = 256;

constexpr int kKnnMaxLevels = 8;

struct KnnArgs {
    int D, L;
    uint64_t nq;
    const double* points;   // nq x D
    const uint64_t* Ks;     // nq
    uint64_t Kmax;
    uint64_t count[kKnnMaxLevels + 1];  // positions per level
    const double* coords[kKnnMaxLevels + 1];
    const uint64_t* ids[kKnnMaxLevels + 1];
    const uint64_t* parentPos[kKnnMaxLevels];
    const double* radiusPos[kKnnMaxLevels];
    const uint64_t* weightPos[kKnnMaxLevels];
    const double* radiusById[kKnnMaxLevels];
    const uint64_t* weightById[kKnnMaxLevels];
    const uint64_t* colRes[kKnnMaxLevels];
    uint64_t emptyCols[kKnnMaxLevels];
    uint8_t* flags;         // per query survivor flags of the middle levels
    uint64_t flagStride;
    uint64_t flagOff[kKnnMaxLevels];  // offset of level l's flags in a query's area
    uint32_t seedMin;       // clusters expanded before the certified radius is taken
    // cluster-major mode (pairs != nullptr): pass 2 of the last level emits
    // (cluster, query) pairs instead of expanding the clusters itself
    uint64_t* pairs;        // (cluster id << kQueryBits) | q
    uint64_t pairCap;
    uint64_t* pairCount;
    uint64_t* cnt;          // pairs per last-level cluster
    uint64_t* overflow;
    uint64_t* thr;          // nq x 2: (key, id) of the K-th seed child, ~0 if fewer than K
    // split mode: levels 0..L-2 here, the last level cluster-major elsewhere.
    // Survivors of level L-2 become (cluster, query) pairs and tau[q] bounds
    // the best M (lb, id) keys of the last level.
    int split;
    uint64_t* tau;          // nq x 2
    uint64_t* work;         // distance / bound evaluations of the flat expansions
    uint64_t* outIds;       // nq x Kmax
    double* outD2;
    uint64_t* outN;         // nq
    uint64_t* survStat;     // [nq * L] optional
    uint64_t* fineStat;     // [nq] optional
};

struct KnnShared {
    uint64_t bk[kKnnQB][kKnnCap];
    uint64_t bv[kKnnQB][kKnnCap];
    double qp[kKnnQB][32];
    uint32_t cnt[kKnnQB];
    uint64_t thrK[kKnnQB], thrV[kKnnQB];
    uint64_t bndK[kKnnQB], bndV[kKnnQB];
    uint32_t M[kKnnQB];
    double ub[kKnnQB];
    uint64_t K[kKnnQB];
    uint32_t i0[kKnnQB];
    uint64_t lstart[kKnnList];      // first finest position of a queued cluster
    uint64_t lpre[kKnnList + 1];    // child counts, then their exclusive prefix
    uint8_t listQ[kKnnList];
    uint32_t listN;
    uint64_t wsum[kKnnThreads / kWave];
    uint32_t seedN[kKnnQB];
    uint64_t surv[kKnnQB];
    uint64_t fine[kKnnQB];
    uint64_t tot[kKnnQB];
    uint32_t need[2][kKnnQB];       // candidates of the current round that beat the threshold (by round parity)
    uint32_t fill[2][kKnnQB];       // slots taken by the current round's inserts
};

__device__ __forceinline__ bool lt_pair(uint64_t ak, uint64_t av, uint64_t bk, uint64_t bv) {
    return ak < bk || (ak == bk && av < bv);
}

// Bitonic sort of the first 64 * R entries of one buffer (entries from cnt
// on read as sentinels), in registers of the calling wave; written back.
template <int R>
__device__ __forceinline__ void knn_sort_regs(uint64_t* bk, uint64_t* bv, uint32_t cnt, int lane) {
    constexpr int N = R * kWave;
    uint64_t K[R], V[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t e = static_cast<uint32_t>(r * kWave + lane);
        K[r] = e < cnt ? bk[e] : ~0ull;
        V[r] = e < cnt ? bv[e] : ~0ull;
    }
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                const int jr = j / kWave;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r & jr) continue;
                    const int r2 = r | jr;
                    const bool up = ((r * kWave + lane) & k) == 0;
                    const bool hiLess = lt_pair(K[r2], V[r2], K[r], V[r]);
                    if (hiLess == up) {
                        const uint64_t tk = K[r], tv = V[r];
                        K[r] = K[r2];
                        V[r] = V[r2];
                        K[r2] = tk;
                        V[r2] = tv;
                    }
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t pk = __shfl_xor(K[r], j, kWave), pv = __shfl_xor(V[r], j, kWave);
                    const bool up = ((r * kWave + lane) & k) == 0;
                    const bool pLess = lt_pair(pk, pv, K[r], V[r]);
                    // the lower index of a pair keeps the smaller entry when up
                    if (pLess == (lower == up)) {
                        K[r] = pk;
                        V[r] = pv;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        bk[r * kWave + lane] = K[r];
        bv[r * kWave + lane] = V[r];
    }
}

static_assert(kKnnQB == kKnnThreads / kWave, "one wave per query buffer");
// Sorts every query's buffer (entries beyond cnt are treated as sentinels) and
// keeps its first M entries; updates the insertion threshold.  Wave i sorts
// query i's buffer in registers with a bitonic network over the smallest
// power of two (at least 64) that holds its cnt entries: lane l holds entries
// r * 64 + l; partners 64 or more apart are in the same lane, nearer ones are
// exchanged with __shfl_xor.  Workgroup barriers at the start and the end.
__device__ void knn_compact(KnnShared& S, int tid) {
    // every wave has read the counts for its own decisions before any is rewritten
    __syncthreads();
    const int lane = tid & (kWave - 1), qi = tid / kWave;
    uint64_t* bk = S.bk[qi];
    uint64_t* bv = S.bv[qi];
    const uint32_t cnt = S.cnt[qi];
    {
        static_assert(kKnnCap == 8 * kWave, "sort sizes");
        if (cnt <= static_cast<uint32_t>(kWave)) knn_sort_regs<1>(bk, bv, cnt, lane);
        else if (cnt <= 2u * kWave) knn_sort_regs<2>(bk, bv, cnt, lane);
        else if (cnt <= 4u * kWave) knn_sort_regs<4>(bk, bv, cnt, lane);
        else knn_sort_regs<8>(bk, bv, cnt, lane);
        wave_lds_sync();
        if (lane == 0) {
            const uint32_t m = S.M[qi];
            if (S.cnt[qi] > m) S.cnt[qi] = m;
            if (m > 0 && S.cnt[qi] >= m) {
                S.thrK[qi] = bk[m - 1];
                S.thrV[qi] = bv[m - 1];
            } else {
                S.thrK[qi] = ~0ull;
                S.thrV[qi] = ~0ull;
            }
        }
    }
    __syncthreads();
}

// One round of insertions, up to one candidate per query slot per thread
// (have[i], key[i], val[i]): the candidates that beat their query's threshold
// are counted first, the buffers are compacted only when one of them would
// overflow (the tightened thresholds then apply), and every remaining
// candidate gets a slot.  S.cnt is not written between the first barrier and
// the decision of any wave: slots are taken on a separate counter (fill) and
// added to cnt by the first threads after the second barrier, which the next
// round's first barrier orders before any read; after the last round of a
// sequence the caller places one more barrier before S.cnt is read again.
// Rounds alternate the parity `par` of their counters; the counters are zero
// between rounds.  Contains workgroup barriers: every thread calls it.
__device__ void knn_round(KnnShared& S, int tid, const bool (&have)[kKnnQB], const uint64_t (&key)[kKnnQB],
                          const uint64_t (&val)[kKnnQB], int par) {
    bool pass[kKnnQB];
#pragma unroll
    for (int i = 0; i < kKnnQB; ++i) {
        pass[i] = have[i] && S.M[i] > 0 && lt_pair(key[i], val[i], S.thrK[i], S.thrV[i]);
        if (pass[i]) atomicAdd(&S.need[par][i], 1u);
    }
    __syncthreads();
    bool over = false;
#pragma unroll
    for (int i = 0; i < kKnnQB; ++i) over = over || S.cnt[i] + S.need[par][i] > static_cast<uint32_t>(kKnnCap);
    if (over) {
        knn_compact(S, tid);
#pragma unroll
        for (int i = 0; i < kKnnQB; ++i) pass[i] = pass[i] && lt_pair(key[i], val[i], S.thrK[i], S.thrV[i]);
    }
#pragma unroll
    for (int i = 0; i < kKnnQB; ++i)
        if (pass[i]) {
            const uint32_t p = S.cnt[i] + atomicAdd(&S.fill[par][i], 1u);
            S.bk[i][p] = key[i];
            S.bv[i][p] = val[i];
        }
    __syncthreads();
    if (tid < kKnnQB) {
        S.cnt[tid] += S.fill[par][tid];
        S.fill[par][tid] = 0;
        S.need[par][tid] = 0;
    }
}

// knn_round with one candidate per thread, for query slot qi.
__device__ __forceinline__ void knn_round1(KnnShared& S, int tid, int qi, bool have, uint64_t key, uint64_t val,
                                           int par) {
    bool h[kKnnQB];
    uint64_t k[kKnnQB], v[kKnnQB];
#pragma unroll
    for (int i = 0; i < kKnnQB; ++i) {
        h[i] = have && i == qi;
        k[i] = key;
        v[i] = val;
    }
    knn_round(S, tid, h, k, v, par);
}

// Coordinates are loaded into registers once per point (all DM loads issued
// together) and padded with zeros beyond D; a padding dimension adds
// fma(0, 0, acc) = acc, so the sums equal the D-term fma chains exactly.
template <int DM>
__device__ __forceinline__ void load_row(const double* __restrict__ x, int D, double (&v)[DM]) {
#pragma unroll
    for (int d = 0; d < DM; ++d) v[d] = d < D ? x[d] : 0.0;
}

template <int DM>
__device__ __forceinline__ double d2_reg(const double (&x)[DM], const double* q) {
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < DM; ++d) {
        const double diff = q[d] - x[d];
        acc = __builtin_fma(diff, diff, acc);
    }
    return acc;
}

template <int DM>
__device__ __forceinline__ double lb_reg(const double (&x)[DM], const double* q, double r) {
    return sqrt_rd(d2_reg<DM>(x, q)) - r;
}

// Queues `count` points starting at position `start` for query slot i
// (caller guarantees a free slot).
__device__ __forceinline__ void knn_queue(KnnShared& S, int i, uint64_t start, uint64_t count) {
    const uint32_t p = atomicAdd(&S.listN, 1u);
    S.lstart[p] = start;
    S.lpre[p] = count;
    S.listQ[p] = static_cast<uint8_t>(i);
}

// Scores every queued point against its query and inserts it into that
// query's buffer; the queue is empty afterwards.  Key = okey(lb) with
// lb = sqrt_rd(d2) - r when `radius` is given, okey(d2) otherwise.  Entry e's
// points are the flat indices [lpre[e], lpre[e+1]) once the counts are
// scanned in place.
template <int DM>
__device__ void knn_flush(KnnShared& S, int tid, int D, const double* coords, const uint64_t* ids,
                          const double* radius, uint64_t* work) {
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const uint32_t n = S.listN;
    if (n == 0) return;
    constexpr uint32_t kPer = (kKnnList + kKnnThreads - 1) / kKnnThreads;
    const uint32_t b = static_cast<uint32_t>(tid) * kPer;
    uint64_t s = 0;
    for (uint32_t k = 0; k < kPer; ++k)
        if (b + k < n) s += S.lpre[b + k];
    uint64_t x = s;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) S.wsum[wid] = x;
    __syncthreads();
    uint64_t run = x - s;
    for (int w = 0; w < wid; ++w) run += S.wsum[w];
    const uint64_t total = run + s;
    for (uint32_t k = 0; k < kPer; ++k)
        if (b + k < n) {
            const uint64_t c = S.lpre[b + k];
            S.lpre[b + k] = run;
            run += c;
        }
    if (tid == kKnnThreads - 1) S.lpre[n] = total;
    __syncthreads();
    const uint64_t all = S.lpre[n];
    if (work && tid == 0) atomicAdd(reinterpret_cast<unsigned long long*>(work), all);
    // flat index f -> (query slot, point); the next round's point is located
    // and its loads issued before the current round's barriers
    auto locate = [&](uint64_t f, int& i, uint64_t& j) {
        uint32_t lo = 0, hi = n;  // lpre[lo] <= f < lpre[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S.lpre[mid] <= f) lo = mid;
            else hi = mid;
        }
        j = S.lstart[lo] + (f - S.lpre[lo]);
        i = S.listQ[lo];
    };
    int ni = 0;
    uint64_t nj = 0, nid = 0;
    double nx[DM], nr = 0.0;
    bool nhave = static_cast<uint64_t>(tid) < all;
#pragma unroll
    for (int d = 0; d < DM; ++d) nx[d] = 0.0;
    if (nhave) {
        locate(static_cast<uint64_t>(tid), ni, nj);
        load_row<DM>(coords + nj * D, D, nx);
        nid = ids[nj];
        if (radius) nr = radius[nj];
    }
    int par = 0;
    for (uint64_t f0 = 0; f0 < all; f0 += kKnnThreads, par ^= 1) {
        const bool have = nhave;
        const int i = ni;
        const uint64_t val = nid;
        const double r = nr;
        double xv[DM];
#pragma unroll
        for (int d = 0; d < DM; ++d) xv[d] = nx[d];
        const uint64_t fn = f0 + kKnnThreads + tid;
        nhave = fn < all;
        if (nhave) {
            locate(fn, ni, nj);
            load_row<DM>(coords + nj * D, D, nx);
            nid = ids[nj];
            if (radius) nr = radius[nj];
        }
        uint64_t key = 0;
        if (have) {
            const double d2 = d2_reg<DM>(xv, S.qp[i]);
            key = radius ? okey(sqrt_rd(d2) - r) : okey(d2);
        }
        knn_round1(S, tid, i, have, key, val, par);
    }
    if (tid == 0) S.listN = 0;
    __syncthreads();
}

// Queues the (cluster, query slot) pair of a surviving cluster.
__device__ __forceinline__ void knn_pair_queue(KnnShared& S, int i, uint64_t cid) {
    const uint32_t p = atomicAdd(&S.listN, 1u);
    S.lstart[p] = cid;
    S.listQ[p] = static_cast<uint8_t>(i);
}

// Appends the queued pairs to the global pair list (one atomic per flush); the
// queue is empty afterwards.  (Their per-cluster counts are added by the
// caller, one atomic per cluster and workgroup.)
__device__ void knn_emit_pairs(const KnnArgs& A, KnnShared& S, int tid, uint64_t q0) {
    const uint32_t n = S.listN;
    if (n == 0) return;
    if (tid == 0) S.wsum[0] = atomicAdd(reinterpret_cast<unsigned long long*>(A.pairCount), n);
    __syncthreads();
    const uint64_t base = S.wsum[0];
    for (uint32_t e = tid; e < n; e += kKnnThreads) {
        const uint64_t pos = base + e;
        if (pos < A.pairCap) A.pairs[pos] = (S.lstart[e] << kQueryBits) | (q0 + S.listQ[e]);
    }
    if (tid == 0 && base + n > A.pairCap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
    __syncthreads();
    if (tid == 0) S.listN = 0;
    __syncthreads();
}

// STEP (KNN.cu:75-96) over each query's sorted buffer, wave i for query i:
// phase 1 takes clusters while the exclusive child count is below K, ub is
// the largest lb + 2r of phase 1 (at least 0).  Sets S.ub and S.i0 (first
// entry after phase 1).
__device__ void knn_step(KnnShared& S, int tid, const uint64_t* weightById, const double* radiusById) {
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    if (wid < kKnnQB) {
        const int i = wid;
        const uint32_t m = S.cnt[i];
        const uint64_t K = S.K[i];
        uint64_t carry = 0;
        double best = 0.0;
        uint32_t i0 = m;
        bool done = false;
        for (uint32_t b = 0; b < m && !done; b += kWave) {
            const uint32_t e = b + lane;
            uint64_t w = 0;
            double c = -__builtin_inf();
            if (e < m) {
                const uint64_t id = S.bv[i][e];
                w = weightById[id];
                c = okey_inv(S.bk[i][e]) + 2.0 * radiusById[id];
            }
            uint64_t x = w;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint64_t y = __shfl_up(x, o, kWave);
                if (lane >= o) x += y;
            }
            const uint64_t ex = carry + x - w;
            const bool in1 = e < m && ex < K;
            best = fmax(best, wave_max_f64(in1 ? c : -__builtin_inf()));
            const uint64_t firstOut = __ballot(e < m && ex >= K);
            if (firstOut) {
                i0 = b + (__ffsll(static_cast<unsigned long long>(firstOut)) - 1);
                done = true;
            }
            carry += __shfl(x, kWave - 1, kWave);
        }
        if (lane == 0) {
            S.ub[i] = best;
            S.i0[i] = i0;
        }
    }
    __syncthreads();
}

// Seed clusters of the last level: the surviving (lb <= ub) prefix of the
// sorted buffer, never shorter than phase 1, up to seedMin entries.  Sets
// S.seedN and the last seed's (key, id) in S.bndK / S.bndV.
__device__ void knn_pick_seeds(KnnShared& S, int tid, uint64_t q0, uint64_t nq, uint32_t seedMin) {
    if (tid < kKnnQB) {
        const int i = tid;
        uint32_t ns = 0;
        if (q0 + i < nq && S.K[i] > 0) {
            const uint32_t m = S.cnt[i];
            uint32_t lim = S.i0[i] > seedMin ? S.i0[i] : seedMin;
            if (lim > m) lim = m;
            const uint64_t ubk = okey(S.ub[i]);
            uint32_t lo = S.i0[i] < lim ? S.i0[i] : lim, hi = lim;  // first entry above ub
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (S.bk[i][mid] <= ubk) lo = mid + 1;
                else hi = mid;
            }
            ns = lo;
        }
        S.seedN[i] = ns;
        S.bndK[i] = ns > 0 ? S.bk[i][ns - 1] : 0;
        S.bndV[i] = ns > 0 ? S.bv[i][ns - 1] : 0;
        if (i == 0) S.listN = 0;
    }
    __syncthreads();
}

// Expands the seeds (their finest children) into a fresh top-K buffer; the
// compacted buffer's threshold is the certified radius.
template <int DM>
__device__ void knn_expand_seeds(KnnShared& S, int tid, uint64_t q0, uint64_t nq, int D, const uint64_t* colRes,
                                 const double* coords, const uint64_t* ids, uint64_t* work) {
    for (int i = 0; i < kKnnQB; ++i)
        for (uint32_t e = tid; e < S.seedN[i]; e += kKnnThreads) {
            const uint64_t cid = S.bv[i][e];
            knn_queue(S, i, colRes[cid], colRes[cid + 1] - colRes[cid]);
        }
    __syncthreads();
    if (tid < kKnnQB) {
        S.M[tid] = static_cast<uint32_t>(q0 + tid < nq ? S.K[tid] : 0);
        S.cnt[tid] = 0;
        S.thrK[tid] = ~0ull;
        S.thrV[tid] = ~0ull;
    }
    __syncthreads();
    knn_flush<DM>(S, tid, D, coords, ids, nullptr, work);
    knn_compact(S, tid);
}

// Writes each query's sorted buffer as its (provisional) top-K.
__device__ void knn_write_topk(const KnnShared& S, int tid, uint64_t q0, uint64_t nq, uint64_t Kmax,
                               uint64_t* outIds, double* outD2, uint64_t* outN, uint64_t* thr) {
    for (int i = 0; i < kKnnQB; ++i) {
        const uint64_t q = q0 + i;
        if (q >= nq) continue;
        const uint32_t k = S.cnt[i];
        for (uint32_t e = tid; e < k; e += kKnnThreads) {
            outIds[q * Kmax + e] = S.bv[i][e];
            outD2[q * Kmax + e] = okey_inv(S.bk[i][e]);
        }
        if (tid == 0) {
            outN[q] = k;
            if (thr) {
                thr[2 * q] = S.thrK[i];
                thr[2 * q + 1] = S.thrV[i];
            }
        }
    }
}

template <int DM>
__global__ __launch_bounds__(kKnnThreads) void k_knn_batch(KnnArgs A) {
    __shared__ KnnShared S;
    if (threadIdx.x < 2 * kKnnQB) {  // zero between rounds from here on
        (&S.need[0][0])[threadIdx.x] = 0;
        (&S.fill[0][0])[threadIdx.x] = 0;
    }
    const int tid = threadIdx.x;
    const int D = A.D, L = A.L;
    const int Lend = A.split ? L - 1 : L;
    const uint64_t ngroups = (A.nq + kKnnQB - 1) / kKnnQB;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t q0 = g * kKnnQB;
        if (tid < kKnnQB) {
            const uint64_t q = q0 + tid;
            S.K[tid] = q < A.nq ? A.Ks[q] : 0;
            S.surv[tid] = 0;
            S.fine[tid] = 0;
        }
        if (tid == 0) S.listN = 0;
        for (int t = tid; t < kKnnQB * 32; t += kKnnThreads) {
            const int i = t / 32, d = t % 32;
            const uint64_t q = q0 + i;
            S.qp[i][d] = q < A.nq && d < D ? A.points[q * D + d] : 0.0;
        }
        __syncthreads();
        for (int l = 0; l < Lend; ++l) {
            const uint64_t n = A.count[l];
            const double* X = A.coords[l];
            const double* R = A.radiusPos[l];
            const uint64_t* I = A.ids[l];
            const uint64_t* Pp = A.parentPos[l];
            // ---- pass 1: best M candidates by (lb, id)
            if (tid < kKnnQB) {
                const uint64_t q = q0 + tid;
                // K + empty clusters bound phase 1 (one spare entry); the last
                // level keeps at least seedMin clusters for the certified radius
                uint64_t m = q < A.nq ? S.K[tid] + A.emptyCols[l] + 1 : 0;
                if (l == L - 1 && q < A.nq && S.K[tid] > 0 && m < A.seedMin) m = A.seedMin;
                if (m > static_cast<uint64_t>(kKnnMaxM)) m = kKnnMaxM;
                S.M[tid] = static_cast<uint32_t>(m);
                S.cnt[tid] = 0;
                S.thrK[tid] = ~0ull;
                S.thrV[tid] = ~0ull;
            }
            __syncthreads();
            int par = 0;
            for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads, par ^= 1) {
                const uint64_t j = j0 + tid;
                bool have[kKnnQB];
                uint64_t key[kKnnQB], val[kKnnQB];
#pragma unroll
                for (int i = 0; i < kKnnQB; ++i) {
                    have[i] = false;
                    key[i] = 0;
                    val[i] = 0;
                }
                if (j < n) {
                    double xv[DM];
                    load_row<DM>(X + j * D, D, xv);
                    const double r = R[j];
                    const uint64_t id = I[j];
#pragma unroll
                    for (int i = 0; i < kKnnQB; ++i) {
                        const uint64_t q = q0 + i;
                        if (q >= A.nq) continue;
                        if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                        have[i] = true;
                        key[i] = okey(lb_reg<DM>(xv, S.qp[i], r));
                        val[i] = id;
                    }
                }
                knn_round(S, tid, have, key, val, par);
            }
            __syncthreads();
            knn_compact(S, tid);
            knn_step(S, tid, A.weightById[l], A.radiusById[l]);
            if (A.split && l == L - 2) {
                // ---- tau: best M (lb, id) keys among the last-level children of
                // the first surviving clusters (in (lb, id) order) that hold at
                // least M of them -- an upper bound on the last level's M-th key
                if (tid < kKnnQB) {
                    const int i = tid;
                    const uint64_t q = q0 + i;
                    uint64_t m = q < A.nq && S.K[i] > 0 ? S.K[i] + A.emptyCols[L - 1] + 1 : 0;
                    if (m > 0 && m < A.seedMin) m = A.seedMin;
                    if (m > static_cast<uint64_t>(kKnnMaxM)) m = kKnnMaxM;
                    const uint64_t ubk = okey(S.ub[i]);
                    uint32_t e = 0;
                    uint64_t got = 0;
                    if (m > 0)
                        for (; e < S.cnt[i] && got < m && S.bk[i][e] <= ubk; ++e) got += A.weightById[l][S.bv[i][e]];
                    S.seedN[i] = e;  // clusters whose children are scored
                    S.M[i] = static_cast<uint32_t>(m);
                }
                __syncthreads();
                for (int i = 0; i < kKnnQB; ++i)
                    for (uint32_t e = tid; e < S.seedN[i]; e += kKnnThreads) {
                        const uint64_t cid = S.bv[i][e];
                        knn_queue(S, i, A.colRes[l][cid], A.colRes[l][cid + 1] - A.colRes[l][cid]);
                    }
                __syncthreads();
                if (tid < kKnnQB) {
                    S.cnt[tid] = 0;
                    S.thrK[tid] = ~0ull;
                    S.thrV[tid] = ~0ull;
                }
                __syncthreads();
                knn_flush<DM>(S, tid, D, A.coords[L - 1], A.ids[L - 1], A.radiusPos[L - 1], A.work);
                knn_compact(S, tid);
                if (tid < kKnnQB && q0 + tid < A.nq) {
                    A.tau[2 * (q0 + tid)] = S.thrK[tid];
                    A.tau[2 * (q0 + tid) + 1] = S.thrV[tid];
                }
                // ---- pass 2: survivors of level L-2 become (cluster, query) pairs
                for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads) {
                    const uint64_t j = j0 + tid;
                    if (j < n) {
                        double xv[DM];
                        load_row<DM>(X + j * D, D, xv);
                        const double r = R[j];
                        const uint64_t id = I[j];
                        uint64_t kept = 0;
                        for (int i = 0; i < kKnnQB; ++i) {
                            const uint64_t q = q0 + i;
                            if (q >= A.nq) continue;
                            if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                            if (!(lb_reg<DM>(xv, S.qp[i], r) <= S.ub[i])) continue;
                            atomicAdd(reinterpret_cast<unsigned long long*>(&S.surv[i]), 1ull);
                            knn_pair_queue(S, i, id);
                            ++kept;
                        }
                        // pairs of this cluster for the grouping scan, one atomic for the workgroup's queries
                        if (kept) atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[id]), kept);
                    }
                    // every thread reads the count before any queues again
                    __syncthreads();
                    const bool due = S.listN >= static_cast<uint32_t>(kKnnFlush);
                    __syncthreads();
                    if (due) knn_emit_pairs(A, S, tid, q0);
                }
                knn_emit_pairs(A, S, tid, q0);
                if (tid < kKnnQB && A.survStat && q0 + tid < A.nq) A.survStat[(q0 + tid) * L + l] = S.surv[tid];
                __syncthreads();
                continue;
            }
            if (l < L - 1) {
                // ---- pass 2: flag survivors (lb <= ub) of a middle level
                for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads) {
                    const uint64_t j = j0 + tid;
                    if (j < n) {
                        double xv[DM];
                        load_row<DM>(X + j * D, D, xv);
                        const double r = R[j];
                        const uint64_t id = I[j];
                        for (int i = 0; i < kKnnQB; ++i) {
                            const uint64_t q = q0 + i;
                            if (q >= A.nq) continue;
                            bool s = true;
                            if (l > 0) s = A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]] != 0;
                            if (s) s = lb_reg<DM>(xv, S.qp[i], r) <= S.ub[i];
                            A.flags[q * A.flagStride + A.flagOff[l] + id] = s ? 1 : 0;
                            if (s) atomicAdd(reinterpret_cast<unsigned long long*>(&S.surv[i]), 1ull);
                        }
                    }
                }
                __syncthreads();
                if (tid < kKnnQB && A.survStat && q0 + tid < A.nq) A.survStat[(q0 + tid) * L + l] = S.surv[tid];
                if (tid < kKnnQB) S.surv[tid] = 0;
                __syncthreads();
                continue;
            }
            // ---- last level (whole level in this kernel): seeds, certified radius
            knn_pick_seeds(S, tid, q0, A.nq, A.seedMin);
            knn_expand_seeds<DM>(S, tid, q0, A.nq, D, A.colRes[L - 1], A.coords[L], A.ids[L], A.work);
            // ---- pass 2 of the last level: other survivors within the certified radius
            for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads) {
                const uint64_t j = j0 + tid;
                if (j < n) {
                    double xv[DM];
                    load_row<DM>(X + j * D, D, xv);
                    const double r = R[j];
                    const uint64_t id = I[j];
                    uint64_t kept = 0;
                    for (int i = 0; i < kKnnQB; ++i) {
                        const uint64_t q = q0 + i;
                        if (q >= A.nq) continue;
                        if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                        const double lb = lb_reg<DM>(xv, S.qp[i], r);
                        if (!(lb <= S.ub[i])) continue;
                        atomicAdd(reinterpret_cast<unsigned long long*>(&S.surv[i]), 1ull);
                        atomicAdd(reinterpret_cast<unsigned long long*>(&S.fine[i]), A.weightPos[l][j]);
                        if (S.M[i] == 0) continue;
                        if (S.seedN[i] > 0 && !lt_pair(S.bndK[i], S.bndV[i], okey(lb), id)) continue;  // seed
                        const uint64_t tk = S.thrK[i];
                        if (tk != ~0ull) {
                            const double rho = __builtin_sqrt(okey_inv(tk));
                            if (lb > rho + 1e-9 * (1.0 + rho + r)) continue;
                        }
                        if (A.pairs) {
                            knn_pair_queue(S, i, id);
                            ++kept;
                        } else {
                            knn_queue(S, i, A.colRes[L - 1][id], A.colRes[L - 1][id + 1] - A.colRes[L - 1][id]);
                        }
                    }
                    if (kept) atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[id]), kept);
                }
                // every thread reads the count before any queues again
                __syncthreads();
                const bool due = S.listN >= static_cast<uint32_t>(kKnnFlush);
                __syncthreads();
                if (due) {
                    if (A.pairs) knn_emit_pairs(A, S, tid, q0);
                    else knn_flush<DM>(S, tid, D, A.coords[L], A.ids[L], nullptr, A.work);
                }
            }
            if (A.pairs) knn_emit_pairs(A, S, tid, q0);
            else knn_flush<DM>(S, tid, D, A.coords[L], A.ids[L], nullptr, A.work);
            knn_compact(S, tid);  // cluster-major: the seeds' top-K, refined later
            knn_write_topk(S, tid, q0, A.nq, A.Kmax, A.outIds, A.outD2, A.outN, A.thr);
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t q = q0 + i;
                if (q >= A.nq || tid != 0) continue;
                if (A.survStat) A.survStat[q * L + l] = S.surv[i];
                if (A.fineStat) A.fineStat[q] = S.fine[i];
            }
            __syncthreads();
        }
    }
}

// --------------------------------------------- cluster-major last level (split)
// Work item = (cluster p of level L-2, slice of <= 64 queries that kept p),
// wave w taking items w, w + waves, ...  The slice's queries first test p's
// enclosing ball and box; those that may still reach a child are staged in LDS
// (point and thresholds), and the wave then runs lane = child of p, one staged
// query at a time.  Three passes share the structure:
//   kCand   records (okey(lb), q, id) with (lb, id) <= tau[q]: a superset of
//           each query's best M clusters, selected exactly afterwards;
//   kLCount per cluster, the number of queries that must refine it
//           (survivor, not a seed, lb within the certified radius); kLStats
//           also counts the survivors (lb <= ub) and their points per query;
//   kLWrite the same test, writing those queries into per-cluster lists.
// lb is computed exactly (sqrt_rd, as everywhere) only where the squared
// distance does not already decide the test: with t the threshold and
// B = t + r + 1e-9(|t| + r), d2 > B^2 implies sqrt_rd(d2) - r > t by far more
// than the rounding of the bounds.
enum LastPass { kCand = 0, kLCount = 1, kLWrite = 2, kLStats = 3 };

struct KnnLastArgs {
    int D, L;
    uint64_t nq;
    const uint64_t* items;
    const uint64_t* nItems;
    uint64_t itemCap;
    const uint64_t* off;
    const uint32_t* qlist;
    uint64_t qcap;
    const uint64_t* colResPrev;  // P[L-2] column pointers
    const double* coords;        // level L-1 positions
    const double* radius;
    const uint64_t* ids;
    const uint64_t* weight;
    const double* points;
    const uint64_t* Ks;
    const uint64_t* tau;     // kCand
    const double* ub;        // kLCount / kLWrite
    const uint64_t* bnd;     // last seed (key, id); (0, 0) = none
    const uint64_t* thr;     // seed top-K threshold (key of rho^2, id)
    // kCand
    uint64_t* recK;
    uint64_t* recV;
    uint64_t recCap;
    uint64_t* recCount;
    uint64_t* cntQ;
    // kLCount / kLWrite
    uint64_t* cnt;
    const uint64_t* offNext;
    uint32_t* qlistNext;
    uint64_t cap;
    uint64_t* overflow;
    uint64_t* survStat;
    uint64_t* fineStat;
    // kLCount stores each item's per-row pair masks (item i's rows from
    // maskOff[i], each wave reserving all its items' rows on maskNext) when
    // they fit in maskCap; kLWrite reads them instead of recomputing
    uint64_t* masks;
    uint64_t* maskOff;
    uint64_t* maskNext;
    uint64_t maskCap;
    // beside each stored mask row: the row's first slot in its cluster's query
    // list (the count pass's atomic returns it), so that kLWrite, when every
    // item's rows were stored, writes the lists without atomics
    uint64_t* maskPos;
    // (L-2)-level clusters by id: centre and the radius of a ball holding their children's balls
    const double* enclC;
    const double* enclR;
    const float* enclBox;    // lo[D] | hi[D] of the children's balls
};

constexpr int kRecIdBits = 40;
constexpr int kRecStage = 256;

__device__ __forceinline__ double guard_bound(double t, double r) {
    return t + r + 1e-9 * (__builtin_fabs(t) + r);
}

// kLWrite for a chunk of nc child rows: lane k holds row k's pair mask
// (queries of the slice that refine child cid); each pair takes the next slot
// of the child's query list (cnt counted back down).
__device__ __forceinline__ void write_pairs(const KnnLastArgs& A, uint64_t cid, uint64_t myMask, int nc, uint64_t q,
                                            int lane, uint64_t ltmask) {
    uint64_t myBase = 0;
    const bool any = lane < nc && myMask;
    if (any) {
        const uint64_t pc = __popcll(myMask);
        myBase = A.offNext[cid] + (atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[cid]), 0ull - pc) - pc);
    }
    // only the rows that hold pairs
    for (uint64_t nz = __ballot(any); nz; nz &= nz - 1) {
        const int k = __builtin_ctzll(nz);
        const uint64_t m = readlane_u64(myMask, k);
        const uint64_t base = readlane_u64(myBase, k);
        if ((m >> lane) & 1ull) {
            const uint64_t o = base + __popcll(m & ltmask);
            if (o < A.cap) A.qlistNext[o] = static_cast<uint32_t>(q);
            else atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
        }
    }
}

// LDS row of a staged query: point (DM) | G | five pass parameters.
template <int DM>
constexpr int last_row() { return DM + 6; }

template <int DM, int PASS>
__global__ __launch_bounds__(kStepWaves * kWave) void k_knn_last(KnnLastArgs A) {
    constexpr bool COUNT = PASS == kLCount || PASS == kLStats;
    constexpr bool STATS = PASS == kLStats;
    constexpr int RS = last_row<DM>();
    __shared__ double s_q[kStepWaves][kWave * RS];
    __shared__ uint64_t s_rk[kStepWaves][PASS == kCand ? kRecStage : 1];
    __shared__ uint64_t s_rv[kStepWaves][PASS == kCand ? kRecStage : 1];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int D = A.D;
    uint64_t nItems = *A.nItems;
    if (nItems > A.itemCap) nItems = A.itemCap;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    double* sq = s_q[wave];
    const ConstF64 encC = (ConstF64)A.enclC;
    const ConstF32 encB = (ConstF32)A.enclBox;
    // kLWrite: every item's mask rows were stored (the count pass reserved
    // within the capacity), so their list slots are known
    const bool allStored = PASS == kLWrite && A.masks && *A.maskNext <= A.maskCap;
    uint64_t staged = 0;
    auto flushRec = [&]() {
        if (staged == 0) return;
        wave_lds_sync();
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(A.recCount), staged);
        base = __shfl(base, 0, kWave);
        for (uint64_t o = lane; o < staged; o += kWave)
            if (base + o < A.recCap) {
                A.recK[base + o] = s_rk[wave][o];
                A.recV[base + o] = s_rv[wave][o];
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.cntQ[s_rv[wave][o] >> kRecIdBits]), 1ull);
            }
        if (lane == 0 && base + staged > A.recCap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 4ull);
        wave_lds_sync();
        staged = 0;
    };
    const uint64_t gw = blockIdx.x * static_cast<uint64_t>(kStepWaves) + wave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    // count passes: one reservation of pair-mask rows for all of this wave's items
    uint64_t maskBase = 0;
    if (COUNT && A.masks) {
        uint64_t rows = 0;
        for (uint64_t item = gw; item < nItems; item += nw) {
            const uint64_t p = A.items[item] >> kSliceBits;
            rows += A.colResPrev[p + 1] - A.colResPrev[p];
        }
        if (lane == 0 && rows) maskBase = atomicAdd(reinterpret_cast<unsigned long long*>(A.maskNext), rows);
        maskBase = uniform_u64(__shfl(maskBase, 0, kWave));
    }
    for (uint64_t item = gw; item < nItems; item += nw) {
        const uint64_t it = A.items[item];
        const uint64_t p = it >> kSliceBits;
        const uint64_t start = (it & ((1ull << kSliceBits) - 1)) * kWave;
        const uint64_t g0 = A.off[p];
        const uint64_t gn = A.off[p + 1] - g0;
        // a list slot past the written pairs (a pair buffer overflow, rerun by the
        // host) may hold a stale query number: only queries below nq are taken
        const bool inList = start + lane < gn && g0 + start + lane < A.qcap;
        const uint64_t ql = inList ? A.qlist[g0 + start + lane] : 0;
        const bool valid = inList && ql < A.nq;
        const uint64_t q = valid ? ql : 0;
        const uint64_t jb = A.colResPrev[p], je = A.colResPrev[p + 1];
        uint64_t mo = 0;
        bool stored = false;
        if (PASS != kCand && A.masks) {
            if (COUNT) {
                mo = maskBase;
                maskBase += je - jb;
                if (lane == 0) A.maskOff[item] = mo;
            } else {
                mo = A.maskOff[item];
            }
            stored = mo + (je - jb) <= A.maskCap;
        }
        if (PASS == kLWrite && stored) {
            // pair masks of the count pass: only the writes remain
            for (uint64_t j0 = jb; j0 < je; j0 += kWave) {
                const int nc = static_cast<int>(je - j0 < static_cast<uint64_t>(kWave) ? je - j0 : kWave);
                const uint64_t myMask = lane < nc ? A.masks[mo + (j0 - jb) + lane] : 0;
                const bool any = myMask != 0;
                if (allStored) {
                    // slots known from the count pass
                    const uint64_t myBase = any ? A.offNext[A.ids[j0 + lane]] + A.maskPos[mo + (j0 - jb) + lane] : 0;
                    for (uint64_t nz = __ballot(any); nz; nz &= nz - 1) {
                        const int k = __builtin_ctzll(nz);
                        const uint64_t m = readlane_u64(myMask, k);
                        if ((m >> lane) & 1ull) {
                            const uint64_t o = readlane_u64(myBase, k) + __popcll(m & ltmask);
                            if (o < A.cap) A.qlistNext[o] = static_cast<uint32_t>(q);
                            else atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
                        }
                    }
                } else {
                    write_pairs(A, any ? A.ids[j0 + lane] : 0, myMask, nc, q, lane, ltmask);
                }
            }
            continue;
        }
        double qp[DM];
#pragma unroll
        for (int d = 0; d < DM; ++d) qp[d] = (d < D) ? A.points[q * D + d] : 0.0;
        const uint64_t K = valid ? A.Ks[q] : 0;
        // per-query thresholds (lane = query of the slice)
        uint64_t tK = ~0ull, tV = 0, bK = 0, bV = 0;
        double ub = -__builtin_inf(), rho = __builtin_inf(), tPar = __builtin_inf();
        bool need = false;
        const double E = A.enclR[p];
        if (PASS == kCand) {
            if (valid && K > 0) {
                tK = A.tau[2 * q];
                tV = A.tau[2 * q + 1];
                need = true;
                if (tK != ~0ull) tPar = okey_inv(tK);
            }
        } else if (valid) {
            ub = A.ub[q];
            bK = A.bnd[2 * q];
            bV = A.bnd[2 * q + 1];
            const uint64_t rk = A.thr[2 * q];
            if (rk != ~0ull) rho = __builtin_sqrt(okey_inv(rk));
            need = STATS || K > 0;
            // every child radius is at most E, so this bounds each child's threshold
            tPar = (STATS || rho == __builtin_inf()) ? ub : fmin(ub, rho + 1e-9 * (1.0 + rho + E));
        }
        // Whole parent: every child ball lies in the ball (centre C, radius E)
        // and in the box of the parent, so a child's lb is at least |q - C| - E
        // and, for q outside the box, at least the distance to the box; when
        // either exceeds the threshold by far more than the rounding of both
        // sides, no child of this parent can pass for this query.
        if (need) {
            double c[DM];
            scalar_row<DM, false>(c, encC, static_cast<uint32_t>(p), D);
            double d2 = 0.0, b2 = 0.0;
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                const double diff = qp[d] - c[d];
                d2 = __builtin_fma(diff, diff, d2);
                if (d < D) {
                    const double lo = encB[p * 2 * D + d], hi = encB[p * 2 * D + D + d];
                    const double g = fmax(fmax(lo - qp[d], qp[d] - hi), 0.0);
                    b2 = __builtin_fma(g, g, b2);
                }
            }
            const double dP = __builtin_sqrt(d2), dB = __builtin_sqrt(b2);
            const double slack = 1e-9 * (1.0 + __builtin_fabs(tPar) + dP + E);
            if (dP - E > tPar + slack || dB > fmax(tPar, 0.0) + slack) need = false;
        }
        const uint64_t act = __ballot(need);
        if (act == 0) {
            if (COUNT && stored)
                for (uint64_t j = jb + lane; j < je; j += kWave) A.masks[mo + (j - jb)] = 0;
            continue;
        }
        // Per test, d2 > (G + r(1 + 2e-9))^2 with G = T + 2e-9|T| rejects
        // everything the exact guard rejects for any per-child threshold t <= T.
        {
            double* row = sq + lane * RS;
#pragma unroll
            for (int d = 0; d < DM; ++d) row[d] = qp[d];
            row[DM] = tPar + 2e-9 * __builtin_fabs(tPar);
            if (PASS == kCand) {
                row[DM + 1] = __longlong_as_double(tK);
                row[DM + 2] = __longlong_as_double(tV);
                row[DM + 3] = __longlong_as_double(q);
            } else {
                row[DM + 1] = ub;
                row[DM + 2] = rho;
                row[DM + 3] = __longlong_as_double(bK);
                row[DM + 4] = __longlong_as_double(bV);
                row[DM + 5] = __longlong_as_double(K > 0 ? 1 : 0);
            }
        }
        wave_lds_sync();
        uint64_t mine = 0, fine = 0;
        for (uint64_t j0 = jb; j0 < je; j0 += kWave) {
            const int nc = static_cast<int>(je - j0 < static_cast<uint64_t>(kWave) ? je - j0 : kWave);
            const bool child = lane < nc;
            // lane = child of the chunk
            double xv[DM];
            double r = 0.0, rg = -__builtin_inf();
            uint64_t id = 0, w = 0;
            if (child) {
                load_row<DM>(A.coords + (j0 + lane) * D, D, xv);
                r = A.radius[j0 + lane];
                rg = r + 2e-9 * r;
                id = A.ids[j0 + lane];
                if (STATS) w = A.weight[j0 + lane];
            } else {
#pragma unroll
                for (int d = 0; d < DM; ++d) xv[d] = 0.0;
            }
            uint64_t myMask = 0;
            for (uint64_t m = act; m; m &= m - 1) {
                const int i = __builtin_ctzll(m);
                const double* qr = sq + i * RS;
                const double d2 = d2_reg<DM>(xv, qr);
                const double Bp = qr[DM] + rg;
                const bool pre = Bp >= 0.0 && d2 <= Bp * Bp;
                if (__ballot(pre) == 0) continue;
                if (PASS == kCand) {
                    const uint64_t tKi = __double_as_longlong(qr[DM + 1]);
                    const uint64_t tVi = __double_as_longlong(qr[DM + 2]);
                    bool emit = false;
                    uint64_t key = 0;
                    if (pre) {
                        const double B = guard_bound(okey_inv(tKi), r);
                        if (tKi == ~0ull || (B >= 0.0 && d2 <= B * B)) {
                            key = okey(sqrt_rd(d2) - r);
                            emit = key < tKi || (key == tKi && id <= tVi);
                        }
                    }
                    const uint64_t em = __ballot(emit);
                    if (em) {
                        const uint64_t qi = __double_as_longlong(qr[DM + 3]);
                        const uint64_t cnt = __popcll(em);
                        if (staged + cnt > static_cast<uint64_t>(kRecStage)) flushRec();
                        if (emit) {
                            const uint64_t o = staged + __popcll(em & ltmask);
                            s_rk[wave][o] = key;
                            s_rv[wave][o] = (qi << kRecIdBits) | id;
                        }
                        staged += cnt;
                    }
                } else {
                    const double ubi = qr[DM + 1], rhoi = qr[DM + 2];
                    const uint64_t bKi = __double_as_longlong(qr[DM + 3]);
                    const uint64_t bVi = __double_as_longlong(qr[DM + 4]);
                    const bool pairable = __double_as_longlong(qr[DM + 5]) != 0;
                    const bool rhoInf = rhoi == __builtin_inf();
                    bool surv = false, pair = false;
                    if (pre) {
                        const double tp = rhoInf ? ubi : fmin(ubi, rhoi + 1e-9 * (1.0 + rhoi + r));
                        const double t = STATS ? ubi : tp;
                        const double B = guard_bound(t, r);
                        if (B >= 0.0 && d2 <= B * B) {
                            const double lb = sqrt_rd(d2) - r;
                            surv = lb <= ubi;
                            const bool seed = !lt_pair(bKi, bVi, okey(lb), id);
                            pair = surv && pairable && !seed && (rhoInf || lb <= rhoi + 1e-9 * (1.0 + rhoi + r));
                        }
                    }
                    if (pair) myMask |= 1ull << i;
                    if (STATS) {
                        const uint64_t sm = __ballot(surv);
                        const uint64_t fw = wave_sum_u64(surv ? w : 0);
                        if (lane == i) {
                            mine += __popcll(sm);
                            fine += fw;
                        }
                    }
                }
            }
            if (COUNT) {
                uint64_t slot = 0;
                if (child && myMask)
                    slot = atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[id]),
                                     static_cast<unsigned long long>(__popcll(myMask)));
                if (stored && child) {
                    A.masks[mo + (j0 - jb) + lane] = myMask;
                    A.maskPos[mo + (j0 - jb) + lane] = slot;
                }
            } else if (PASS == kLWrite) {
                write_pairs(A, id, myMask, nc, q, lane, ltmask);
            }
        }
        wave_lds_sync();
        if (STATS && valid) {
            if (A.survStat && mine)
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.survStat[q * A.L + A.L - 1]), mine);
            if (A.fineStat && fine) atomicAdd(reinterpret_cast<unsigned long long*>(&A.fineStat[q]), fine);
        }
    }
    if (PASS == kCand) flushRec();
}

// Last level, query-major: exact best M from the grouped candidate records,
// STEP, seeds and their expansion (the seed top-K and certified radius).
struct KnnSelectArgs {
    int D, L;
    uint64_t nq;
    const double* points;
    const uint64_t* Ks;
    uint64_t Kmax;
    uint64_t emptyLast;
    uint32_t seedMin;
    const uint64_t* offQ;
    const uint64_t* gK;
    const uint64_t* gV;
    const uint64_t* weightById;  // level L-1
    const double* radiusById;
    const uint64_t* colRes;      // P[L-1]
    const double* coordsFine;
    const uint64_t* idsFine;
    double* ub;
    uint64_t* bnd;
    uint64_t* thr;
    uint64_t* outIds;
    double* outD2;
    uint64_t* outN;
    uint64_t* survStat;
    uint64_t* fineStat;
    const double* radius;        // proposed search radii (learned mode), or null (seeds)
    uint64_t* work;              // distance evaluations of the seed expansion
};

template <int DM>
__global__ __launch_bounds__(kKnnThreads) void k_knn_select(KnnSelectArgs A) {
    __shared__ KnnShared S;
    if (threadIdx.x < 2 * kKnnQB) {  // zero between rounds from here on
        (&S.need[0][0])[threadIdx.x] = 0;
        (&S.fill[0][0])[threadIdx.x] = 0;
    }
    const int tid = threadIdx.x;
    const int D = A.D;
    const uint64_t ngroups = (A.nq + kKnnQB - 1) / kKnnQB;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t q0 = g * kKnnQB;
        if (tid < kKnnQB) {
            const uint64_t q = q0 + tid;
            const uint64_t K = q < A.nq ? A.Ks[q] : 0;
            S.K[tid] = K;
            uint64_t m = K > 0 ? K + A.emptyLast + 1 : 0;
            if (m > 0 && m < A.seedMin) m = A.seedMin;
            if (m > static_cast<uint64_t>(kKnnMaxM)) m = kKnnMaxM;
            S.M[tid] = static_cast<uint32_t>(m);
            S.cnt[tid] = 0;
            S.thrK[tid] = ~0ull;
            S.thrV[tid] = ~0ull;
            if (q < A.nq) {
                if (A.survStat) A.survStat[q * A.L + A.L - 1] = 0;
                if (A.fineStat) A.fineStat[q] = 0;
            }
        }
        if (tid == 0) S.listN = 0;
        for (int t = tid; t < kKnnQB * 32; t += kKnnThreads) {
            const int i = t / 32, d = t % 32;
            const uint64_t q = q0 + i;
            S.qp[i][d] = q < A.nq && d < D ? A.points[q * D + d] : 0.0;
        }
        __syncthreads();
        // every query slot's records in the same rounds (one per thread and slot)
        uint64_t r0[kKnnQB], rn[kKnnQB], rmax = 0;
#pragma unroll
        for (int i = 0; i < kKnnQB; ++i) {
            const uint64_t q = q0 + i;
            r0[i] = 0;
            rn[i] = 0;
            if (q < A.nq && S.M[i] > 0) {
                r0[i] = A.offQ[q];
                rn[i] = A.offQ[q + 1] - r0[i];
            }
            rmax = rn[i] > rmax ? rn[i] : rmax;
        }
        int par = 0;
        for (uint64_t b = 0; b < rmax; b += kKnnThreads, par ^= 1) {
            bool have[kKnnQB];
            uint64_t key[kKnnQB], val[kKnnQB];
#pragma unroll
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t r = b + tid;
                have[i] = r < rn[i];
                key[i] = have[i] ? A.gK[r0[i] + r] : 0;
                val[i] = have[i] ? A.gV[r0[i] + r] : 0;
            }
            knn_round(S, tid, have, key, val, par);
        }
        __syncthreads();
        knn_compact(S, tid);
        knn_step(S, tid, A.weightById, A.radiusById);
        if (A.radius) {
            // learned mode: no seeds; every survivor within the proposed radius
            // is refined and each point within it becomes a record
            if (tid < kKnnQB && q0 + tid < A.nq) {
                const uint64_t q = q0 + tid;
                const double r = A.radius[q];
                A.ub[q] = S.ub[tid];
                A.bnd[2 * q] = 0;
                A.bnd[2 * q + 1] = 0;
                A.outN[q] = 0;
                A.thr[2 * q] = S.K[tid] > 0 ? okey(r * r) : 0;
                A.thr[2 * q + 1] = ~0ull;
            }
            __syncthreads();
            continue;
        }
        knn_pick_seeds(S, tid, q0, A.nq, A.seedMin);
        if (tid < kKnnQB && q0 + tid < A.nq) {
            const uint64_t q = q0 + tid;
            A.ub[q] = S.ub[tid];
            A.bnd[2 * q] = S.bndK[tid];
            A.bnd[2 * q + 1] = S.bndV[tid];
        }
        knn_expand_seeds<DM>(S, tid, q0, A.nq, D, A.colRes, A.coordsFine, A.idsFine, A.work);
        knn_write_topk(S, tid, q0, A.nq, A.Kmax, A.outIds, A.outD2, A.outN, A.thr);
        __syncthreads();
    }
}

// ------------------------------------------------------ cluster-major refine
// Work item = (last-level cluster c, slice of up to 64 queries that kept c as
// a survivor within their certified radius).  Lane = query; the cluster's
// children are staged in LDS 64 at a time and scored against every lane's
// query with the same fma chain as the per-query refine.  A child is a record
// for query q only if (d2, id) precedes q's seed threshold, so records are the
// few points that can still displace a seed result.
struct KnnClusterArgs {
    int D;
    uint64_t nq;
    const uint64_t* items;
    const uint64_t* nItems;
    uint64_t itemCap;
    const uint64_t* off;
    const uint32_t* qlist;
    uint64_t qcap;
    const uint64_t* colRes;   // P[L-1] column pointers
    const uint64_t* leafOff;  // first kd leaf of every last-level cluster
    const float* leafBox;     // outward-rounded float box lo[D] | hi[D] of every kd leaf
    const double* coords;     // finest positions
    const uint64_t* ids;
    const double* points;     // nq x D
    const uint64_t* thr;      // nq x 2
    uint64_t* recK;           // squared-distance keys
    uint64_t* recV;           // (q << kRecIdBits) | id
    uint64_t recCap;
    uint64_t* recCount;
    uint64_t* cntQ;           // records per query
    uint64_t* overflow;
    uint64_t* work;           // distance evaluations (child rows x queries)
    // row-major refine of the clusters refined by at most rowMax queries:
    // their ids in per-tile slots of the group scan's list (kScanThreads slots
    // and a count per tile)
    const uint64_t* sparse;
    const uint64_t* sparseCnt;
    uint64_t sparseTiles;
    uint64_t rowMax;
};

// Records of the cluster-major refine staged per wave (one reservation per
// flush of up to kClusterRec records rather than one per row that emits:
// with learned radii most rows near a query emit).
constexpr int kClusterRec = 64;

template <int DM>
__global__ __launch_bounds__(kStepWaves * kWave) void k_knn_cluster(KnnClusterArgs A) {
    __shared__ double s_x[kStepWaves][kWave * DM];
    __shared__ uint64_t s_id[kStepWaves][kWave];
    __shared__ uint64_t s_rk[kStepWaves][kClusterRec];
    __shared__ uint64_t s_rv[kStepWaves][kClusterRec];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t gw = blockIdx.x * static_cast<uint64_t>(kStepWaves) + wave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = A.D;
    uint64_t nItems = *A.nItems;
    if (nItems > A.itemCap) nItems = A.itemCap;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    double* sx = s_x[wave];
    uint64_t* rk = s_rk[wave];
    uint64_t* rv = s_rv[wave];
    uint32_t staged = 0;
    auto flush = [&]() {
        if (staged == 0) return;
        wave_lds_sync();
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(A.recCount), staged);
        base = uniform_u64(__shfl(base, 0, kWave));
        if (static_cast<uint32_t>(lane) < staged) {
            if (base + lane < A.recCap) {
                A.recK[base + lane] = rk[lane];
                A.recV[base + lane] = rv[lane];
            } else {
                // not stored: take it back out of its query's count (the items add
                // their records to the counts as they emit them), so that the
                // counts, and the offsets scanned from them, cover stored records only
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.cntQ[rv[lane] >> kRecIdBits]), ~0ull);
            }
        }
        if (lane == 0 && base + staged > A.recCap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 4ull);
        wave_lds_sync();
        staged = 0;
    };
    uint64_t wTests = 0, wRows = 0;
    if (D != DM)
        for (int t = lane; t < kWave * DM; t += kWave) sx[t] = 0.0;
    for (uint64_t item = gw; item < nItems; item += nw) {
        const uint64_t it = A.items[item];
        const uint64_t c = it >> kSliceBits;
        const uint64_t start = (it & ((1ull << kSliceBits) - 1)) * kWave;
        const uint64_t g0 = A.off[c];
        const uint64_t gn = A.off[c + 1] - g0;
        // a list slot past the written pairs (a pair buffer overflow, rerun by the
        // host) may hold a stale query number: only queries below nq are taken
        const bool inList = start + lane < gn && g0 + start + lane < A.qcap;
        const uint64_t ql = inList ? A.qlist[g0 + start + lane] : 0;
        const bool valid = inList && ql < A.nq;
        const uint64_t q = valid ? ql : 0;
        double qp[DM];
#pragma unroll
        for (int d = 0; d < DM; ++d) qp[d] = (d < D) ? A.points[q * D + d] : 0.0;
        const uint64_t tK = valid ? A.thr[2 * q] : 0, tV = valid ? A.thr[2 * q + 1] : 0;
        // the threshold as a double: keys of non-negative distances order like
        // the doubles, so only ties and NaNs (tK = ~0 maps to a NaN) need the keys;
        // an idle lane compares above -inf and never emits
        const double tD = valid ? okey_inv(tK) : -__builtin_inf();
        const uint64_t jb = A.colRes[c], je = A.colRes[c + 1];
        uint64_t mine = 0;
        for (uint64_t j0 = jb; j0 < je; j0 += kWave) {
            const int nc = static_cast<int>(je - j0 < static_cast<uint64_t>(kWave) ? je - j0 : kWave);
            {
                const double* src = A.coords + j0 * D;
                const int tot = nc * D;
                if (D == DM) {
                    for (int t = lane; t < tot; t += kWave) sx[t] = src[t];
                } else {
                    for (int t = lane; t < tot; t += kWave) sx[(t / D) * DM + t % D] = src[t];
                }
                if (lane < nc) s_id[wave][lane] = A.ids[j0 + lane];
            }
            wave_lds_sync();
            for (int k = 0; k < nc; ++k) {
                const double* x = sx + k * DM;
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < DM; ++d) {
                    const double diff = qp[d] - x[d];
                    acc = __builtin_fma(diff, diff, acc);
                }
                // a record when (acc, id) precedes the lane's threshold (tD as a
                // double; the keys only for ties and NaNs)
                bool emit = acc < tD;
                const bool tie = !(emit || acc > tD);
                if (__ballot(tie)) {
                    if (tie) {
                        const uint64_t key = okey(acc);
                        emit = key < tK || (key == tK && s_id[wave][k] < tV);
                    }
                }
                const uint64_t em = __ballot(emit);
                if (em) {
                    if (staged + __popcll(em) > static_cast<uint32_t>(kClusterRec)) flush();
                    if (emit) {
                        const uint32_t o = staged + __popcll(em & ltmask);
                        rk[o] = okey(acc);
                        rv[o] = (q << kRecIdBits) | s_id[wave][k];
                        ++mine;
                    }
                    staged += __popcll(em);
                }
            }
            wave_lds_sync();
        }
        // (records past the capacity are taken back out by the flush that drops them)
        if (valid && mine) atomicAdd(reinterpret_cast<unsigned long long*>(&A.cntQ[q]), mine);
        wRows += je - jb;
        wTests += (je - jb) * static_cast<uint64_t>(__popcll(__ballot(valid)));
    }
    flush();
    if (lane == 0 && A.work) atomicAdd(reinterpret_cast<unsigned long long*>(A.work), wTests);
    (void)wRows;
}

// ------------------------------------------------------ row-major refine
// Last-level clusters refined by at most kRowQ queries (most of them when the
// queries are few against the clusters, e.g. a thousand KNN points over a
// 500M-point index) leave lane = query mostly idle.  The group scan lists them
// apart (no work items) and they are refined here: a unit of kRowUnit list
// slots, lane = (cluster slot, query slot) pair, query points and thresholds
// in LDS.  The kd leaves of the unit's clusters are then tested with lane =
// (cluster slot, leaf) -- one box load per lane, against the slot's queries;
// the (pair, leaf) tuples that a box cannot rule out are queued in LDS and scored kWave / kLeafRows at a time, lane =
// (tuple, row of the leaf), with the same fma chain and threshold rule, so that
// the records are the cluster-major refine's.
constexpr int kRowQ = 4;
static_assert(kRowUnit * kRowQ == kWave, "one list unit per wave pass");
constexpr int kRowGroup = kWave / kLeafRows;  // tuples scored per wave pass
constexpr int kRowRec = kWave;
// tuple queue: fewer than kWave left before a leaf pass adds up to kWave * kRowQ
constexpr int kRowTupCap = kWave + kWave * kRowQ;

__device__ __forceinline__ uint64_t bperm_u64(uint64_t v, int src) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(static_cast<uint32_t>(v)));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(static_cast<uint32_t>(v >> 32)));
    return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}
__device__ __forceinline__ double bperm_f64(double v, int src) {
    return __longlong_as_double(static_cast<long long>(bperm_u64(static_cast<uint64_t>(__double_as_longlong(v)), src)));
}

template <int DM, bool DENSE>
__global__ __launch_bounds__(kStepWaves * kWave) void k_knn_rows(KnnClusterArgs A) {
    __shared__ double s_qp[kStepWaves][kWave][DM];  // query point of every pair lane
    __shared__ double s_td[kStepWaves][kWave];      // its threshold as a double (-inf: no pair)
    __shared__ uint32_t s_tup[kStepWaves][kRowTupCap];
    __shared__ uint64_t s_rk[kStepWaves][kRowRec];
    __shared__ uint64_t s_rv[kStepWaves][kRowRec];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t gw = blockIdx.x * static_cast<uint64_t>(kStepWaves) + wave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = A.D;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    // sparse: units of kRowUnit list slots; dense: the work items (cluster,
    // slice of 64 queries), each taken as kRowUnit slots of one cluster
    uint64_t nUnits = A.sparseTiles * (kScanThreads / kRowUnit);
    if (DENSE) {
        nUnits = *A.nItems;
        if (nUnits > A.itemCap) nUnits = A.itemCap;
    }
    double(*qs)[DM] = s_qp[wave];
    double* tds = s_td[wave];
    uint32_t* tup = s_tup[wave];
    uint64_t* rk = s_rk[wave];
    uint64_t* rv = s_rv[wave];
    uint32_t staged = 0;
    uint64_t wTests = 0;
    // staged records: one reservation per flush; cntQ counts the stored ones
    auto flush = [&]() {
        if (staged == 0) return;
        wave_lds_sync();
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(A.recCount), staged);
        base = uniform_u64(__shfl(base, 0, kWave));
        for (uint32_t o0 = 0; o0 < staged; o0 += kWave) {
            const uint32_t o = o0 + lane;
            const bool stored = o < staged && base + o < A.recCap;
            const uint64_t qr = stored ? rv[o] >> kRecIdBits : ~0ull;
            if (stored) {
                A.recK[base + o] = rk[o];
                A.recV[base + o] = rv[o];
            }
            // one count update per distinct query of the pass, not per record
            for (uint64_t left = __ballot(stored); left;) {
                const uint64_t qv = readlane_u64(qr, __builtin_ctzll(left));
                const uint64_t same = __ballot(qr == qv);
                if (lane == __builtin_ctzll(left))
                    atomicAdd(reinterpret_cast<unsigned long long*>(&A.cntQ[qv]),
                              static_cast<unsigned long long>(__popcll(same)));
                left &= ~same;
            }
        }
        if (lane == 0 && base + staged > A.recCap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 4ull);
        wave_lds_sync();
        staged = 0;
    };
    // the list entry of this lane's slot in unit uu and the tile's list end,
    // loaded one unit ahead (the slot is read whatever the count: it lies in
    // the list's allocation)
    // units of kRowUnit list slots, wave-strided over a resident grid: step u
    // is unit u / tiles of tile u % tiles, so that every wave meets the first
    // units of tiles (the full ones) as often as the last
    const uint64_t tiles = A.sparseTiles;
    auto unit_slot = [&](uint64_t uu) {
        return (uu % tiles) * kScanThreads + (uu / tiles) * kRowUnit;
    };
    uint64_t nextEnd = 0, nextC = 0;
    auto fetch = [&](uint64_t uu) {
        if (!DENSE && uu < nUnits) {
            const uint64_t s0 = unit_slot(uu);
            nextEnd = s0 / kScanThreads * kScanThreads + A.sparseCnt[s0 / kScanThreads];
            nextC = A.sparse[s0 + lane / kRowQ];
        }
    };
    fetch(gw);
    for (uint64_t u = gw; u < nUnits; u += nw) {
        uint64_t j = lane % kRowQ;
        bool inSlot = true;
        uint64_t cl = 0;
        if (DENSE) {
            // lane = query lane of the item's slice (the list position)
            const uint64_t it = A.items[u];
            cl = it >> kSliceBits;
            j = (it & ((1ull << kSliceBits) - 1)) * kWave + lane;
        } else {
            const uint64_t s0 = unit_slot(u);
            const uint64_t tileEnd = uniform_u64(nextEnd);
            cl = nextC;
            fetch(u + nw);
            if (s0 >= tileEnd) continue;
            inSlot = s0 + lane / kRowQ < tileEnd;
        }
        uint64_t g0 = 0, gn = 0, jb = 0, je = 0, leaf0 = 0;
        if (inSlot) {
            g0 = A.off[cl];
            gn = A.off[cl + 1] - g0;
            jb = A.colRes[cl];
            je = A.colRes[cl + 1];
            leaf0 = A.leafOff[cl];
        }
        // list slots that were written (an overflowed list is rerun by the
        // host) and hold a query below nq
        const bool inList = inSlot && (DENSE || gn <= A.rowMax) && j < gn && g0 + j < A.qcap;
        const uint64_t ql = inList ? A.qlist[g0 + j] : 0;
        const bool valid = inList && ql < A.nq;
        const uint64_t q = valid ? ql : 0;
        const uint64_t tK = valid ? A.thr[2 * q] : 0, tV = valid ? A.thr[2 * q + 1] : 0;
#pragma unroll
        for (int d = 0; d < DM; ++d) qs[lane][d] = (d < D) ? A.points[q * D + d] : 0.0;
        tds[lane] = valid ? okey_inv(tK) : -__builtin_inf();
        // (slot, leaf) pairs of the unit, slot-major: each slot's leaf count on
        // its first lane (0 without a query), inclusive prefix over the slots
        const uint64_t vm = __ballot(valid);
        uint32_t x = (lane % kRowQ == 0 && ((vm >> lane) & 0xFull))
                         ? static_cast<uint32_t>((je - jb + kLeafRows - 1) / kLeafRows)
                         : 0u;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, kWave);
            if (lane >= o) x += y;
        }
        const uint32_t total = __builtin_amdgcn_readlane(x, kWave - 1);
        wave_lds_sync();
        uint32_t queued = 0;
        // scores tuples [0, n) of the queue (lane = tuple g, row r of its leaf)
        auto score = [&](uint32_t n) {
            wave_lds_sync();
            for (uint32_t b0 = 0; b0 < n; b0 += kRowGroup) {
                const uint32_t g = b0 + lane / kLeafRows;
                const uint32_t tp = g < n ? tup[g] : 0u;
                const int p = static_cast<int>(tp & (kWave - 1));
                const uint64_t t = tp / kWave;
                // (every lane takes part in a bpermute: an inactive source lane reads as 0)
                const uint64_t pjb = bperm_u64(jb, p), pje = bperm_u64(je, p);
                const uint64_t tKp = bperm_u64(tK, p), tVp = bperm_u64(tV, p);
                const uint64_t qHi = bperm_u64(q, p) << kRecIdBits;
                const uint64_t row = pjb + t * kLeafRows + (lane % kLeafRows);
                const bool act = g < n && row < pje;
                double xr[DM];
#pragma unroll
                for (int d = 0; d < DM; ++d) xr[d] = (act && d < D) ? A.coords[row * D + d] : 0.0;
                const uint64_t id = act ? A.ids[row] : 0;
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < DM; ++d) {
                    if (d < D) {
                        const double diff = qs[p][d] - xr[d];
                        acc = __builtin_fma(diff, diff, acc);
                    }
                }
                wTests += __popcll(__ballot(act));
                const double tDp = tds[p];
                bool emit = act && acc < tDp;
                if (act && !(acc < tDp || acc > tDp)) {
                    const uint64_t key = okey(acc);
                    emit = key < tKp || (key == tKp && id < tVp);
                }
                const uint64_t em = __ballot(emit);
                if (em) {
                    if (staged + __popcll(em) > static_cast<uint32_t>(kRowRec)) flush();
                    if (emit) {
                        const uint32_t o = staged + __popcll(em & ltmask);
                        rk[o] = okey(acc);
                        rv[o] = qHi | id;
                    }
                    staged += __popcll(em);
                }
            }
        };
        for (uint32_t f0 = 0; f0 < total; f0 += kWave) {
            // lane = (slot s, leaf t): the leaf's box against the slot's queries.
            // The squared distance to the (outward-rounded) box bounds every
            // row's fma-chain acc from below up to a relative 2(D+2) ulps of both
            // chains (and an absolute 1e-300 for underflow), so a query whose
            // threshold lies below it by more than that margin cannot emit any
            // of the leaf's rows
            const uint32_t f = f0 + lane;
            const bool fv = f < total;
            // slot of f: the slots whose inclusive prefix is at most f (binary
            // search over the slots' last lanes)
            int sl = 0;
#pragma unroll
            for (int step = kRowUnit / 2; step; step >>= 1) {
                const uint32_t v = static_cast<uint32_t>(
                    __builtin_amdgcn_ds_bpermute(((sl + step) * kRowQ - 1) << 2, static_cast<int>(x)));
                if (f >= v) sl += step;
            }
            const uint32_t ex = static_cast<uint32_t>(
                __builtin_amdgcn_ds_bpermute((sl * kRowQ - 1) << 2, static_cast<int>(x)));
            const uint64_t t = f - (sl ? ex : 0u);
            const uint64_t sLeaf0 = bperm_u64(leaf0, sl * kRowQ);
            float lo[DM], hi[DM];
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                const float* bx = A.leafBox + (sLeaf0 + t) * 2 * D;
                lo[d] = (fv && d < D) ? bx[d] : 0.0f;
                hi[d] = (fv && d < D) ? bx[D + d] : 0.0f;
            }
#pragma unroll 1
            for (int k = 0; k < kRowQ; ++k) {
                const int pl = sl * kRowQ + k;
                double lb = 0.0;
#pragma unroll
                for (int d = 0; d < DM; ++d) {
                    if (d < D) {
                        const double qd = qs[pl][d];
                        const double g = __builtin_fmax(
                            __builtin_fmax(static_cast<double>(lo[d]) - qd, qd - static_cast<double>(hi[d])), 0.0);
                        lb = __builtin_fma(g, g, lb);
                    }
                }
                const bool need = fv && !(lb * (1.0 - 1e-12) - 1e-300 > tds[pl]);
                const uint64_t m = __ballot(need);
                if (need) tup[queued + __popcll(m & ltmask)] = static_cast<uint32_t>(t * kWave) | pl;
                queued += __popcll(m);
            }
            if (queued >= static_cast<uint32_t>(kWave)) {
                // full passes now; the rest (fewer than kRowGroup) to the front
                const uint32_t done = queued / kRowGroup * kRowGroup;
                score(done);
                const uint32_t rest = queued - done;
                const uint32_t v = static_cast<uint32_t>(lane) < rest ? tup[done + lane] : 0u;
                wave_lds_sync();
                if (static_cast<uint32_t>(lane) < rest) tup[lane] = v;
                queued = rest;
            }
        }
        if (queued) score(queued);
        wave_lds_sync();
    }
    flush();
    if (lane == 0 && A.work) atomicAdd(reinterpret_cast<unsigned long long*>(A.work), wTests);
}

template <bool DENSE>
static auto knn_rows_kernel(int D) -> void (*)(KnnClusterArgs) {
    if (D <= 4) return k_knn_rows<4, DENSE>;
    if (D <= 8) return k_knn_rows<8, DENSE>;
    if (D <= 16) return k_knn_rows<16, DENSE>;
    return k_knn_rows<32, DENSE>;
}

// Counting-sort scatter of the records into per-query segments; leaves cntQ
// zeroed.  Chunks of kAggChunk records; with nq <= kAggLds a chunk's records are
// ranked per query in LDS and each query's run is reserved with one global
// atomic (slots count down from cntQ[q]), else one atomic per record.
__global__ __launch_bounds__(kAggThreads) void k_rec_scatter(const uint64_t* __restrict__ recK,
                                                             const uint64_t* __restrict__ recV, const uint64_t* dn,
                                                             uint64_t cap, uint64_t nq, uint64_t* __restrict__ cntQ,
                                                             const uint64_t* __restrict__ offQ,
                                                             uint64_t* __restrict__ gK, uint64_t* __restrict__ gV) {
    __shared__ uint32_t lc[kAggLds];
    const int tid = threadIdx.x;
    uint64_t n = *dn;
    if (n > cap) n = cap;
    const uint64_t idMask = (1ull << kRecIdBits) - 1;
    if (nq > kAggLds) {
        for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + tid; i < n;
             i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
            const uint64_t v = recV[i];
            const uint64_t q = v >> kRecIdBits;
            const uint64_t pos = atomicAdd(reinterpret_cast<unsigned long long*>(&cntQ[q]), ~0ull) - 1;
            if (offQ[q] + pos < cap) {  // the counts cover stored records only; never write past them
                gK[offQ[q] + pos] = recK[i];
                gV[offQ[q] + pos] = v & idMask;
            }
        }
        return;
    }
    for (uint64_t i = tid; i < nq; i += kAggThreads) lc[i] = 0;
    __syncthreads();
    const uint64_t chunks = (n + kAggChunk - 1) / kAggChunk;
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        uint64_t v[kAggPer];
        uint32_t rank[kAggPer];
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t i = c * kAggChunk + static_cast<uint64_t>(r) * kAggThreads + tid;
            v[r] = i < n ? recV[i] : ~0ull;
            const uint64_t q = v[r] >> kRecIdBits;
            rank[r] = q < nq ? atomicAdd(&lc[q], 1u) : 0;
        }
        __syncthreads();
        for (uint64_t q = tid; q < nq; q += kAggThreads) {
            const uint32_t k = lc[q];
            if (k) lc[q] = static_cast<uint32_t>(
                       atomicAdd(reinterpret_cast<unsigned long long*>(&cntQ[q]), 0ull - k) - k);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t q = v[r] >> kRecIdBits;
            if (q < nq) {
                const uint64_t i = c * kAggChunk + static_cast<uint64_t>(r) * kAggThreads + tid;
                const uint64_t pos = offQ[q] + lc[q] + rank[r];
                if (pos < cap) {
                    gK[pos] = recK[i];
                    gV[pos] = v[r] & idMask;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t q = v[r] >> kRecIdBits;
            if (q < nq) lc[q] = 0;
        }
        __syncthreads();
    }
}

// Final top-K of queries that have records: the seed results plus the records,
// by (squared distance, id).  Queries without records keep the seed results.
struct KnnMergeArgs {
    uint64_t nq;
    const uint64_t* Ks;
    uint64_t Kmax;
    const uint64_t* offQ;
    const uint64_t* gK;
    const uint64_t* gV;
    uint64_t* outIds;
    double* outD2;
    uint64_t* outN;
};

__global__ __launch_bounds__(kKnnThreads) void k_knn_merge(KnnMergeArgs A) {
    __shared__ KnnShared S;
    if (threadIdx.x < 2 * kKnnQB) {  // zero between rounds from here on
        (&S.need[0][0])[threadIdx.x] = 0;
        (&S.fill[0][0])[threadIdx.x] = 0;
    }
    __shared__ uint64_t s_r0[kKnnQB], s_rn[kKnnQB];
    __shared__ uint32_t s_any;
    const int tid = threadIdx.x;
    const uint64_t ngroups = (A.nq + kKnnQB - 1) / kKnnQB;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t q0 = g * kKnnQB;
        if (tid == 0) s_any = 0;
        __syncthreads();
        if (tid < kKnnQB) {
            const uint64_t q = q0 + tid;
            uint64_t r0 = 0, rn = 0;
            if (q < A.nq) {
                r0 = A.offQ[q];
                rn = A.offQ[q + 1] - r0;
            }
            s_r0[tid] = r0;
            s_rn[tid] = rn;
            S.M[tid] = static_cast<uint32_t>(q < A.nq && rn ? A.Ks[q] : 0);
            S.cnt[tid] = 0;
            S.thrK[tid] = ~0ull;
            S.thrV[tid] = ~0ull;
            if (rn) atomicOr(&s_any, 1u);
        }
        __syncthreads();
        if (!s_any) continue;
        for (int i = 0; i < kKnnQB; ++i) {
            if (S.M[i] == 0) continue;
            const uint64_t q = q0 + i;
            const uint32_t n0 = static_cast<uint32_t>(A.outN[q]);
            for (uint32_t e = tid; e < n0; e += kKnnThreads) {
                S.bk[i][e] = okey(A.outD2[q * A.Kmax + e]);
                S.bv[i][e] = A.outIds[q * A.Kmax + e];
            }
            if (tid == 0) S.cnt[i] = n0;
        }
        __syncthreads();
        // every query slot's records in the same rounds (one per thread and slot)
        uint64_t rmax = 0;
#pragma unroll
        for (int i = 0; i < kKnnQB; ++i)
            if (S.M[i] > 0 && s_rn[i] > rmax) rmax = s_rn[i];
        int par = 0;
        for (uint64_t b = 0; b < rmax; b += kKnnThreads, par ^= 1) {
            bool have[kKnnQB];
            uint64_t key[kKnnQB], val[kKnnQB];
#pragma unroll
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t r = b + tid;
                have[i] = S.M[i] > 0 && r < s_rn[i];
                key[i] = have[i] ? A.gK[s_r0[i] + r] : 0;
                val[i] = have[i] ? A.gV[s_r0[i] + r] : 0;
            }
            knn_round(S, tid, have, key, val, par);
        }
        __syncthreads();
        knn_compact(S, tid);
        for (int i = 0; i < kKnnQB; ++i) {
            if (S.M[i] == 0) continue;
            const uint64_t q = q0 + i;
            const uint32_t k = S.cnt[i];
            for (uint32_t e = tid; e < k; e += kKnnThreads) {
                A.outIds[q * A.Kmax + e] = S.bv[i][e];
                A.outD2[q * A.Kmax + e] = okey_inv(S.bk[i][e]);
            }
            if (tid == 0) A.outN[q] = k;
        }
        __syncthreads();
    }
}

// Learned mode: a query is exact when at least K candidates lie within its
// radius (its records); otherwise it falls back.
__global__ void k_certify(uint64_t nq, const uint64_t* __restrict__ Ks, const uint64_t* __restrict__ offQ,
                          uint8_t* __restrict__ fallback) {
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; q < nq;
         q += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        fallback[q] = (Ks[q] > 0 && offQ[q + 1] - offQ[q] < Ks[q]) ? 1 : 0;
}

// Row moves by an index list (rows of `width` words): gather dst[i] = src[idx[i]],
// scatter dst[idx[i]] = src[i].
template <bool SCATTER>
__global__ void k_move_rows(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx, uint64_t n,
                            uint64_t width, uint64_t* __restrict__ dst) {
    const uint64_t total = n * width;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t i = e / width, j = e - i * width;
        if (SCATTER) dst[idx[i] * width + j] = src[e];
        else dst[e] = src[idx[i] * width + j];
    }
}

void move_rows(tblaeq_ctx* ctx, bool scatter, const void* src, const uint64_t* idx, uint64_t n, uint64_t width,
               void* dst) {
    const uint64_t total = n * width;
    if (total == 0) return;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * 8;
    const uint64_t want = (total + 255) / 256;
    const int grid = static_cast<int>(want < cap ? want : cap);
    if (scatter)
        hipLaunchKernelGGL(k_move_rows<true>, dim3(grid), dim3(256), 0, ctx->stream,
                           static_cast<const uint64_t*>(src), idx, n, width, static_cast<uint64_t*>(dst));
    else
        hipLaunchKernelGGL(k_move_rows<false>, dim3(grid), dim3(256), 0, ctx->stream,
                           static_cast<const uint64_t*>(src), idx, n, width, static_cast<uint64_t*>(dst));
}

}  // namespace tblaeq

using namespace tblaeq;

namespace {
struct ParamSpec {
    int id;
    uint64_t tblaeq_batch::*field;
    uint64_t lo, hi;
};
const ParamSpec kParams[] = {
    {TBLAEQ_PARAM_KNN_SEED, &tblaeq_batch::knnSeed, 1, static_cast<uint64_t>(kKnnMaxM)},
    {TBLAEQ_PARAM_KNN_WG_PER_CU, &tblaeq_batch::knnWgPerCu, 1, 16},
    {TBLAEQ_PARAM_KNN_MODE, &tblaeq_batch::knnMode, 0, 1},
    {TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU, &tblaeq_batch::rangeBlocksPerCu, 1, 64},
    {TBLAEQ_PARAM_STEP_BLOCKS_PER_CU, &tblaeq_batch::stepBlocksPerCu, 1, 64},
    {TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU, &tblaeq_batch::radixBlocksPerCu, 1, 32},
    {TBLAEQ_PARAM_HIT_STAGE, &tblaeq_batch::hitStage, 256, 1024},
    {TBLAEQ_PARAM_RANGE_ROW_QUERIES, &tblaeq_batch::rangeRowMax, 0, 64},
    {TBLAEQ_PARAM_KNN_ROW_QUERIES, &tblaeq_batch::knnRowMax, 0, static_cast<uint64_t>(kRowQ)},
    {TBLAEQ_PARAM_KNN_LEAF_REFINE, &tblaeq_batch::knnLeafRefine, 0, 1},
};
const ParamSpec* find_param(int id) {
    for (const auto& p : kParams)
        if (p.id == id) return &p;
    return nullptr;
}
}  // namespace

extern "C" int tblaeq_batch_knn_max_k(const tblaeq_batch* b, uint64_t* out) {
    if (!b || !out) {
        set_error("tblaeq_batch_knn_max_k: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    uint64_t maxEmpty = 0;
    for (int l = 0; l < b->L; ++l) maxEmpty = std::max(maxEmpty, b->emptyCols[l]);
    *out = static_cast<uint64_t>(kKnnMaxM) > maxEmpty + 1 ? static_cast<uint64_t>(kKnnMaxM) - maxEmpty - 1 : 0;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_set_param(tblaeq_batch* b, int param, uint64_t value) {
    if (!b) {
        set_error("tblaeq_batch_set_param: null engine");
        return TBLAEQ_ERR_INVALID;
    }
    const ParamSpec* p = find_param(param);
    if (!p) {
        set_error("tblaeq_batch_set_param: unknown parameter");
        return TBLAEQ_ERR_INVALID;
    }
    const bool stageOk = param != TBLAEQ_PARAM_HIT_STAGE || value == 256 || value == 512 || value == 1024;
    if (value < p->lo || value > p->hi || !stageOk) {
        set_error("tblaeq_batch_set_param: value out of range");
        return TBLAEQ_ERR_INVALID;
    }
    if (b->*(p->field) != value) ++b->paramVersion;
    b->*(p->field) = value;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_get_param(const tblaeq_batch* b, int param, uint64_t* value) {
    if (!b || !value) {
        set_error("tblaeq_batch_get_param: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const ParamSpec* p = find_param(param);
    if (!p) {
        set_error("tblaeq_batch_get_param: unknown parameter");
        return TBLAEQ_ERR_INVALID;
    }
    *value = b->*(p->field);
    return TBLAEQ_OK;
}

static int knn_impl(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, uint64_t Kmax,
                    const double* d_radius, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                    uint8_t* d_fallback, uint64_t* d_status, uint64_t* d_level_surv, uint64_t* d_fine_count);

// Context workspace of a KNN batch of nq queries:
// thr | tau | bnd [2nq each] | ub [nq] | cntQ [nq] | offQ [nq+1] | flags [nq * (points of levels 0..L-2) bytes].
static uint64_t knn_ws_bytes(const tblaeq_batch* b, uint64_t nq) {
    uint64_t stride = 0;
    for (int l = 0; l + 1 < b->L; ++l) stride += b->count[l];
    if (stride == 0) stride = 1;
    const uint64_t words = 6 * nq + nq + nq + nq + 1;
    return 8 * words + nq * stride + 64;
}

extern "C" int tblaeq_batch_knn_graph(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq,
                                      uint64_t Kmax, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                                      uint64_t* d_status) {
    if (!b) return TBLAEQ_ERR_INVALID;
    tblaeq_ctx* ctx = b->ctx;
    const bool same = b->knnExec && b->kParamVersion == b->paramVersion && b->kPoints == d_points && b->kK == d_K &&
                      b->kNq == nq && b->kKmax == Kmax && b->kIds == d_out_ids && b->kD2 == d_out_d2 &&
                      b->kN == d_out_n && b->kStatus == d_status && b->kCtxStatus == ctx->status &&
                      b->kCtxWs == ctx->ws && ctx->wsBytes >= knn_ws_bytes(b, nq);
    if (!same) {
        if (b->knnExec) TB_HIP(hipGraphExecDestroy(b->knnExec));
        if (b->knnGraph) TB_HIP(hipGraphDestroy(b->knnGraph));
        b->knnExec = nullptr;
        b->knnGraph = nullptr;
        // the workspace must not be allocated inside the capture
        int rc = ws_reserve(ctx, knn_ws_bytes(b, nq));
        if (rc) return rc;
        // the row-major refine's occupancy, queried outside the capture
        resident_blocks(reinterpret_cast<const void*>(knn_rows_kernel<false>(b->D)), kStepWaves * kWave);
        resident_blocks(reinterpret_cast<const void*>(knn_rows_kernel<true>(b->D)), kStepWaves * kWave);
        const bool timing = ctx->timing;
        ctx->timing = false;  // no event records inside the graph
        const uint64_t before = ctx->scanLaunches;
        TB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = knn_impl(b, d_points, d_K, nq, Kmax, nullptr, d_out_ids, d_out_d2, d_out_n, nullptr, d_status, nullptr,
                      nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->timing = timing;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        TB_HIP(e);
        b->knnGraph = g;
        TB_HIP(hipGraphInstantiate(&b->knnExec, g, nullptr, nullptr, 0));
        b->kScanLaunches = ctx->scanLaunches - before;
        ctx->scanLaunches = before;
        b->kParamVersion = b->paramVersion;
        b->kPoints = d_points;
        b->kK = d_K;
        b->kNq = nq;
        b->kKmax = Kmax;
        b->kIds = d_out_ids;
        b->kD2 = d_out_d2;
        b->kN = d_out_n;
        b->kStatus = d_status;
        b->kCtxStatus = ctx->status;
        b->kCtxWs = ctx->ws;
    }
    // keep the host count of scan launches (epoch wrap protection) in step with replays
    for (uint64_t i = 0; i < b->kScanLaunches; ++i) {
        int rc = scan_prepare(ctx, 1);
        if (rc) return rc;
    }
    TimedScope ts(ctx, kKBatch0 + 6);
    TB_HIP(hipGraphLaunch(b->knnExec, ctx->stream));
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_knn(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq,
                                uint64_t Kmax, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                                uint64_t* d_status, uint64_t* d_level_surv, uint64_t* d_fine_count) {
    return knn_impl(b, d_points, d_K, nq, Kmax, nullptr, d_out_ids, d_out_d2, d_out_n, nullptr, d_status,
                    d_level_surv, d_fine_count);
}

extern "C" int tblaeq_batch_knn_radius(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq,
                                       uint64_t Kmax, const double* d_radius, uint64_t* d_out_ids, double* d_out_d2,
                                       uint64_t* d_out_n, uint8_t* d_fallback, uint64_t* d_status,
                                       uint64_t* d_level_surv, uint64_t* d_fine_count) {
    if (!d_radius || !d_fallback) {
        set_error("tblaeq_batch_knn_radius: radius and fallback arrays are required");
        return TBLAEQ_ERR_INVALID;
    }
    return knn_impl(b, d_points, d_K, nq, Kmax, d_radius, d_out_ids, d_out_d2, d_out_n, d_fallback, d_status,
                    d_level_surv, d_fine_count);
}

static int knn_impl(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, uint64_t Kmax,
                    const double* d_radius, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                    uint8_t* d_fallback, uint64_t* d_status, uint64_t* d_level_surv, uint64_t* d_fine_count) {
    if (!b || !d_points || !d_K || nq == 0 || !d_out_ids || !d_out_d2 || !d_out_n || !d_status) {
        set_error("tblaeq_batch_knn: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const bool clusterMajor = b->knnMode == 0;
    if (clusterMajor && (!b->pairs || !b->hits || nq > b->maxQ)) {
        set_error("tblaeq_batch_knn: engine not reserved for this many queries");
        return TBLAEQ_ERR_INVALID;
    }
    if (nq >= (1ull << kQueryBits)) {
        set_error("tblaeq_batch_knn: at most 2^24 - 1 queries per batch");
        return TBLAEQ_ERR_INVALID;
    }
    const int L = b->L;
    uint64_t maxEmpty = 0;
    for (int l = 0; l < L; ++l) maxEmpty = std::max(maxEmpty, b->emptyCols[l]);
    if (Kmax + maxEmpty + 1 > static_cast<uint64_t>(kKnnMaxM)) {
        set_error("tblaeq_batch_knn: K plus empty clusters exceeds the LDS selection size (256)");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    if (L > kKnnMaxLevels) {
        set_error("tblaeq_batch_knn: at most 8 levels");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    if (b->count[L] >= (1ull << kRecIdBits)) {
        set_error("tblaeq_batch_knn: finest level exceeds 2^40 points");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    tblaeq_ctx* ctx = b->ctx;
    KnnArgs A{};
    uint64_t stride = 0;
    for (int l = 0; l + 1 < L; ++l) {
        A.flagOff[l] = stride;
        stride += b->count[l];
    }
    if (stride == 0) stride = 1;
    // split: levels 0..L-2 query-major, the last level cluster-major
    const bool split = clusterMajor && L >= 2;
    // (the row-major refine's LDS tuples hold a leaf index below 2^26)
    const uint64_t rowMax = split && b->maxLeaves < (1ull << 26) ? b->knnRowMax : 0;
    int rc = ws_reserve(ctx, knn_ws_bytes(b, nq));
    if (rc) return rc;
    uint64_t* thr = static_cast<uint64_t*>(ctx->ws);
    uint64_t* tau = thr + 2 * nq;
    uint64_t* bnd = tau + 2 * nq;
    double* ubq = reinterpret_cast<double*>(bnd + 2 * nq);
    uint64_t* cntQ = reinterpret_cast<uint64_t*>(ubq + nq);
    uint64_t* offQ = cntQ + nq;
    uint8_t* flags = reinterpret_cast<uint8_t*>(offQ + nq + 1);
    uint64_t* ctl = b->ctl;
    launch_fill(b, ctl, 0, kCtlWords + kQueueSlots);
    A.D = b->D;
    A.L = L;
    A.nq = nq;
    A.points = d_points;
    A.Ks = d_K;
    A.Kmax = Kmax;
    for (int l = 0; l <= L; ++l) {
        A.count[l] = b->count[l];
        A.coords[l] = b->coords[l];
        A.ids[l] = b->ids[l];
        if (l < L) {
            A.parentPos[l] = b->parentPos[l];
            A.radiusPos[l] = b->radiusPos[l];
            A.weightPos[l] = b->weightPos[l];
            A.radiusById[l] = b->radiusById[l];
            A.weightById[l] = b->nnzPerCol[l];
            A.colRes[l] = b->colRes[l];
            A.emptyCols[l] = b->emptyCols[l];
        }
    }
    A.flags = flags;
    A.flagStride = stride;
    A.seedMin = static_cast<uint32_t>(b->knnSeed);
    A.outIds = d_out_ids;
    A.outD2 = d_out_d2;
    A.outN = d_out_n;
    A.survStat = d_level_surv;
    A.fineStat = d_fine_count;
    A.split = split ? 1 : 0;
    A.tau = tau;
    A.work = ctl + 9;
    if (clusterMajor) {
        A.pairs = b->pairs;
        A.pairCap = b->pairCap;
        A.pairCount = ctl + 0;
        A.cnt = b->cnt;
        A.overflow = ctl + 3;
        A.thr = thr;
        launch_fill(b, cntQ, 0, nq);
    }
    const uint64_t groups = (nq + kKnnQB - 1) / kKnnQB;
    const uint64_t wgGrid = std::min<uint64_t>(groups, static_cast<uint64_t>(ctx->cuCount) * b->knnWgPerCu);
    const uint64_t recCap = std::min(b->hitCap, b->pairCap);
    const dim3 tK(kKnnThreads);
    {
        TimedScope ts(ctx, kKBatch0 + 5);
        const dim3 g(static_cast<unsigned>(wgGrid));
        if (A.D <= 4) hipLaunchKernelGGL(k_knn_batch<4>, g, tK, 0, ctx->stream, A);
        else if (A.D <= 8) hipLaunchKernelGGL(k_knn_batch<8>, g, tK, 0, ctx->stream, A);
        else if (A.D <= 16) hipLaunchKernelGGL(k_knn_batch<16>, g, tK, 0, ctx->stream, A);
        else hipLaunchKernelGGL(k_knn_batch<32>, g, tK, 0, ctx->stream, A);
        TB_HIP(hipGetLastError());
    }
    const uint64_t stepBlocks = std::min<uint64_t>((b->itemCap + kStepWaves - 1) / kStepWaves,
                                                   static_cast<uint64_t>(ctx->cuCount) * b->stepBlocksPerCu);
    const dim3 sg(static_cast<unsigned>(stepBlocks ? stepBlocks : 1)), st(kStepWaves * kWave);
    // refine groups: set 1 after a split last level, set 0 otherwise
    const uint64_t* rItems = b->items;
    const uint64_t* rN = ctl + 1;
    const uint64_t* rOff = b->off;
    const uint32_t* rQ = b->qlist;
    if (split) {
        if ((rc = launch_grouping(b, b->count[L - 2]))) return rc;
        KnnLastArgs LA{};
        LA.D = b->D;
        LA.L = L;
        LA.nq = nq;
        LA.items = b->items;
        LA.nItems = ctl + 1;
        LA.itemCap = b->itemCap;
        LA.off = b->off;
        LA.qlist = b->qlist;
        LA.qcap = b->pairCap;
        LA.colResPrev = b->colRes[L - 2];
        LA.coords = b->coords[L - 1];
        LA.radius = b->radiusPos[L - 1];
        LA.ids = b->ids[L - 1];
        LA.weight = b->weightPos[L - 1];
        LA.points = d_points;
        LA.Ks = d_K;
        LA.tau = tau;
        LA.ub = ubq;
        LA.bnd = bnd;
        LA.thr = thr;
        LA.recK = b->hits;
        LA.recV = b->hitsTmp;
        LA.recCap = recCap;
        LA.recCount = ctl + 2;
        LA.cntQ = cntQ;
        LA.cnt = b->cnt;
        LA.offNext = b->off2;
        LA.qlistNext = b->qlist2;
        LA.cap = b->pairCap;
        LA.overflow = ctl + 3;
        LA.masks = nullptr;  // kCand writes its records into b->hits
        LA.maskPos = b->hitsTmp;  // free once the records are grouped
        LA.maskOff = b->maskOff;
        LA.maskNext = b->queue + kMaskQueue0;
        LA.maskCap = b->hitCap;
        LA.enclC = b->enclC;
        LA.enclR = b->enclR;
        LA.enclBox = b->enclBox;
        auto launchLast = [&](int pass) {
            TimedScope ts(ctx, kKBatch0 + 9);
#define TB_LAST(DMV)                                                                             \
    do {                                                                                         \
        if (pass == kCand) hipLaunchKernelGGL((k_knn_last<DMV, kCand>), sg, st, 0, ctx->stream, LA);     \
        else if (pass == kLCount)                                                                \
            hipLaunchKernelGGL((k_knn_last<DMV, kLCount>), sg, st, 0, ctx->stream, LA);          \
        else if (pass == kLStats)                                                                \
            hipLaunchKernelGGL((k_knn_last<DMV, kLStats>), sg, st, 0, ctx->stream, LA);          \
        else hipLaunchKernelGGL((k_knn_last<DMV, kLWrite>), sg, st, 0, ctx->stream, LA);         \
    } while (0)
            if (LA.D <= 4) TB_LAST(4);
            else if (LA.D <= 8) TB_LAST(8);
            else if (LA.D <= 16) TB_LAST(16);
            else TB_LAST(32);
#undef TB_LAST
        };
        // candidates for each query's best M last-level clusters, grouped per query
        launchLast(kCand);
        TB_HIP(hipGetLastError());
        if ((rc = launch_count_scan(b, cntQ, nq, offQ))) return rc;
        {
            const uint64_t g = std::min<uint64_t>((recCap + kAggChunk - 1) / kAggChunk,
                                                  static_cast<uint64_t>(ctx->cuCount) * 4);
            hipLaunchKernelGGL(k_rec_scatter, dim3(static_cast<unsigned>(g ? g : 1)), dim3(kAggThreads), 0,
                               ctx->stream, b->hits, b->hitsTmp, ctl + 2, recCap, nq, cntQ, offQ, b->pairs,
                               b->recAux);
        }
        launch_fill(b, ctl + 2, 0, 1);
        KnnSelectArgs SA{};
        SA.D = b->D;
        SA.L = L;
        SA.nq = nq;
        SA.points = d_points;
        SA.Ks = d_K;
        SA.Kmax = Kmax;
        SA.emptyLast = b->emptyCols[L - 1];
        SA.seedMin = static_cast<uint32_t>(b->knnSeed);
        SA.offQ = offQ;
        SA.gK = b->pairs;
        SA.gV = b->recAux;
        SA.weightById = b->nnzPerCol[L - 1];
        SA.radiusById = b->radiusById[L - 1];
        SA.colRes = b->colRes[L - 1];
        SA.coordsFine = b->coords[L];
        SA.idsFine = b->ids[L];
        SA.ub = ubq;
        SA.bnd = bnd;
        SA.thr = thr;
        SA.outIds = d_out_ids;
        SA.outD2 = d_out_d2;
        SA.outN = d_out_n;
        SA.survStat = d_level_surv;
        SA.fineStat = d_fine_count;
        SA.radius = d_radius;
        SA.work = ctl + 9;
        {
            TimedScope ts(ctx, kKBatch0 + 10);
            const dim3 g(static_cast<unsigned>(wgGrid));
            if (SA.D <= 4) hipLaunchKernelGGL(k_knn_select<4>, g, tK, 0, ctx->stream, SA);
            else if (SA.D <= 8) hipLaunchKernelGGL(k_knn_select<8>, g, tK, 0, ctx->stream, SA);
            else if (SA.D <= 16) hipLaunchKernelGGL(k_knn_select<16>, g, tK, 0, ctx->stream, SA);
            else hipLaunchKernelGGL(k_knn_select<32>, g, tK, 0, ctx->stream, SA);
            TB_HIP(hipGetLastError());
        }
        // survivors of the last level: statistics and the pairs to refine
        LA.survStat = d_level_surv;
        LA.fineStat = d_fine_count;
        LA.masks = b->hits;  // free again: the records went to b->pairs / b->recAux
        launchLast(d_level_surv || d_fine_count ? kLStats : kLCount);
        TB_HIP(hipGetLastError());
        if ((rc = launch_group_scan(b, b->count[L - 1], b->off2, b->items2, ctl + 4, rowMax))) return rc;
        LA.survStat = nullptr;
        LA.fineStat = nullptr;
        launchLast(kLWrite);
        TB_HIP(hipGetLastError());
        // the list writes count the per-cluster pairs back to zero only when
        // they take their slots by atomics
        launch_fill(b, b->cnt, 0, b->count[L - 1]);
        rItems = b->items2;
        rN = ctl + 4;
        rOff = b->off2;
        rQ = b->qlist2;
    } else if (clusterMajor) {
        if ((rc = launch_grouping(b, b->count[L - 1]))) return rc;
    }
    if (clusterMajor) {
        KnnClusterArgs C{};
        // clusters of the row-major list: only after a split last level
        const SparseList sl = sparse_list(b, b->count[L - 1]);
        C.rowMax = rowMax;
        C.sparse = sl.slots;
        C.sparseCnt = sl.counts;
        C.sparseTiles = (b->count[L - 1] + kScanThreads - 1) / kScanThreads;
        C.D = b->D;
        C.nq = nq;
        C.items = rItems;
        C.nItems = rN;
        C.itemCap = b->itemCap;
        C.off = rOff;
        C.qlist = rQ;
        C.qcap = b->pairCap;
        C.colRes = b->colRes[L - 1];
        C.leafOff = b->leafOff;
        C.leafBox = b->leafBox;
        C.coords = b->coords[L];
        C.ids = b->ids[L];
        C.points = d_points;
        C.thr = thr;
        C.recK = b->hits;
        C.recV = b->hitsTmp;
        C.recCap = recCap;
        C.recCount = ctl + 2;
        C.cntQ = cntQ;
        C.overflow = ctl + 3;
        C.work = ctl + 8;
        if (b->knnLeafRefine && b->maxLeaves < (1ull << 26)) {
            // the work items refined like the row-major list: kd-leaf box tests
            // per (query, leaf), then the kept leaves' rows
            TimedScope ts(ctx, kKBatch0 + 7);
            void (*k)(KnnClusterArgs) = knn_rows_kernel<true>(C.D);
            const uint64_t blocks = std::min<uint64_t>(
                static_cast<uint64_t>(sg.x),
                static_cast<uint64_t>(ctx->cuCount) *
                    static_cast<uint64_t>(resident_blocks(reinterpret_cast<const void*>(k), kStepWaves * kWave)));
            hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(blocks ? blocks : 1)), st, 0, ctx->stream, C);
            TB_HIP(hipGetLastError());
        } else {
            TimedScope ts(ctx, kKBatch0 + 7);
            if (C.D <= 4) hipLaunchKernelGGL(k_knn_cluster<4>, sg, st, 0, ctx->stream, C);
            else if (C.D <= 8) hipLaunchKernelGGL(k_knn_cluster<8>, sg, st, 0, ctx->stream, C);
            else if (C.D <= 16) hipLaunchKernelGGL(k_knn_cluster<16>, sg, st, 0, ctx->stream, C);
            else hipLaunchKernelGGL(k_knn_cluster<32>, sg, st, 0, ctx->stream, C);
            TB_HIP(hipGetLastError());
        }
        if (rowMax) {
            TimedScope ts(ctx, kKBatch0 + 13);
            void (*k)(KnnClusterArgs) = knn_rows_kernel<false>(C.D);
            // wave-strided units: a resident grid, so that every wave's share is
            // processed in one round
            const uint64_t units = C.sparseTiles * (kScanThreads / kRowUnit);
            uint64_t blocks = std::min<uint64_t>((units + kStepWaves - 1) / kStepWaves,
                                                 static_cast<uint64_t>(ctx->cuCount) * b->stepBlocksPerCu);
            blocks = std::min<uint64_t>(blocks, static_cast<uint64_t>(ctx->cuCount) *
                                                    static_cast<uint64_t>(resident_blocks(
                                                        reinterpret_cast<const void*>(k), kStepWaves * kWave)));
            hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(blocks ? blocks : 1)), st, 0, ctx->stream, C);
            TB_HIP(hipGetLastError());
        }
        if ((rc = launch_count_scan(b, cntQ, nq, offQ))) return rc;
        {
            TimedScope ts(ctx, kKBatch0 + 8);
            const uint64_t g = std::min<uint64_t>((recCap + kAggChunk - 1) / kAggChunk,
                                                  static_cast<uint64_t>(ctx->cuCount) * 4);
            hipLaunchKernelGGL(k_rec_scatter, dim3(static_cast<unsigned>(g ? g : 1)), dim3(kAggThreads), 0,
                               ctx->stream, b->hits, b->hitsTmp, ctl + 2, recCap, nq, cntQ, offQ, b->pairs,
                               b->recAux);
            KnnMergeArgs M{};
            M.nq = nq;
            M.Ks = d_K;
            M.Kmax = Kmax;
            M.offQ = offQ;
            M.gK = b->pairs;
            M.gV = b->recAux;
            M.outIds = d_out_ids;
            M.outD2 = d_out_d2;
            M.outN = d_out_n;
            const uint64_t mg = std::min<uint64_t>(groups, static_cast<uint64_t>(ctx->cuCount) * 4);
            hipLaunchKernelGGL(k_knn_merge, dim3(static_cast<unsigned>(mg)), dim3(kKnnThreads), 0, ctx->stream, M);
            TB_HIP(hipGetLastError());
        }
    }
    if (d_fallback) {
        if (split && d_radius) {
            const uint64_t g = std::min<uint64_t>((nq + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 4);
            hipLaunchKernelGGL(k_certify, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream, nq, d_K,
                               offQ, d_fallback);
        } else {
            TB_HIP(hipMemsetAsync(d_fallback, 0, nq, ctx->stream));
        }
    }
    launch_copy_status(b, d_status);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}
