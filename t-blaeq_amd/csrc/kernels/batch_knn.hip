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
};

__device__ __forceinline__ bool lt_pair(uint64_t ak, uint64_t av, uint64_t bk, uint64_t bv) {
    return ak < bk || (ak == bk && av < bv);
}

// Sorts every query's buffer (entries beyond cnt are treated as sentinels) and
// keeps its first M entries; updates the insertion threshold.
__device__ void knn_compact(KnnShared& S, int tid) {
    for (int i = 0; i < kKnnQB; ++i)
        for (int e = tid; e < kKnnCap; e += kKnnThreads)
            if (e >= static_cast<int>(S.cnt[i])) {
                S.bk[i][e] = ~0ull;
                S.bv[i][e] = ~0ull;
            }
    __syncthreads();
    for (int k = 2; k <= kKnnCap; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < kKnnQB * (kKnnCap / 2); t += kKnnThreads) {
                const int qi = t / (kKnnCap / 2);
                const int tt = t % (kKnnCap / 2);
                const int a = 2 * tt - (tt & (j - 1));
                const int b = a + j;
                const bool up = (a & k) == 0;
                const uint64_t ak = S.bk[qi][a], av = S.bv[qi][a], bk = S.bk[qi][b], bv = S.bv[qi][b];
                const bool gt = lt_pair(bk, bv, ak, av);
                if (gt == up) {
                    S.bk[qi][a] = bk;
                    S.bv[qi][a] = bv;
                    S.bk[qi][b] = ak;
                    S.bv[qi][b] = av;
                }
            }
            __syncthreads();
        }
    }
    if (tid < kKnnQB) {
        const uint32_t m = S.M[tid];
        if (S.cnt[tid] > m) S.cnt[tid] = m;
        if (m > 0 && S.cnt[tid] >= m) {
            S.thrK[tid] = S.bk[tid][m - 1];
            S.thrV[tid] = S.bv[tid][m - 1];
        } else {
            S.thrK[tid] = ~0ull;
            S.thrV[tid] = ~0ull;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void knn_insert(KnnShared& S, int i, uint64_t key, uint64_t val) {
    if (S.M[i] == 0) return;
    if (!lt_pair(key, val, S.thrK[i], S.thrV[i])) return;
    const uint32_t p = atomicAdd(&S.cnt[i], 1u);
    if (p < kKnnCap) {
        S.bk[i][p] = key;
        S.bv[i][p] = val;
    }
}

__device__ __forceinline__ bool knn_needs_compact(const KnnShared& S) {
    bool need = false;
    for (int i = 0; i < kKnnQB; ++i) need = need || (S.cnt[i] > static_cast<uint32_t>(kKnnCap - kKnnThreads));
    return need;
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

// Queues the children of last-level cluster `cid` for query slot i (caller
// guarantees a free slot).
__device__ __forceinline__ void knn_queue(const KnnArgs& A, KnnShared& S, int i, uint64_t cid) {
    const uint64_t cs = A.colRes[A.L - 1][cid], ce = A.colRes[A.L - 1][cid + 1];
    const uint32_t p = atomicAdd(&S.listN, 1u);
    S.lstart[p] = cs;
    S.lpre[p] = ce - cs;
    S.listQ[p] = static_cast<uint8_t>(i);
}

// Scores every queued child against its query and inserts it into that query's
// top-K buffer; the queue is empty afterwards.  Entry e's children are the flat
// indices [lpre[e], lpre[e+1]) once the counts are scanned in place.
template <int DM>
__device__ void knn_flush(const KnnArgs& A, KnnShared& S, int tid) {
    const int D = A.D, L = A.L, lane = tid & (kWave - 1), wid = tid / kWave;
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
    for (uint64_t f0 = 0; f0 < all; f0 += kKnnThreads) {
        if (knn_needs_compact(S)) knn_compact(S, tid);
        const uint64_t f = f0 + tid;
        if (f < all) {
            uint32_t lo = 0, hi = n;  // lpre[lo] <= f < lpre[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (S.lpre[mid] <= f) lo = mid;
                else hi = mid;
            }
            const uint64_t j = S.lstart[lo] + (f - S.lpre[lo]);
            const int i = S.listQ[lo];
            double xv[DM];
            load_row<DM>(A.coords[L] + j * D, D, xv);
            const double d2 = d2_reg<DM>(xv, S.qp[i]);
            knn_insert(S, i, okey(d2), A.ids[L][j]);
        }
        __syncthreads();
    }
    if (tid == 0) S.listN = 0;
    __syncthreads();
}

// Queues the (cluster, query slot) pair of a surviving last-level cluster.
__device__ __forceinline__ void knn_pair_queue(KnnShared& S, int i, uint64_t cid) {
    const uint32_t p = atomicAdd(&S.listN, 1u);
    S.lstart[p] = cid;
    S.listQ[p] = static_cast<uint8_t>(i);
}

// Appends the queued pairs to the global pair list (one atomic per flush) and
// counts them per cluster for the grouping scan; the queue is empty afterwards.
__device__ void knn_emit_pairs(const KnnArgs& A, KnnShared& S, int tid, uint64_t q0) {
    const uint32_t n = S.listN;
    if (n == 0) return;
    if (tid == 0) S.wsum[0] = atomicAdd(reinterpret_cast<unsigned long long*>(A.pairCount), n);
    __syncthreads();
    const uint64_t base = S.wsum[0];
    for (uint32_t e = tid; e < n; e += kKnnThreads) {
        const uint64_t pos = base + e;
        if (pos < A.pairCap) {
            const uint64_t cid = S.lstart[e];
            A.pairs[pos] = (cid << kQueryBits) | (q0 + S.listQ[e]);
            atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[cid]), 1ull);
        }
    }
    if (tid == 0 && base + n > A.pairCap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
    __syncthreads();
    if (tid == 0) S.listN = 0;
    __syncthreads();
}

template <int DM>
__global__ __launch_bounds__(kKnnThreads) void k_knn_batch(KnnArgs A) {
    __shared__ KnnShared S;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int D = A.D, L = A.L;
    const uint64_t ngroups = (A.nq + kKnnQB - 1) / kKnnQB;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t q0 = g * kKnnQB;
        if (tid < kKnnQB) {
            const uint64_t q = q0 + tid;
            S.K[tid] = q < A.nq ? A.Ks[q] : 0;
            S.surv[tid] = 0;
            S.fine[tid] = 0;
        }
        for (int t = tid; t < kKnnQB * 32; t += kKnnThreads) {
            const int i = t / 32, d = t % 32;
            const uint64_t q = q0 + i;
            S.qp[i][d] = q < A.nq && d < D ? A.points[q * D + d] : 0.0;
        }
        __syncthreads();
        for (int l = 0; l < L; ++l) {
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
                S.tot[tid] = 0;
            }
            __syncthreads();
            for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads) {
                if (knn_needs_compact(S)) knn_compact(S, tid);
                const uint64_t j = j0 + tid;
                if (j < n) {
                    double xv[DM];
                    load_row<DM>(X + j * D, D, xv);
                    const double r = R[j];
                    const uint64_t id = I[j];
                    for (int i = 0; i < kKnnQB; ++i) {
                        const uint64_t q = q0 + i;
                        if (q >= A.nq) continue;
                        if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                        knn_insert(S, i, okey(lb_reg<DM>(xv, S.qp[i], r)), id);
                    }
                }
                __syncthreads();
            }
            knn_compact(S, tid);
            // ---- STEP: phase 1 over the sorted best-M list (wave i handles query i)
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
                        w = A.weightById[l][id];
                        c = okey_inv(S.bk[i][e]) + 2.0 * A.radiusById[l][id];
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
                    if (i0 < m) {
                        S.bndK[i] = S.bk[i][i0];
                        S.bndV[i] = S.bv[i][i0];
                    } else {
                        S.bndK[i] = ~0ull;
                        S.bndV[i] = ~0ull;
                    }
                }
            }
            __syncthreads();
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
            // ---- last level: seed clusters = the surviving prefix of the sorted
            // list, at least phase 1 and up to seedMin clusters
            if (tid < kKnnQB) {
                const int i = tid;
                uint32_t ns = 0;
                if (q0 + i < A.nq && S.K[i] > 0) {
                    const uint32_t m = S.cnt[i];
                    uint32_t lim = S.i0[i] > A.seedMin ? S.i0[i] : A.seedMin;
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
                if (ns > 0) {
                    S.bndK[i] = S.bk[i][ns - 1];
                    S.bndV[i] = S.bv[i][ns - 1];
                }
                if (i == 0) S.listN = 0;
            }
            __syncthreads();
            for (int i = 0; i < kKnnQB; ++i)
                for (uint32_t e = tid; e < S.seedN[i]; e += kKnnThreads) knn_queue(A, S, i, S.bv[i][e]);
            __syncthreads();
            if (tid < kKnnQB) {
                S.M[tid] = static_cast<uint32_t>(q0 + tid < A.nq ? S.K[tid] : 0);
                S.cnt[tid] = 0;
                S.thrK[tid] = ~0ull;
                S.thrV[tid] = ~0ull;
            }
            __syncthreads();
            knn_flush<DM>(A, S, tid);
            knn_compact(S, tid);  // thr = K-th seed child: the certified radius
            // ---- pass 2 of the last level: other survivors within the certified radius
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
                        if (A.pairs) knn_pair_queue(S, i, id);
                        else knn_queue(A, S, i, id);
                    }
                }
                __syncthreads();
                if (S.listN >= static_cast<uint32_t>(kKnnFlush)) {
                    if (A.pairs) knn_emit_pairs(A, S, tid, q0);
                    else knn_flush<DM>(A, S, tid);
                }
            }
            if (A.pairs) knn_emit_pairs(A, S, tid, q0);
            else knn_flush<DM>(A, S, tid);
            knn_compact(S, tid);  // cluster-major: the seeds' top-K, refined later
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t q = q0 + i;
                if (q >= A.nq) continue;
                const uint32_t k = S.cnt[i];
                for (uint32_t e = tid; e < k; e += kKnnThreads) {
                    A.outIds[q * A.Kmax + e] = S.bv[i][e];
                    A.outD2[q * A.Kmax + e] = okey_inv(S.bk[i][e]);
                }
                if (tid == 0) {
                    A.outN[q] = k;
                    if (A.thr) {
                        A.thr[2 * q] = S.thrK[i];
                        A.thr[2 * q + 1] = S.thrV[i];
                    }
                    if (A.survStat) A.survStat[q * L + l] = S.surv[i];
                    if (A.fineStat) A.fineStat[q] = S.fine[i];
                }
            }
            __syncthreads();
        }
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
    const uint64_t* items;
    const uint64_t* nItems;
    uint64_t itemCap;
    const uint64_t* off;
    const uint32_t* qlist;
    const uint64_t* colRes;   // P[L-1] column pointers
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
};

constexpr int kRecIdBits = 40;

template <int DM>
__global__ __launch_bounds__(kStepWaves * kWave) void k_knn_cluster(KnnClusterArgs A) {
    __shared__ double s_x[kStepWaves][kWave * DM];
    __shared__ uint64_t s_id[kStepWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t gw = blockIdx.x * static_cast<uint64_t>(kStepWaves) + wave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = A.D;
    uint64_t nItems = *A.nItems;
    if (nItems > A.itemCap) nItems = A.itemCap;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    double* sx = s_x[wave];
    if (D != DM)
        for (int t = lane; t < kWave * DM; t += kWave) sx[t] = 0.0;
    for (uint64_t item = gw; item < nItems; item += nw) {
        const uint64_t it = A.items[item];
        const uint64_t c = it >> kSliceBits;
        const uint64_t start = (it & ((1ull << kSliceBits) - 1)) * kWave;
        const uint64_t g0 = A.off[c];
        const uint64_t gn = A.off[c + 1] - g0;
        const bool valid = start + lane < gn;
        const uint64_t q = valid ? A.qlist[g0 + start + lane] : 0;
        double qp[DM];
#pragma unroll
        for (int d = 0; d < DM; ++d) qp[d] = (d < D) ? A.points[q * D + d] : 0.0;
        const uint64_t tK = valid ? A.thr[2 * q] : 0, tV = valid ? A.thr[2 * q + 1] : 0;
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
                const uint64_t key = okey(acc);
                const uint64_t id = s_id[wave][k];
                const bool emit = valid && (key < tK || (key == tK && id < tV));
                const uint64_t m = __ballot(emit);
                if (m) {
                    uint64_t base = 0;
                    if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(A.recCount), __popcll(m));
                    base = __shfl(base, 0, kWave);
                    if (emit) {
                        const uint64_t pos = base + __popcll(m & ltmask);
                        if (pos < A.recCap) {
                            A.recK[pos] = key;
                            A.recV[pos] = (q << kRecIdBits) | id;
                            ++mine;
                        }
                    }
                    if (lane == 0 && base + __popcll(m) > A.recCap)
                        atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 4ull);
                }
            }
            wave_lds_sync();
        }
        if (valid && mine) atomicAdd(reinterpret_cast<unsigned long long*>(&A.cntQ[q]), mine);
    }
}

// Counting-sort scatter of the records into per-query segments; leaves cntQ zeroed.
__global__ void k_rec_scatter(const uint64_t* __restrict__ recK, const uint64_t* __restrict__ recV, const uint64_t* dn,
                              uint64_t cap, uint64_t* __restrict__ cntQ, const uint64_t* __restrict__ offQ,
                              uint64_t* __restrict__ gK, uint64_t* __restrict__ gV) {
    uint64_t n = *dn;
    if (n > cap) n = cap;
    const uint64_t idMask = (1ull << kRecIdBits) - 1;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t v = recV[i];
        const uint64_t q = v >> kRecIdBits;
        const uint64_t pos = atomicAdd(reinterpret_cast<unsigned long long*>(&cntQ[q]), ~0ull) - 1;
        gK[offQ[q] + pos] = recK[i];
        gV[offQ[q] + pos] = v & idMask;
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
        for (int i = 0; i < kKnnQB; ++i) {
            if (S.M[i] == 0) continue;
            const uint64_t r0 = s_r0[i], rn = s_rn[i];
            for (uint64_t b = 0; b < rn; b += kKnnThreads) {
                if (knn_needs_compact(S)) knn_compact(S, tid);
                const uint64_t r = b + tid;
                if (r < rn) knn_insert(S, i, A.gK[r0 + r], A.gV[r0 + r]);
                __syncthreads();
            }
        }
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

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_batch_set_param(tblaeq_batch* b, int param, uint64_t value) {
    if (!b) {
        set_error("tblaeq_batch_set_param: null engine");
        return TBLAEQ_ERR_INVALID;
    }
    switch (param) {
    case TBLAEQ_PARAM_KNN_SEED:
        if (value < 1 || value > static_cast<uint64_t>(kKnnMaxM)) break;
        b->knnSeed = value;
        return TBLAEQ_OK;
    case TBLAEQ_PARAM_KNN_WG_PER_CU:
        if (value < 1 || value > 16) break;
        b->knnWgPerCu = value;
        return TBLAEQ_OK;
    case TBLAEQ_PARAM_KNN_MODE:
        if (value > 1) break;
        b->knnMode = value;
        return TBLAEQ_OK;
    default:
        set_error("tblaeq_batch_set_param: unknown parameter");
        return TBLAEQ_ERR_INVALID;
    }
    set_error("tblaeq_batch_set_param: value out of range");
    return TBLAEQ_ERR_INVALID;
}

extern "C" int tblaeq_batch_get_param(const tblaeq_batch* b, int param, uint64_t* value) {
    if (!b || !value) {
        set_error("tblaeq_batch_get_param: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    switch (param) {
    case TBLAEQ_PARAM_KNN_SEED: *value = b->knnSeed; return TBLAEQ_OK;
    case TBLAEQ_PARAM_KNN_WG_PER_CU: *value = b->knnWgPerCu; return TBLAEQ_OK;
    case TBLAEQ_PARAM_KNN_MODE: *value = b->knnMode; return TBLAEQ_OK;
    default:
        set_error("tblaeq_batch_get_param: unknown parameter");
        return TBLAEQ_ERR_INVALID;
    }
}

extern "C" int tblaeq_batch_knn(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq,
                                uint64_t Kmax, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                                uint64_t* d_status, uint64_t* d_level_surv, uint64_t* d_fine_count) {
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
    // workspace: thr [2nq] | cntQ [nq] | offQ [nq+1] | flags [nq*stride bytes]
    const uint64_t words = 2 * nq + nq + nq + 1;
    int rc = ws_reserve(ctx, 8 * words + nq * stride + 64);
    if (rc) return rc;
    uint64_t* thr = static_cast<uint64_t*>(ctx->ws);
    uint64_t* cntQ = thr + 2 * nq;
    uint64_t* offQ = cntQ + nq;
    uint8_t* flags = reinterpret_cast<uint8_t*>(offQ + nq + 1);
    uint64_t* ctl = b->ctl;
    launch_fill(b, ctl, 0, 8);
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
    {
        const uint64_t grid = std::min<uint64_t>(groups, static_cast<uint64_t>(ctx->cuCount) * b->knnWgPerCu);
        TimedScope ts(ctx, kKBatch0 + 5);
        const dim3 g(static_cast<unsigned>(grid)), t(kKnnThreads);
        if (A.D <= 4) hipLaunchKernelGGL(k_knn_batch<4>, g, t, 0, ctx->stream, A);
        else if (A.D <= 8) hipLaunchKernelGGL(k_knn_batch<8>, g, t, 0, ctx->stream, A);
        else if (A.D <= 16) hipLaunchKernelGGL(k_knn_batch<16>, g, t, 0, ctx->stream, A);
        else hipLaunchKernelGGL(k_knn_batch<32>, g, t, 0, ctx->stream, A);
        TB_HIP(hipGetLastError());
    }
    if (clusterMajor) {
        if ((rc = launch_grouping(b, b->count[L - 1]))) return rc;
        KnnClusterArgs C{};
        C.D = b->D;
        C.items = b->items;
        C.nItems = ctl + 1;
        C.itemCap = b->itemCap;
        C.off = b->off;
        C.qlist = b->qlist;
        C.colRes = b->colRes[L - 1];
        C.coords = b->coords[L];
        C.ids = b->ids[L];
        C.points = d_points;
        C.thr = thr;
        C.recK = b->hits;
        C.recV = b->hitsTmp;
        C.recCap = std::min(b->hitCap, b->pairCap);
        C.recCount = ctl + 2;
        C.cntQ = cntQ;
        C.overflow = ctl + 3;
        {
            uint64_t blocks = std::min<uint64_t>((b->itemCap + kStepWaves - 1) / kStepWaves,
                                                 static_cast<uint64_t>(ctx->cuCount) * 16);
            const dim3 g(static_cast<unsigned>(blocks ? blocks : 1)), t(kStepWaves * kWave);
            TimedScope ts(ctx, kKBatch0 + 7);
            if (C.D <= 4) hipLaunchKernelGGL(k_knn_cluster<4>, g, t, 0, ctx->stream, C);
            else if (C.D <= 8) hipLaunchKernelGGL(k_knn_cluster<8>, g, t, 0, ctx->stream, C);
            else if (C.D <= 16) hipLaunchKernelGGL(k_knn_cluster<16>, g, t, 0, ctx->stream, C);
            else hipLaunchKernelGGL(k_knn_cluster<32>, g, t, 0, ctx->stream, C);
            TB_HIP(hipGetLastError());
        }
        if ((rc = launch_count_scan(b, cntQ, nq, offQ))) return rc;
        {
            TimedScope ts(ctx, kKBatch0 + 8);
            const uint64_t g = std::min<uint64_t>((C.recCap + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 8);
            hipLaunchKernelGGL(k_rec_scatter, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream,
                               b->hits, b->hitsTmp, ctl + 2, C.recCap, cntQ, offQ, b->pairs, b->recAux);
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
    launch_copy_status(b, d_status);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}
