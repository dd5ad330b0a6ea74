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
// The last level is fused with the exact top-K refine: the children of the
// phase-1 clusters (at least K points) give an upper bound rho on the K-th
// nearest distance, and a surviving cluster is expanded only if its lower bound
// is within rho (plus a relative margin of 1e-9 far above the rounding error of
// the bounds), since every child x of cluster c satisfies |q-x| >= |q-c| - r_c.
// The top-K is then the same set the per-query path computes over all
// survivors' children, ordered by (squared distance, id).
#include <algorithm>

#include "../batch_internal.hpp"

namespace tblaeq {

constexpr int kKnnQB = 4;        // queries per workgroup
constexpr int kKnnCap = 512;     // LDS candidates per query
constexpr int kKnnMaxM = kKnnCap - 256;
constexpr int kKnnThreads = 256;
constexpr int kKnnList = 1024;   // clusters queued for expansion per round

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
    uint64_t* p1;           // per query phase-1 cluster ids (kKnnMaxM each)
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
    uint32_t listPos[kKnnList];
    uint8_t listQ[kKnnList];
    uint32_t listN;
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

__device__ double knn_lb_s(const double* x, const double* q, double r, int D) {
    double dist = 0.0;
    for (int d = 0; d < D; ++d) {
        const double diff = q[d] - x[d];
        dist = __builtin_fma(diff, diff, dist);
    }
    return sqrt_rd(dist) - r;
}

__device__ double knn_d2_s(const double* x, const double* q, int D) {
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
        const double diff = q[d] - x[d];
        acc = __builtin_fma(diff, diff, acc);
    }
    return acc;
}

// Inserts the children (finest positions) of cluster `cid` (level L-1) for query slot i.
__device__ void knn_expand(const KnnArgs& A, KnnShared& S, int tid, int i, uint64_t cid) {
    const int D = A.D, L = A.L;
    const uint64_t cs = A.colRes[L - 1][cid], ce = A.colRes[L - 1][cid + 1];
    for (uint64_t c0 = cs; c0 < ce; c0 += kKnnThreads) {
        if (knn_needs_compact(S)) knn_compact(S, tid);
        const uint64_t j = c0 + tid;
        if (j < ce) {
            const double d2 = knn_d2_s(A.coords[L] + j * D, S.qp[i], D);
            knn_insert(S, i, okey(d2), A.ids[L][j]);
        }
        __syncthreads();
    }
}

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
        for (int t = tid; t < kKnnQB * D; t += kKnnThreads) {
            const int i = t / D, d = t % D;
            const uint64_t q = q0 + i;
            S.qp[i][d] = q < A.nq ? A.points[q * D + d] : 0.0;
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
                // K + empty clusters bound phase 1; one more entry gives the first
                // cluster after phase 1 (the boundary used at the last level)
                uint64_t m = q < A.nq ? S.K[tid] + A.emptyCols[l] + 1 : 0;
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
                    const double* x = X + j * D;
                    const double r = R[j];
                    const uint64_t id = I[j];
                    for (int i = 0; i < kKnnQB; ++i) {
                        const uint64_t q = q0 + i;
                        if (q >= A.nq) continue;
                        if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                        knn_insert(S, i, okey(knn_lb_s(x, S.qp[i], r, D)), id);
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
                        const double* x = X + j * D;
                        const double r = R[j];
                        const uint64_t id = I[j];
                        for (int i = 0; i < kKnnQB; ++i) {
                            const uint64_t q = q0 + i;
                            if (q >= A.nq) continue;
                            bool s = true;
                            if (l > 0) s = A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]] != 0;
                            if (s) s = knn_lb_s(x, S.qp[i], r, D) <= S.ub[i];
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
            // ---- last level: phase-1 clusters to global scratch, then top-K over their children
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t q = q0 + i;
                if (q >= A.nq) continue;
                for (uint32_t e = tid; e < S.i0[i]; e += kKnnThreads) A.p1[q * kKnnMaxM + e] = S.bv[i][e];
            }
            __syncthreads();
            if (tid < kKnnQB) {
                S.M[tid] = static_cast<uint32_t>(q0 + tid < A.nq ? S.K[tid] : 0);
                S.cnt[tid] = 0;
                S.thrK[tid] = ~0ull;
                S.thrV[tid] = ~0ull;
            }
            __syncthreads();
            for (int i = 0; i < kKnnQB; ++i) {
                const uint64_t q = q0 + i;
                if (q >= A.nq || S.M[i] == 0) continue;
                const uint32_t np1 = S.i0[i];
                for (uint32_t e = 0; e < np1; ++e) knn_expand(A, S, tid, i, A.p1[q * kKnnMaxM + e]);
            }
            // ---- pass 2 of the last level: survivors outside phase 1 within the certified radius
            if (tid == 0) S.listN = 0;
            __syncthreads();
            for (uint64_t j0 = 0; j0 < n; j0 += kKnnThreads) {
                const uint64_t j = j0 + tid;
                if (j < n) {
                    const double* x = X + j * D;
                    const double r = R[j];
                    const uint64_t id = I[j];
                    for (int i = 0; i < kKnnQB; ++i) {
                        const uint64_t q = q0 + i;
                        if (q >= A.nq) continue;
                        if (l > 0 && !A.flags[q * A.flagStride + A.flagOff[l - 1] + Pp[j]]) continue;
                        const double lb = knn_lb_s(x, S.qp[i], r, D);
                        if (!(lb <= S.ub[i])) continue;
                        atomicAdd(reinterpret_cast<unsigned long long*>(&S.surv[i]), 1ull);
                        atomicAdd(reinterpret_cast<unsigned long long*>(&S.fine[i]), A.weightPos[l][j]);
                        if (S.M[i] == 0) continue;
                        if (lt_pair(okey(lb), id, S.bndK[i], S.bndV[i])) continue;  // phase-1 cluster, done
                        const uint64_t tk = S.thrK[i];
                        if (tk != ~0ull) {
                            const double rho = __builtin_sqrt(okey_inv(tk));
                            if (lb > rho + 1e-9 * (1.0 + rho + r)) continue;
                        }
                        const uint32_t p = atomicAdd(&S.listN, 1u);
                        if (p < kKnnList) {
                            S.listPos[p] = static_cast<uint32_t>(tid);
                            S.listQ[p] = static_cast<uint8_t>(i);
                        }
                    }
                }
                __syncthreads();
                const uint32_t ln = S.listN < kKnnList ? S.listN : kKnnList;
                for (uint32_t e = 0; e < ln; ++e) {
                    const uint64_t pj = j0 + S.listPos[e];
                    knn_expand(A, S, tid, S.listQ[e], I[pj]);
                }
                __syncthreads();
                if (tid == 0) S.listN = 0;
                __syncthreads();
            }
            knn_compact(S, tid);
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
                    if (A.survStat) A.survStat[q * L + l] = S.surv[i];
                    if (A.fineStat) A.fineStat[q] = S.fine[i];
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_batch_knn(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq,
                                uint64_t Kmax, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                                uint64_t* d_level_surv, uint64_t* d_fine_count) {
    if (!b || !d_points || !d_K || nq == 0 || !d_out_ids || !d_out_d2 || !d_out_n) {
        set_error("tblaeq_batch_knn: invalid argument");
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
    tblaeq_ctx* ctx = b->ctx;
    KnnArgs A{};
    uint64_t stride = 0;
    for (int l = 0; l + 1 < L; ++l) {
        A.flagOff[l] = stride;
        stride += b->count[l];
    }
    if (stride == 0) stride = 1;
    const uint64_t flagBytes = nq * stride;
    const uint64_t p1Words = nq * kKnnMaxM;
    int rc = ws_reserve(ctx, 8 * p1Words + flagBytes + 64);
    if (rc) return rc;
    uint64_t* p1 = static_cast<uint64_t*>(ctx->ws);
    uint8_t* flags = reinterpret_cast<uint8_t*>(p1 + p1Words);
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
    A.p1 = p1;
    A.outIds = d_out_ids;
    A.outD2 = d_out_d2;
    A.outN = d_out_n;
    A.survStat = d_level_surv;
    A.fineStat = d_fine_count;
    const uint64_t groups = (nq + kKnnQB - 1) / kKnnQB;
    uint64_t grid = std::min<uint64_t>(groups, static_cast<uint64_t>(ctx->cuCount) * 4);
    TimedScope ts(ctx, kKBatch0 + 5);
    hipLaunchKernelGGL(k_knn_batch, dim3(static_cast<unsigned>(grid)), dim3(kKnnThreads), 0, ctx->stream, A);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}
