// Block merge sort, KNN lower-bound selection + STEP, and exact refine.
//
// Sorting is over (key, value) pairs compared lexicographically.  A tile of
// 2048 pairs (32 KB of LDS) is sorted by a bitonic network of 64-wide waves;
// tiles are merged by merge-path passes (each workgroup emits 2048 outputs after
// a binary search for its split on the diagonal).  When only the first `keep`
// pairs matter (KNN STEP's closest clusters, a top-K), each tile keeps its first
// `keep` pairs and tiles of survivors are sorted again until one tile remains.
//
// Replaces warpSort / blockMergeSortKernel / gpuSort / hostSerialMergeSort and
// the host STEP loop of the reference (MergeSort.cuh:162-313,326-444,628-696;
// KNN.cu:26-110; L0Strategy.cu:103-158).  The reference sort is unstable; here
// equal lower bounds are ordered by point id, so the selection is deterministic.
#include <vector>

#include "../ctx.hpp"

namespace tblaeq {

enum TileMode { kTilePairs = 0, kTileKnnLb = 1, kTileD2 = 2 };

struct TileArgs {
    int D;
    const uint64_t* inK;
    const uint64_t* inV;  // may be null (keys only; values read as 0)
    uint64_t n;
    const uint64_t* dCount;  // device element count (overrides n when set)
    const uint64_t* ids;
    const double* vals;
    const double* q;
    const double* radius;
    uint64_t keep;
    int noPad;       // write only the tile's real elements (full sorts), not `keep` padded entries
    uint64_t* outK;
    uint64_t* outV;  // may be null
};

__device__ __forceinline__ bool pair_gt(uint64_t ak, uint64_t av, uint64_t bk, uint64_t bv) {
    return ak > bk || (ak == bk && av > bv);
}

__device__ void bitonic_lds(uint64_t* sk, uint64_t* sv, int tid) {
    for (int k = 2; k <= kSortTile; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < kSortTile / 2; t += kSortThreads) {
                const int i = 2 * t - (t & (j - 1));
                const int p = i + j;
                const bool up = (i & k) == 0;
                const uint64_t ak = sk[i], av = sv[i], bk = sk[p], bv = sv[p];
                if (pair_gt(ak, av, bk, bv) == up) {
                    sk[i] = bk;
                    sv[i] = bv;
                    sk[p] = ak;
                    sv[p] = av;
                }
            }
            __syncthreads();
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(kSortThreads) void k_tilesort(TileArgs A) {
    __shared__ uint64_t sk[kSortTile];
    __shared__ uint64_t sv[kSortTile];
    __shared__ double s_q[64];
    const int tid = threadIdx.x;
    const int D = A.D;
    if (MODE != kTilePairs && tid < D) s_q[tid] = A.q[tid];
    __syncthreads();
    const uint64_t n = A.dCount ? *A.dCount : A.n;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
    for (int i = tid; i < kSortTile; i += kSortThreads) {
        const uint64_t e = base + i;
        uint64_t k = ~0ull, v = ~0ull;
        if (e < n) {
            if (MODE == kTilePairs) {
                k = A.inK[e];
                v = A.inV ? A.inV[e] : 0ull;
            } else {
                const uint64_t id = A.ids[e];
                const double* x = A.vals + e * D;
                const double val = (MODE == kTileKnnLb) ? knn_lb(x, s_q, A.radius[id], D) : point_dist2(x, s_q, D);
                k = okey(val);
                v = id;
            }
        }
        sk[i] = k;
        sv[i] = v;
    }
    __syncthreads();
    bitonic_lds(sk, sv, tid);
    const uint64_t ob = static_cast<uint64_t>(blockIdx.x) * A.keep;
    uint64_t lim = A.keep;
    if (A.noPad) lim = n > base ? (n - base < lim ? n - base : lim) : 0;
    for (uint64_t i = tid; i < lim; i += kSortThreads) {
        A.outK[ob + i] = sk[i];
        if (A.outV) A.outV[ob + i] = sv[i];
    }
}

// Number of elements taken from A among the first `diag` outputs of a stable
// merge (A wins ties).
__device__ uint64_t merge_split(const uint64_t* ak, const uint64_t* av, uint64_t na, const uint64_t* bk,
                                const uint64_t* bv, uint64_t nb, uint64_t diag) {
    uint64_t lo = diag > nb ? diag - nb : 0;
    uint64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const uint64_t bi = diag - mid - 1;
        const uint64_t avv = av ? av[mid] : 0, bvv = bv ? bv[bi] : 0;
        if (!pair_gt(ak[mid], avv, bk[bi], bvv)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSortThreads) void k_merge(const uint64_t* __restrict__ inK,
                                                        const uint64_t* __restrict__ inV, uint64_t* __restrict__ outK,
                                                        uint64_t* __restrict__ outV, uint64_t n, uint64_t width) {
    __shared__ uint64_t sk[kSortTile];
    __shared__ uint64_t sv[kSortTile];
    __shared__ uint64_t s_split[2];
    const int tid = threadIdx.x;
    const uint64_t outStart = static_cast<uint64_t>(blockIdx.x) * kSortTile;
    if (outStart >= n) return;
    const uint64_t a0 = (outStart / (2 * width)) * (2 * width);
    const uint64_t aEnd = a0 + width < n ? a0 + width : n;
    const uint64_t bEnd = a0 + 2 * width < n ? a0 + 2 * width : n;
    const uint64_t na = aEnd - a0, nb = bEnd - aEnd;
    const uint64_t d0 = outStart - a0;
    const uint64_t d1 = d0 + kSortTile < na + nb ? d0 + kSortTile : na + nb;
    const uint64_t* AK = inK + a0;
    const uint64_t* AV = inV ? inV + a0 : nullptr;
    const uint64_t* BK = inK + aEnd;
    const uint64_t* BV = inV ? inV + aEnd : nullptr;
    if (tid < 2) s_split[tid] = merge_split(AK, AV, na, BK, BV, nb, tid == 0 ? d0 : d1);
    __syncthreads();
    const uint64_t ia0 = s_split[0], ia1 = s_split[1];
    const uint64_t ib0 = d0 - ia0, ib1 = d1 - ia1;
    const int la = static_cast<int>(ia1 - ia0), lb = static_cast<int>(ib1 - ib0);
    for (int i = tid; i < la; i += kSortThreads) {
        sk[i] = AK[ia0 + i];
        sv[i] = AV ? AV[ia0 + i] : 0;
    }
    for (int i = tid; i < lb; i += kSortThreads) {
        sk[la + i] = BK[ib0 + i];
        sv[la + i] = BV ? BV[ib0 + i] : 0;
    }
    __syncthreads();
    constexpr int kPer = kSortTile / kSortThreads;  // 8 outputs per thread
    const int total = la + lb;
    const int diag = tid * kPer;
    if (diag >= total) return;
    // split of this thread's diagonal inside the LDS copies
    int lo = diag > lb ? diag - lb : 0, hi = diag < la ? diag : la;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int bi = diag - mid - 1;
        if (!pair_gt(sk[mid], sv[mid], sk[la + bi], sv[la + bi])) lo = mid + 1;
        else hi = mid;
    }
    int ia = lo, ib = diag - lo;
    const int cnt = total - diag < kPer ? total - diag : kPer;
    for (int r = 0; r < cnt; ++r) {
        bool takeA;
        if (ia >= la) takeA = false;
        else if (ib >= lb) takeA = true;
        else takeA = !pair_gt(sk[ia], sv[ia], sk[la + ib], sv[la + ib]);
        const int src = takeA ? ia++ : la + ib++;
        const uint64_t o = outStart + diag + r;
        outK[o] = sk[src];
        if (outV) outV[o] = sv[src];
    }
}

// STEP over the closest clusters, sorted by (lb, id) (KNN.cu:75-96 in closed
// form): phase 1 takes clusters while the exclusive running child count is
// below K and raises ub = max(ub, lb + 2r) starting from 0; every cluster with
// lb <= ub is kept afterwards (the compaction applies that test).
__global__ __launch_bounds__(kSortThreads) void k_step(const uint64_t* __restrict__ sk,
                                                       const uint64_t* __restrict__ sv, uint64_t m, uint64_t K,
                                                       const uint64_t* __restrict__ weight,
                                                       const double* __restrict__ radius, double* ubOut) {
    __shared__ uint64_t s_tot[kSortThreads / kWave];
    __shared__ double s_max[kSortThreads / kWave];
    __shared__ uint64_t s_carry;
    __shared__ int s_done;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    constexpr int kPer = 4;
    double best = 0.0;
    if (tid == 0) {
        s_carry = 0;
        s_done = 0;
    }
    __syncthreads();
    for (uint64_t base = 0; base < m; base += kSortThreads * kPer) {
        uint64_t w[kPer];
        double c[kPer];
        uint64_t tsum = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint64_t e = base + static_cast<uint64_t>(tid) * kPer + i;
            w[i] = 0;
            c[i] = -__builtin_inf();
            if (e < m && sv[e] != ~0ull) {
                const uint64_t id = sv[e];
                w[i] = weight[id];
                c[i] = okey_inv(sk[e]) + 2.0 * radius[id];
            }
            tsum += w[i];
        }
        // exclusive scan of per-thread sums across the block
        uint64_t x = tsum;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint64_t y = __shfl_up(x, o, kWave);
            if (lane >= o) x += y;
        }
        if (lane == kWave - 1) s_tot[wid] = x;
        __syncthreads();
        uint64_t ex = s_carry + x - tsum;
        for (int ww = 0; ww < wid; ++ww) ex += s_tot[ww];
        double mx = -__builtin_inf();
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            if (ex < K && c[i] > mx) mx = c[i];
            ex += w[i];
        }
        mx = wave_max_f64(mx);
        if (lane == 0) s_max[wid] = mx;
        __syncthreads();
        if (tid == 0) {
            uint64_t tot = 0;
            for (int ww = 0; ww < kSortThreads / kWave; ++ww) tot += s_tot[ww];
            s_carry += tot;
            if (s_carry >= K) s_done = 1;
        }
        for (int ww = 0; ww < kSortThreads / kWave; ++ww) best = fmax(best, s_max[ww]);
        __syncthreads();
        if (s_done) break;
    }
    if (tid == 0) *ubOut = best;
}

int launch_tilesort(tblaeq_ctx* ctx, int mode, const TileArgs& A, uint64_t bound) {
    const uint64_t tiles = bound == 0 ? 1 : (bound + kSortTile - 1) / kSortTile;
    const dim3 grid(static_cast<unsigned>(tiles)), block(kSortThreads);
    TimedScope ts(ctx, kKTileSort);
    switch (mode) {
        case kTilePairs: hipLaunchKernelGGL(k_tilesort<kTilePairs>, grid, block, 0, ctx->stream, A); break;
        case kTileKnnLb: hipLaunchKernelGGL(k_tilesort<kTileKnnLb>, grid, block, 0, ctx->stream, A); break;
        default: hipLaunchKernelGGL(k_tilesort<kTileD2>, grid, block, 0, ctx->stream, A); break;
    }
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

// Merge passes over tile-sorted data (tiles of 2048) in k/v; tmp is the same
// size.  Result ends in k/v.
int merge_passes(tblaeq_ctx* ctx, uint64_t* k, uint64_t* v, uint64_t* tk, uint64_t* tv, uint64_t n) {
    uint64_t* sk = k;
    uint64_t* svv = v;
    uint64_t* dk = tk;
    uint64_t* dv = tv;
    const unsigned blocks = static_cast<unsigned>((n + kSortTile - 1) / kSortTile);
    for (uint64_t width = kSortTile; width < n; width <<= 1) {
        {
            TimedScope ts(ctx, kKMerge);
            hipLaunchKernelGGL(k_merge, dim3(blocks), dim3(kSortThreads), 0, ctx->stream, sk, svv, dk, dv, n, width);
        }
        TB_HIP(hipGetLastError());
        std::swap(sk, dk);
        std::swap(svv, dv);
    }
    if (sk != k) {
        TB_HIP(hipMemcpyAsync(k, sk, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        if (v) TB_HIP(hipMemcpyAsync(v, svv, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return TBLAEQ_OK;
}

// Smallest `keep` pairs of the source described by A (mode KnnLb/D2/Pairs,
// `bound` elements at most), sorted, written to outK/outV (keep entries, padded
// with sentinel pairs).  Uses scratch starting at ws (4 * roundsCap words).
int select_smallest(tblaeq_ctx* ctx, int mode, TileArgs A, uint64_t bound, uint64_t keep, uint64_t* outK,
                    uint64_t* outV) {
    if (keep == 0) return TBLAEQ_OK;
    const uint64_t tiles0 = bound == 0 ? 1 : (bound + kSortTile - 1) / kSortTile;
    if (keep > kSortTile / 2) {
        // Full sort of everything, then the first `keep`.
        const uint64_t n = tiles0 * kSortTile;
        int rc = ws_reserve(ctx, n * 32);
        if (rc) return rc;
        uint64_t* k = static_cast<uint64_t*>(ctx->ws);
        uint64_t* v = k + n;
        A.keep = kSortTile;
        A.outK = k;
        A.outV = v;
        if ((rc = launch_tilesort(ctx, mode, A, bound))) return rc;
        if ((rc = merge_passes(ctx, k, v, v + n, v + 2 * n, n))) return rc;
        TB_HIP(hipMemcpyAsync(outK, k, keep * 8, hipMemcpyDeviceToDevice, ctx->stream));
        TB_HIP(hipMemcpyAsync(outV, v, keep * 8, hipMemcpyDeviceToDevice, ctx->stream));
        return TBLAEQ_OK;
    }
    uint64_t m = tiles0 * keep;
    int rc = ws_reserve(ctx, m * 32);
    if (rc) return rc;
    uint64_t* bufK[2] = {static_cast<uint64_t*>(ctx->ws), static_cast<uint64_t*>(ctx->ws) + m};
    uint64_t* bufV[2] = {bufK[1] + m, bufK[1] + 2 * m};
    A.keep = keep;
    A.outK = bufK[0];
    A.outV = bufV[0];
    if ((rc = launch_tilesort(ctx, mode, A, bound))) return rc;
    int cur = 0;
    uint64_t tiles = tiles0;
    while (tiles > 1) {
        TileArgs P{};
        P.inK = bufK[cur];
        P.inV = bufV[cur];
        P.n = m;
        P.keep = keep;
        P.outK = bufK[cur ^ 1];
        P.outV = bufV[cur ^ 1];
        if ((rc = launch_tilesort(ctx, kTilePairs, P, m))) return rc;
        tiles = (m + kSortTile - 1) / kSortTile;
        m = tiles * keep;
        cur ^= 1;
    }
    TB_HIP(hipMemcpyAsync(outK, bufK[cur], keep * 8, hipMemcpyDeviceToDevice, ctx->stream));
    TB_HIP(hipMemcpyAsync(outV, bufV[cur], keep * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return TBLAEQ_OK;
}

__global__ void k_fill_f64(double* p, double v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ void k_unkey(const uint64_t* k, double* d2, uint64_t n) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) d2[i] = okey_inv(k[i]);
}

}  // namespace tblaeq

using namespace tblaeq;

namespace tblaeq {
int launch_compact_knn(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids, const double* vals,
                       const uint64_t* d_count, uint64_t count_bound, const double* d_q, const double* d_ub,
                       uint64_t* surv_ids, double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                       uint64_t* d_out_count);
}

extern "C" int tblaeq_knn_lower_bound_sort_step(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                                const double* vals, const uint64_t* d_count, uint64_t count_bound,
                                                const double* d_q, uint64_t K, uint64_t* surv_ids,
                                                double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                                                uint64_t* d_out_count) {
    if (!ctx || !lv || !d_count || !d_q || !surv_ids || !surv_vals || !surv_off || !d_surv_count || !d_out_count ||
        lv->D == 0 || lv->D > 64) {
        set_error("tblaeq_knn_lower_bound_sort_step: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    // Phase 1 of STEP ends once the running child count reaches K, so it never
    // spans more than K non-empty clusters plus the empty ones of the level.
    uint64_t keep = K + lv->empty_cols;
    if (keep > count_bound) keep = count_bound;
    double* d_ub = &ctx->ctrl->ub;
    int rc;
    if (keep == 0) {
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, ctx->stream, d_ub, 0.0);
        TB_HIP(hipGetLastError());
    } else {
        const uint64_t tiles0 = count_bound == 0 ? 1 : (count_bound + kSortTile - 1) / kSortTile;
        const uint64_t selWords = keep > kSortTile / 2 ? tiles0 * kSortTile * 4 : tiles0 * keep * 4;
        rc = ws_reserve(ctx, (selWords + 2 * keep) * 8);
        if (rc) return rc;
        uint64_t* outK = static_cast<uint64_t*>(ctx->ws) + selWords;
        uint64_t* outV = outK + keep;
        TileArgs A{};
        A.D = static_cast<int>(lv->D);
        A.dCount = d_count;
        A.ids = ids;
        A.vals = vals;
        A.q = d_q;
        A.radius = lv->radius;
        rc = select_smallest(ctx, kTileKnnLb, A, count_bound, keep, outK, outV);
        if (rc) return rc;
        TimedScope ts(ctx, kKStep);
        hipLaunchKernelGGL(k_step, dim3(1), dim3(kSortThreads), 0, ctx->stream, outK, outV, keep, K,
                           lv->nnz_per_col, lv->radius, d_ub);
        TB_HIP(hipGetLastError());
    }
    return launch_compact_knn(ctx, lv, ids, vals, d_count, count_bound, d_q, d_ub, surv_ids, surv_vals, surv_off,
                              d_surv_count, d_out_count);
}

namespace tblaeq {
int launch_compact_box(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals, uint64_t count,
                       const double* d_lo, const double* d_hi, uint64_t* out_ids, uint64_t* d_out_count);
}

extern "C" int tblaeq_refine_range(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals,
                                   uint64_t count, const double* d_lo, const double* d_hi, uint64_t* out_ids,
                                   uint64_t out_cap, uint64_t* n_out) {
    if (!ctx || D == 0 || D > 64 || !d_lo || !d_hi || !n_out || (count && (!ids || !vals))) {
        set_error("tblaeq_refine_range: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t n = count == 0 ? 1 : count;
    int rc = ws_reserve(ctx, 2 * n * 8);
    if (rc) return rc;
    uint64_t* hits = static_cast<uint64_t*>(ctx->ws);
    uint64_t* tmp = hits + n;
    uint64_t* dcnt = &ctx->ctrl->aux;
    rc = launch_compact_box(ctx, D, ids, vals, count, d_lo, d_hi, hits, dcnt);
    if (rc) return rc;
    TB_HIP(hipMemcpyAsync(ctx->hostPinned, dcnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    TB_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t h = ctx->hostPinned[0];
    *n_out = h;
    if (h == 0) return TBLAEQ_OK;
    if (h > out_cap) {
        set_error("tblaeq_refine_range: output capacity too small");
        return TBLAEQ_ERR_CAPACITY;
    }
    TileArgs A{};
    A.inK = hits;
    A.n = h;
    A.keep = kSortTile;
    A.noPad = 1;
    A.outK = hits;
    if ((rc = launch_tilesort(ctx, kTilePairs, A, h))) return rc;
    const uint64_t tiles = (h + kSortTile - 1) / kSortTile;
    if (tiles > 1) {
        // merge passes need the values pointer null throughout (keys only)
        uint64_t* sk = hits;
        uint64_t* dk = tmp;
        const unsigned blocks = static_cast<unsigned>(tiles);
        for (uint64_t width = kSortTile; width < h; width <<= 1) {
            hipLaunchKernelGGL(k_merge, dim3(blocks), dim3(kSortThreads), 0, ctx->stream, sk, nullptr, dk, nullptr,
                               h, width);
            TB_HIP(hipGetLastError());
            std::swap(sk, dk);
        }
        TB_HIP(hipMemcpyAsync(out_ids, sk, h * 8, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        TB_HIP(hipMemcpyAsync(out_ids, hits, h * 8, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return TBLAEQ_OK;
}

extern "C" int tblaeq_refine_knn_topk(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals,
                                      uint64_t count, const double* d_q, uint64_t K, uint64_t* out_ids,
                                      double* out_d2, uint64_t* n_out) {
    if (!ctx || D == 0 || D > 64 || !d_q || !n_out || (K && (!out_ids || !out_d2)) || (count && (!ids || !vals))) {
        set_error("tblaeq_refine_knn_topk: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t k = K < count ? K : count;
    *n_out = k;
    if (k == 0) return TBLAEQ_OK;
    const uint64_t tiles0 = (count + kSortTile - 1) / kSortTile;
    const uint64_t selWords = k > kSortTile / 2 ? tiles0 * kSortTile * 4 : tiles0 * k * 4;
    int rc = ws_reserve(ctx, (selWords + 2 * k) * 8);
    if (rc) return rc;
    uint64_t* outK = static_cast<uint64_t*>(ctx->ws) + selWords;
    TileArgs A{};
    A.D = static_cast<int>(D);
    A.n = count;
    A.ids = ids;
    A.vals = vals;
    A.q = d_q;
    rc = select_smallest(ctx, kTileD2, A, count, k, outK, out_ids);
    if (rc) return rc;
    hipLaunchKernelGGL(k_unkey, dim3(static_cast<unsigned>((k + 255) / 256)), dim3(256), 0, ctx->stream, outK,
                       out_d2, k);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

extern "C" int tblaeq_sort_pairs(tblaeq_ctx* ctx, uint64_t* keys, uint64_t* vals, uint64_t n) {
    if (!ctx || (n && (!keys || !vals))) {
        set_error("tblaeq_sort_pairs: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    if (n <= 1) return TBLAEQ_OK;
    int rc = ws_reserve(ctx, n * 16);
    if (rc) return rc;
    TileArgs A{};
    A.inK = keys;
    A.inV = vals;
    A.n = n;
    A.keep = kSortTile;
    A.noPad = 1;  // each workgroup reads its whole tile into LDS first, so in place is safe
    A.outK = keys;
    A.outV = vals;
    if ((rc = launch_tilesort(ctx, kTilePairs, A, n))) return rc;
    uint64_t* tk = static_cast<uint64_t*>(ctx->ws);
    return merge_passes(ctx, keys, vals, tk, tk + n, n);
}
