// Fused prune + order-preserving compaction for the plain per-query path.
//
// One launch replaces the reference's mask kernel, host mask download and
// count, warp count kernel, two-kernel prefix sum (with its cudaMalloc/cudaFree)
// and scatter kernel (L0Strategy.cu:83-101,160-172; compact.cuh:34-268;
// GridCompact.cuh:41-100).  Tiles of 1024 points are taken from an atomic
// ticket so that a tile only ever waits on tiles already running; each tile
// publishes its survivor count and survivor child-count through a decoupled
// look-back, so survivors land in grid order together with their prolongation
// offsets, and the element count is read from device memory (no host sync).
#include <mutex>
#include <string>

#include "../ctx.hpp"
#include "../scan_common.hpp"

namespace tblaeq {

enum CompactMode { kModeRange = 0, kModeKnn = 1, kModeBox = 2 };

struct CompactArgs {
    int D;
    const uint64_t* ids;
    const double* vals;
    const uint64_t* dCount;
    uint64_t hostCount;
    int countOnHost;
    const double* a;       // lo (range/box) or query point (knn)
    const double* b;       // hi (range/box)
    const double* ub;      // knn STEP bound (device)
    const double* radius;  // level radius, indexed by id
    const uint64_t* colRes;
    uint64_t* outIds;
    double* outVals;
    uint64_t* outOff;
    uint64_t* dOutCount;
    uint64_t* dOutNnz;
    uint64_t* status;
    ScanCtrl* ctrl;
    uint64_t maxTiles;  // capacity of the status array (never exceeded, whatever the device count says)
    int radiusByPos;  // radius[i] belongs to grid position i (pre-gathered tile)
};

template <int MODE>
__global__ __launch_bounds__(kScanThreads) void k_compact(CompactArgs A) {
    __shared__ double s_a[64];
    __shared__ double s_b[64];
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_cnt[kScanItems][kScanThreads / kWave];
    __shared__ uint64_t s_wsum[kScanItems][kScanThreads / kWave];
    __shared__ uint64_t s_ex0, s_ex1;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int D = A.D;
    if (tid < D) {
        s_a[tid] = A.a[tid];
        if (MODE != kModeKnn) s_b[tid] = A.b[tid];
    }
    __syncthreads();
    const uint64_t n = A.countOnHost ? A.hostCount : *A.dCount;
    uint64_t ntiles = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
    if (ntiles > A.maxTiles) ntiles = A.maxTiles;
    const double ub = (MODE == kModeKnn) ? *A.ub : 0.0;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    const uint32_t epoch = launch_epoch(A.ctrl);

    for (;;) {
        if (tid == 0) s_tile = atomicAdd(reinterpret_cast<unsigned long long*>(&A.ctrl->tileCounter), 1ull);
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= ntiles) break;

        bool f[kScanItems];
        uint64_t w[kScanItems], id[kScanItems], winc[kScanItems], ball[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const uint64_t e = tile * kScanTile + static_cast<uint64_t>(i) * kScanThreads + tid;
            f[i] = false;
            w[i] = 0;
            id[i] = 0;
            if (e < n) {
                const double* v = A.vals + e * D;
                id[i] = A.ids[e];
                if (MODE == kModeRange) {
                    const double r = A.radiusByPos ? A.radius[e] : A.radius[id[i]];
                    f[i] = box_dist2(v, s_a, s_b, D) <= r * r;
                } else if (MODE == kModeKnn) {
                    f[i] = knn_lb(v, s_a, A.radius[id[i]], D) <= ub;
                } else {
                    f[i] = in_box(v, s_a, s_b, D);
                }
                if (MODE != kModeBox && f[i]) w[i] = A.colRes[id[i] + 1] - A.colRes[id[i]];
            }
            ball[i] = __ballot(f[i]);
            uint64_t x = w[i];
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint64_t y = __shfl_up(x, o, kWave);
                if (lane >= o) x += y;
            }
            winc[i] = x;
            if (lane == kWave - 1) s_wsum[i][wid] = x;
            if (lane == 0) s_cnt[i][wid] = __popcll(ball[i]);
        }
        __syncthreads();

        if (wid == 0) {
            constexpr int kSlots = kScanItems * (kScanThreads / kWave);
            uint64_t c = 0, s = 0;
            if (lane < kSlots) {
                c = s_cnt[lane / (kScanThreads / kWave)][lane % (kScanThreads / kWave)];
                s = s_wsum[lane / (kScanThreads / kWave)][lane % (kScanThreads / kWave)];
            }
            c = wave_sum_u64(c);
            s = wave_sum_u64(s);
            uint64_t* st = A.status + tile * 4;
            uint64_t ex0 = 0, ex1 = 0;
            if (tile == 0) {
                if (lane == 0) {
                    st_release(st + 2, status_word(epoch, c));
                    st_release(st + 3, status_word(epoch, s));
                }
            } else {
                if (lane == 0) {
                    st_release(st + 0, status_word(epoch, c));
                    st_release(st + 1, status_word(epoch, s));
                }
                lookback2(A.status, epoch, tile, lane, ex0, ex1);
                if (lane == 0) {
                    st_release(st + 2, status_word(epoch, ex0 + c));
                    st_release(st + 3, status_word(epoch, ex1 + s));
                }
            }
            if (lane == 0) {
                s_ex0 = ex0;
                s_ex1 = ex1;
                if (tile == ntiles - 1) {
                    *A.dOutCount = ex0 + c;
                    if (A.dOutNnz) *A.dOutNnz = ex1 + s;
                }
            }
        }
        __syncthreads();

        uint64_t pc = s_ex0, pw = s_ex1;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            uint64_t oc = pc, ow = pw;
            for (int ww = 0; ww < wid; ++ww) {
                oc += s_cnt[i][ww];
                ow += s_wsum[i][ww];
            }
            if (f[i]) {
                const uint64_t e = tile * kScanTile + static_cast<uint64_t>(i) * kScanThreads + tid;
                const uint64_t o = oc + __popcll(ball[i] & ltmask);
                A.outIds[o] = id[i];
                if (MODE != kModeBox) {
                    const double* v = A.vals + e * D;
                    double* dst = A.outVals + o * D;
                    for (int d = 0; d < D; ++d) dst[d] = v[d];
                    A.outOff[o] = ow + winc[i] - w[i];
                }
            }
            for (int ww = 0; ww < kScanThreads / kWave; ++ww) {
                pc += s_cnt[i][ww];
                pw += s_wsum[i][ww];
            }
        }
        __syncthreads();
    }

    if (tid == 0) scan_exit(A.ctrl);
}

// Launch helper shared by the C entry points.  `bound` sizes the grid; the
// element count comes from dCount (device) or hostCount.
int launch_compact(tblaeq_ctx* ctx, int mode, CompactArgs A, uint64_t bound) {
    const uint64_t tiles = bound == 0 ? 1 : (bound + kScanTile - 1) / kScanTile;
    int rc = scan_prepare(ctx, tiles);
    if (rc) return rc;
    A.status = ctx->status;
    A.ctrl = ctx->ctrl;
    A.maxTiles = tiles;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * 8;
    const unsigned grid = static_cast<unsigned>(tiles < cap ? tiles : cap);
    TimedScope ts(ctx, mode == kModeRange ? kKCompactRange : (mode == kModeKnn ? kKCompactKnn : kKCompactBox));
    switch (mode) {
        case kModeRange: hipLaunchKernelGGL(k_compact<kModeRange>, dim3(grid), dim3(kScanThreads), 0, ctx->stream, A); break;
        case kModeKnn: hipLaunchKernelGGL(k_compact<kModeKnn>, dim3(grid), dim3(kScanThreads), 0, ctx->stream, A); break;
        default: hipLaunchKernelGGL(k_compact<kModeBox>, dim3(grid), dim3(kScanThreads), 0, ctx->stream, A); break;
    }
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_range_prune_compact(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                          const double* vals, const uint64_t* d_count, uint64_t count_bound,
                                          const double* d_lo, const double* d_hi, uint64_t* surv_ids,
                                          double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                                          uint64_t* d_out_count) {
    if (!ctx || !lv || !d_count || !d_lo || !d_hi || !surv_ids || !surv_vals || !surv_off || !d_surv_count ||
        !d_out_count || lv->D == 0 || lv->D > 64) {
        set_error("tblaeq_range_prune_compact: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    CompactArgs A{};
    A.D = static_cast<int>(lv->D);
    A.ids = ids;
    A.vals = vals;
    A.dCount = d_count;
    A.a = d_lo;
    A.b = d_hi;
    A.radius = lv->radius;
    A.colRes = lv->col_res;
    A.outIds = surv_ids;
    A.outVals = surv_vals;
    A.outOff = surv_off;
    A.dOutCount = d_surv_count;
    A.dOutNnz = d_out_count;
    return launch_compact(ctx, kModeRange, A, count_bound);
}

namespace tblaeq {

int launch_compact_knn(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids, const double* vals,
                       const uint64_t* d_count, uint64_t count_bound, const double* d_q, const double* d_ub,
                       uint64_t* surv_ids, double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                       uint64_t* d_out_count) {
    CompactArgs A{};
    A.D = static_cast<int>(lv->D);
    A.ids = ids;
    A.vals = vals;
    A.dCount = d_count;
    A.a = d_q;
    A.ub = d_ub;
    A.radius = lv->radius;
    A.colRes = lv->col_res;
    A.outIds = surv_ids;
    A.outVals = surv_vals;
    A.outOff = surv_off;
    A.dOutCount = d_surv_count;
    A.dOutNnz = d_out_count;
    return launch_compact(ctx, kModeKnn, A, count_bound);
}

}  // namespace tblaeq

extern "C" int tblaeq_knn_radius_compact(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                         const double* vals, const uint64_t* d_count, uint64_t count_bound,
                                         const double* d_q, const double* d_radius, uint64_t* out_ids,
                                         double* out_vals, uint64_t* out_off, uint64_t* d_out_count,
                                         uint64_t* d_out_nnz) {
    if (!ctx || !lv || !ids || !vals || !d_count || !d_q || !d_radius || !out_ids || !out_vals || !out_off ||
        !d_out_count || !d_out_nnz) {
        tblaeq::set_error("tblaeq_knn_radius_compact: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return tblaeq::launch_compact_knn(ctx, lv, ids, vals, d_count, count_bound, d_q, d_radius, out_ids, out_vals,
                                      out_off, d_out_count, d_out_nnz);
}

namespace tblaeq {

int launch_compact_box(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals, uint64_t count,
                       const double* d_lo, const double* d_hi, uint64_t* out_ids, uint64_t* d_out_count) {
    CompactArgs A{};
    A.D = static_cast<int>(D);
    A.ids = ids;
    A.vals = vals;
    A.hostCount = count;
    A.countOnHost = 1;
    A.a = d_lo;
    A.b = d_hi;
    A.outIds = out_ids;
    A.dOutCount = d_out_count;
    return launch_compact(ctx, kModeBox, A, count);
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_range_prune_compact_tile(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                               const double* vals, uint64_t count, const double* d_lo,
                                               const double* d_hi, uint64_t* surv_ids, double* surv_vals,
                                               uint64_t* surv_off, uint64_t* d_surv_count, uint64_t* d_out_count) {
    if (!ctx || !lv || !d_lo || !d_hi || !surv_ids || !surv_vals || !surv_off || !d_surv_count || !d_out_count ||
        !lv->radius || !lv->col_res || lv->D == 0 || lv->D > 64) {
        set_error("tblaeq_range_prune_compact_tile: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    CompactArgs A{};
    A.D = static_cast<int>(lv->D);
    A.ids = ids;
    A.vals = vals;
    A.hostCount = count;
    A.countOnHost = 1;
    A.a = d_lo;
    A.b = d_hi;
    A.radius = lv->radius;
    A.radiusByPos = 1;
    A.colRes = lv->col_res;
    A.outIds = surv_ids;
    A.outVals = surv_vals;
    A.outOff = surv_off;
    A.dOutCount = d_surv_count;
    A.dOutNnz = d_out_count;
    return launch_compact(ctx, kModeRange, A, count);
}

namespace tblaeq {
__global__ __launch_bounds__(256) void k_knn_lb(int D, const uint64_t* __restrict__ ids,
                                                const double* __restrict__ vals, uint64_t n,
                                                const double* __restrict__ q, const double* __restrict__ radius,
                                                double* __restrict__ out) {
    __shared__ double s_q[64];
    if (threadIdx.x < D) s_q[threadIdx.x] = q[threadIdx.x];
    __syncthreads();
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) out[i] = knn_lb(vals + i * D, s_q, radius[ids[i]], D);
}
}  // namespace tblaeq

extern "C" int tblaeq_knn_lower_bounds(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals,
                                       uint64_t count, const double* d_q, const double* radius, double* out_lb) {
    if (!ctx || D == 0 || D > 64 || !d_q || !radius || (count && (!ids || !vals || !out_lb))) {
        set_error("tblaeq_knn_lower_bounds: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    if (count == 0) return TBLAEQ_OK;
    TimedScope ts(ctx, kKKnnLb);
    hipLaunchKernelGGL(k_knn_lb, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, ctx->stream,
                       static_cast<int>(D), ids, vals, count, d_q, radius, out_lb);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

// --------------------------------------------- allocation-free range compaction
//
// tblaeq_range_prune_compact_noalloc: the same kernel as
// tblaeq_range_prune_compact, with its look-back status words and scan control
// block in a file-scope device array instead of a context's allocations.  The
// array's device address is looked up once per device (hipGetSymbolAddress,
// that device current) and passed to the kernel; host code never uses the
// array's name as an address.  Every call on one device shares that device's
// array: the status words carry the launch epoch (advanced on the device by the
// last workgroup to leave), so consecutive calls need no clearing, and a fill
// kernel clears them once every 2^26 calls, before the 27-bit epoch could wrap
// onto a stale word.  Calls that overlap on one device race on the array.

namespace tblaeq {

constexpr uint64_t kNoAllocMaxTiles = 16384;  // grid points per call <= 16384 * 1024
constexpr uint64_t kNoAllocCtrlWords = sizeof(ScanCtrl) / sizeof(uint64_t);
constexpr uint64_t kNoAllocWords = kNoAllocCtrlWords + kNoAllocMaxTiles * 4;  // 512 KiB + 64 B

static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
    uint64_t* ptr = nullptr;  // device address of `scratch` on this device
    int cuCount = 0;
    uint64_t calls = 0;
};

__global__ void k_noalloc_clear(uint64_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = 0;
}

// Per-device set-up (first call on a device): the array's address on the
// current device and its CU count.  Returns null with the error recorded.
static NoAllocDevice* noalloc_device(int* rc) {
    static std::mutex mu;
    static NoAllocDevice devs[kNoAllocMaxDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *rc = hip_fail(e, "hipGetDevice");
        return nullptr;
    }
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("tblaeq_range_prune_compact_noalloc: device index out of range");
        *rc = TBLAEQ_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    NoAllocDevice& d = devs[dev];
    if (!d.ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) {
            *rc = hip_fail(e, "hipDeviceGetAttribute");
            return nullptr;
        }
        d.cuCount = cus > 0 ? cus : 1;
        d.ptr = static_cast<uint64_t*>(p);
    }
    return &d;
}

}  // namespace tblaeq

extern "C" int tblaeq_range_prune_compact_noalloc(void* stream, const tblaeq_level* lv, const uint64_t* ids,
                                                  const double* vals, const uint64_t* d_count, uint64_t count_bound,
                                                  const double* d_lo, const double* d_hi, uint64_t* surv_ids,
                                                  double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                                                  uint64_t* d_out_count) {
    if (!lv || !d_count || !d_lo || !d_hi || !surv_ids || !surv_vals || !surv_off || !d_surv_count ||
        !d_out_count || lv->D == 0 || lv->D > 64 || !lv->radius || !lv->col_res) {
        set_error("tblaeq_range_prune_compact_noalloc: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t tiles = count_bound == 0 ? 1 : (count_bound + kScanTile - 1) / kScanTile;
    if (tiles > kNoAllocMaxTiles) {
        set_error("tblaeq_range_prune_compact_noalloc: count_bound " + std::to_string(count_bound) +
                  " exceeds the fixed scratch (at most " + std::to_string(kNoAllocMaxTiles * kScanTile) +
                  " grid points per call)");
        return TBLAEQ_ERR_CAPACITY;
    }
    int rc = TBLAEQ_OK;
    NoAllocDevice* dev = noalloc_device(&rc);
    if (!dev) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint64_t calls;
    {
        static std::mutex mu;
        std::lock_guard<std::mutex> lock(mu);
        calls = ++dev->calls;
    }
    if (calls % (1ull << 26) == 0)
        hipLaunchKernelGGL(k_noalloc_clear, dim3(64), dim3(256), 0, s, dev->ptr + kNoAllocCtrlWords,
                           kNoAllocMaxTiles * 4);
    CompactArgs A{};
    A.D = static_cast<int>(lv->D);
    A.ids = ids;
    A.vals = vals;
    A.dCount = d_count;
    A.a = d_lo;
    A.b = d_hi;
    A.radius = lv->radius;
    A.colRes = lv->col_res;
    A.outIds = surv_ids;
    A.outVals = surv_vals;
    A.outOff = surv_off;
    A.dOutCount = d_surv_count;
    A.dOutNnz = d_out_count;
    A.ctrl = reinterpret_cast<ScanCtrl*>(dev->ptr);
    A.status = dev->ptr + kNoAllocCtrlWords;
    A.maxTiles = tiles;
    const uint64_t cap = static_cast<uint64_t>(dev->cuCount) * 8;
    const unsigned grid = static_cast<unsigned>(tiles < cap ? tiles : cap);
    hipLaunchKernelGGL(k_compact<kModeRange>, dim3(grid), dim3(kScanThreads), 0, s, A);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tblaeq_range_prune_compact_noalloc launch");
    return TBLAEQ_OK;
}
