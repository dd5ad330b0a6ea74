// Internal context shared by the C-ABI entry points (not part of the public header).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tblaeq.h"
#include "device_common.hpp"

namespace tblaeq {
// Kernel families timed by HIP events when timing is enabled on a context.
enum KernelKind {
    kKCompactRange = 0,
    kKCompactKnn = 1,
    kKCompactBox = 2,
    kKProlong = 3,
    kKTileSort = 4,
    kKMerge = 5,
    kKStep = 6,
    kKKnnLb = 7,
    kKBatch0 = 8,  // batched-path kernels use kKBatch0 + i
    kKindCount = 32
};
constexpr int kTagCount = 16;
struct TimedLaunch {
    int kind, tag;
    hipEvent_t a, b;
};
}  // namespace tblaeq

struct tblaeq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownsStream = false;
    uint64_t scanLaunches = 0;       // scan launches issued (status words are cleared before the epoch wraps)
    uint64_t* status = nullptr;      // 4 words per tile (agg0, agg1, inc0, inc1)
    uint64_t statusTiles = 0;
    tblaeq::ScanCtrl* ctrl = nullptr;  // [kCtrlSlots]
    // growable scratch (sort ping-pong buffers, KNN selection lists)
    void* ws = nullptr;
    size_t wsBytes = 0;
    uint64_t* hostPinned = nullptr;  // small pinned staging for counts
    int cuCount = 0;
    // HIP-event timing of kernel launches on `stream` (off unless enabled)
    bool timing = false;
    int tag = 0;
    std::vector<tblaeq::TimedLaunch> pending;
    std::vector<hipEvent_t> eventPool;
    double accMs[tblaeq::kKindCount][tblaeq::kTagCount] = {};
    uint64_t accN[tblaeq::kKindCount][tblaeq::kTagCount] = {};
};

namespace tblaeq {

constexpr int kCtrlSlots = 8;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;  // 1024 elements per look-back tile
constexpr int kSortTile = 2048;                         // pairs per LDS tile sort
constexpr int kSortThreads = 256;

void set_error(const std::string& msg);
// Sends a printf-style progress message to the callback set by
// tblaeq_set_progress (nothing when none is set).
void progress(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_fail(hipError_t e, const char* what);
// Ensures the look-back status array can hold `tiles` tiles (epochs advance on the device).
int scan_prepare(tblaeq_ctx* ctx, uint64_t tiles);
// Grows the context scratch to at least `bytes`.
int ws_reserve(tblaeq_ctx* ctx, size_t bytes);
// out[i] = the low 32 bits of in[i], i < n, on stream st (ids that fit in 32
// bits, narrowed before they cross to the host).
hipError_t narrow_ids(hipStream_t st, const uint64_t* in, uint32_t* out, uint64_t n);

// Brackets one kernel launch with HIP events on the context stream when timing is on.
struct TimedScope {
    TimedScope(tblaeq_ctx* c, int kind);
    ~TimedScope();
    tblaeq_ctx* ctx;
    int idx = -1;
};

}  // namespace tblaeq

#define TB_HIP(call)                                               \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return tblaeq::hip_fail(e_, #call);  \
    } while (0)
