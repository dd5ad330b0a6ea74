// Context management and error reporting for the C ABI.
#include "ctx.hpp"

#include <cstring>
#include <string>

#include <cstdarg>
#include <cstdio>
#include <mutex>

namespace {
thread_local std::string g_lastError;
std::mutex g_progressMu;
tblaeq_progress_fn g_progress = nullptr;
void* g_progressUser = nullptr;
}

namespace tblaeq {

void set_error(const std::string& msg) { g_lastError = msg; }

void progress(const char* fmt, ...) {
    tblaeq_progress_fn fn;
    void* user;
    {
        std::lock_guard<std::mutex> lk(g_progressMu);
        fn = g_progress;
        user = g_progressUser;
    }
    if (!fn) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    fn(buf, user);
}

int hip_fail(hipError_t e, const char* what) {
    g_lastError = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? TBLAEQ_ERR_NOMEM : TBLAEQ_ERR_HIP;
}

int scan_prepare(tblaeq_ctx* ctx, uint64_t tiles) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(ctx->stream, &cap);
    if (tiles > ctx->statusTiles) {
        if (cap != hipStreamCaptureStatusNone) {
            set_error("look-back status array too small during graph capture (reserve first)");
            return TBLAEQ_ERR_CAPACITY;
        }
        uint64_t n = ctx->statusTiles ? ctx->statusTiles : 1024;
        while (n < tiles) n *= 2;
        TB_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->status) TB_HIP(hipFree(ctx->status));
        ctx->status = nullptr;
        ctx->statusTiles = 0;
        TB_HIP(hipMalloc(&ctx->status, n * 4 * sizeof(uint64_t)));
        TB_HIP(hipMemsetAsync(ctx->status, 0, n * 4 * sizeof(uint64_t), ctx->stream));
        ctx->statusTiles = n;
    }
    // The device epoch has 27 bits; clear every status word well before it can
    // wrap onto a value still present in the array (outside graph capture).
    if (++ctx->scanLaunches % (1ull << 26) == 0 && cap == hipStreamCaptureStatusNone)
        TB_HIP(hipMemsetAsync(ctx->status, 0, ctx->statusTiles * 4 * sizeof(uint64_t), ctx->stream));
    return TBLAEQ_OK;
}

int ws_reserve(tblaeq_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->wsBytes) return TBLAEQ_OK;
    size_t cap = ctx->wsBytes ? ctx->wsBytes : (1u << 20);
    while (cap < bytes) cap *= 2;
    if (ctx->ws) {
        TB_HIP(hipStreamSynchronize(ctx->stream));
        TB_HIP(hipFree(ctx->ws));
    }
    ctx->ws = nullptr;
    ctx->wsBytes = 0;
    TB_HIP(hipMalloc(&ctx->ws, cap));
    ctx->wsBytes = cap;
    return TBLAEQ_OK;
}

static hipEvent_t take_event(tblaeq_ctx* ctx) {
    if (!ctx->eventPool.empty()) {
        hipEvent_t e = ctx->eventPool.back();
        ctx->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedScope::TimedScope(tblaeq_ctx* c, int kind) : ctx(c) {
    if (!ctx->timing) return;
    TimedLaunch t{kind, ctx->tag, take_event(ctx), take_event(ctx)};
    if (!t.a || !t.b) return;
    (void)hipEventRecord(t.a, ctx->stream);
    ctx->pending.push_back(t);
    idx = static_cast<int>(ctx->pending.size()) - 1;
}

TimedScope::~TimedScope() {
    if (idx >= 0) (void)hipEventRecord(ctx->pending[idx].b, ctx->stream);
}

static int collect_timing(tblaeq_ctx* ctx) {
    if (ctx->pending.empty()) return TBLAEQ_OK;
    TB_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess && t.kind >= 0 && t.kind < kKindCount) {
            const int tg = t.tag < 0 ? 0 : (t.tag >= kTagCount ? kTagCount - 1 : t.tag);
            ctx->accMs[t.kind][tg] += ms;
            ctx->accN[t.kind][tg] += 1;
        }
        ctx->eventPool.push_back(t.a);
        ctx->eventPool.push_back(t.b);
    }
    ctx->pending.clear();
    return TBLAEQ_OK;
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_ctx_set_timing(tblaeq_ctx* ctx, int enable) {
    if (!ctx) return TBLAEQ_ERR_INVALID;
    if (!enable) {
        int rc = collect_timing(ctx);
        if (rc) return rc;
    }
    ctx->timing = enable != 0;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_set_tag(tblaeq_ctx* ctx, int tag) {
    if (!ctx) return TBLAEQ_ERR_INVALID;
    ctx->tag = tag;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_timing(tblaeq_ctx* ctx, int kind, int tag, uint64_t* launches, double* ms) {
    if (!ctx || kind < 0 || kind >= kKindCount) return TBLAEQ_ERR_INVALID;
    int rc = collect_timing(ctx);
    if (rc) return rc;
    uint64_t n = 0;
    double t = 0;
    for (int g = 0; g < kTagCount; ++g) {
        if (tag >= 0 && g != tag) continue;
        n += ctx->accN[kind][g];
        t += ctx->accMs[kind][g];
    }
    if (launches) *launches = n;
    if (ms) *ms = t;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_timing_reset(tblaeq_ctx* ctx) {
    if (!ctx) return TBLAEQ_ERR_INVALID;
    int rc = collect_timing(ctx);
    if (rc) return rc;
    for (int k = 0; k < kKindCount; ++k)
        for (int g = 0; g < kTagCount; ++g) {
            ctx->accMs[k][g] = 0;
            ctx->accN[k][g] = 0;
        }
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_create(int device, void* stream, tblaeq_ctx** out) {
    if (!out) {
        set_error("tblaeq_ctx_create: null output");
        return TBLAEQ_ERR_INVALID;
    }
    *out = nullptr;
    TB_HIP(hipSetDevice(device));
    auto* ctx = new tblaeq_ctx();
    ctx->device = device;
    int rc = TBLAEQ_OK;
    do {
        hipError_t e;
        if (stream) {
            ctx->stream = static_cast<hipStream_t>(stream);
        } else {
            e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
            if (e != hipSuccess) { rc = hip_fail(e, "hipStreamCreate"); break; }
            ctx->ownsStream = true;
        }
        e = hipDeviceGetAttribute(&ctx->cuCount, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess) { rc = hip_fail(e, "hipDeviceGetAttribute"); break; }
        e = hipMalloc(&ctx->ctrl, kCtrlSlots * sizeof(ScanCtrl));
        if (e != hipSuccess) { rc = hip_fail(e, "hipMalloc(ctrl)"); break; }
        e = hipMemsetAsync(ctx->ctrl, 0, kCtrlSlots * sizeof(ScanCtrl), ctx->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "hipMemset(ctrl)"); break; }
        e = hipHostMalloc(reinterpret_cast<void**>(&ctx->hostPinned), 64 * sizeof(uint64_t), hipHostMallocDefault);
        if (e != hipSuccess) { rc = hip_fail(e, "hipHostMalloc"); break; }
        rc = scan_prepare(ctx, 1024);
        if (rc) break;
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) { rc = hip_fail(e, "hipStreamSynchronize"); break; }
    } while (false);
    if (rc) {
        tblaeq_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_destroy(tblaeq_ctx* ctx) {
    if (!ctx) return TBLAEQ_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->pending) {
        ctx->eventPool.push_back(t.a);
        ctx->eventPool.push_back(t.b);
    }
    for (auto e : ctx->eventPool) (void)hipEventDestroy(e);
    if (ctx->status) (void)hipFree(ctx->status);
    if (ctx->ctrl) (void)hipFree(ctx->ctrl);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->hostPinned) (void)hipHostFree(ctx->hostPinned);
    if (ctx->ownsStream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_ctx_sync(tblaeq_ctx* ctx) {
    if (!ctx) return TBLAEQ_ERR_INVALID;
    TB_HIP(hipStreamSynchronize(ctx->stream));
    return TBLAEQ_OK;
}

extern "C" void* tblaeq_ctx_stream(tblaeq_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

extern "C" const char* tblaeq_status_string(int s) {
    switch (s) {
        case TBLAEQ_OK: return "ok";
        case TBLAEQ_ERR_INVALID: return "invalid argument";
        case TBLAEQ_ERR_HIP: return "HIP runtime error";
        case TBLAEQ_ERR_UNSUPPORTED: return "unsupported shape";
        case TBLAEQ_ERR_CAPACITY: return "buffer too small";
        case TBLAEQ_ERR_IO: return "index file error";
        case TBLAEQ_ERR_NOMEM: return "out of memory";
        case TBLAEQ_ERR_TIMEOUT: return "timed out";
        case TBLAEQ_ERR_INTERNAL: return "internal error";
        default: return "unknown status";
    }
}

extern "C" const char* tblaeq_last_error(void) { return g_lastError.c_str(); }

extern "C" void tblaeq_set_progress(tblaeq_progress_fn fn, void* user) {
    std::lock_guard<std::mutex> lk(g_progressMu);
    g_progress = fn;
    g_progressUser = user;
}
