// Internal definition of the C handle behind tblaeq_handler.h.
#pragma once
#include <memory>
#include <mutex>
#include <vector>

#include "../../../include/tblaeq_host.hpp"

namespace tblaeq_host {

// Device-resident batched engine state of a handler.
struct BatchState {
    tblaeq_batch* b = nullptr;
    // context (stream) the engine runs on: the index's for slot 0, a context of
    // its own (owned) for the other slots, so that slots run concurrently
    tblaeq_ctx* ctx = nullptr;
    bool ownsCtx = false;
    uint64_t* hStatus = nullptr;  // pinned status words of concurrent runs
    hipStream_t stream() const { return static_cast<hipStream_t>(tblaeq_ctx_stream(ctx)); }
    uint64_t maxQ = 0, pairCap = 0, hitCap = 0;
    int type = 0;
    uint64_t nq = 0, Kmax = 0;
    DeviceBuffer<double> dQ;
    DeviceBuffer<uint64_t> dK, dOff, dOffTmp, dIds, dStatus, dSurv, dFine, dN;
    DeviceBuffer<double> dD2;
    uint64_t lastTotal = 0;
    uint64_t work[2] = {0, 0};  // last run: status words 2 and 3 (summed over sub-batches)
    std::vector<double> hQ;     // host copy of the uploaded queries
    std::vector<uint64_t> hK;
    bool useGraph = true;
    // learned KNN: features, radii, fallback flags, and the fallback rerun's
    // gathered queries and rows (allocated for maxQ on first use)
    DeviceBuffer<float> lX;
    DeviceBuffer<double> lR, lQ, lD2;
    DeviceBuffer<uint8_t> lF;
    DeviceBuffer<uint64_t> lIdx, lK, lIds, lN, lSurv, lFine;
    ~BatchState() {
        if (b) tblaeq_batch_destroy(b);
        if (hStatus) (void)hipHostFree(hStatus);
        // the device buffers are freed after this body; hipFree needs no context
        if (ownsCtx && ctx) tblaeq_ctx_destroy(ctx);
    }
};

constexpr int kBatchSlots = 2;

}  // namespace tblaeq_host

struct tblaeq_handler {
    std::unique_ptr<tblaeq_host::QueryHandler> h;
    // batched engines (slots, each on its own stream), destroyed before the
    // index they use; the batch entry points act on slots[slot]
    std::unique_ptr<tblaeq_host::BatchState> slots[tblaeq_host::kBatchSlots];
    int slot = 0;
    // serialises the per-query traversal (its level buffers are shared): the
    // C entry points and a server's large-K requests take it
    std::mutex plainMu;
};

namespace tblaeq_host {
// Exact top-K (ids, squared distances) of one KNN query through the per-query
// traversal (any K), with its per-level survivors and candidate-superset size;
// takes the handler's plainMu.
void plainKnn(tblaeq_handler* h, const double* point, uint64_t K, std::vector<uint64_t>& ids,
              std::vector<double>& d2, std::vector<uint64_t>* levelSurv, uint64_t* fine);
}  // namespace tblaeq_host
