// Serving loop: continuous batching of range and KNN requests (tblaeq_serve.h).
//
// One worker thread per request kind owns a batched engine with its own HIP
// stream, device query / result buffers and pinned staging.  Requests wait in
// a FIFO per kind; a worker admits the waiting requests into its next batch as
// soon as its previous batch has completed on the GPU (the range worker's
// results of one batch reach the host and their requests while the next one
// runs), bounded by max_batch and by the in-flight HBM budget shared by both
// workers.  A batch whose scratch would
// overflow the engine's capacities is split in halves and rerun, exactly as
// the handler's batched driver does, and the worker's next batches start at
// the size that fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/tblaeq_serve.h"
#include "../ctx.hpp"
#include "handler_internal.hpp"
#include "json.hpp"

using namespace tblaeq_host;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    int type = 0;
    std::vector<double> q;
    uint64_t K = 0;
    uint64_t est = 0;  // admission estimate (bytes)
    Clock::time_point submitted, finished;
    bool done = false;
    int status = TBLAEQ_OK;
    std::vector<uint64_t> ids;
    std::vector<double> d2;
    // traversal requests run the per-query traversal on the worker and keep its
    // fine mesh (ids) and per-level sizes instead of the exact answer
    bool traversal = false;
    std::vector<uint64_t> levelSizes, levelOriginal;
    uint64_t queryUs = 0, totalUs = 0, queryCount = 0;
    std::string policy;
    std::string error;  // message of a failed request
    bool orphan = false;  // its caller stopped waiting: dropped when it completes
};

template <class T>
struct Pinned {
    T* p = nullptr;
    size_t n = 0;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t count) {
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T), 0),
                 "hipHostMalloc");
        n = count;
    }
};

// Range results of one batch on their way to the host.  The range worker
// keeps two: a batch's ids (narrowed to 32 bits when every id of the index
// fits) are copied right after the batch on the worker's stream and scattered
// into its requests on the host while the next batch runs on the GPU.  (The
// copies stay on the batch's stream: on a stream of their own they may share
// a hardware queue with the next batch and wait for it.)
struct Flight {
    DeviceBuffer<uint64_t> dOff, dIds;
    DeviceBuffer<uint32_t> dIds32;
    Pinned<uint64_t> hOff, hIds;
    Pinned<uint32_t> hIds32;
    hipEvent_t copied = nullptr;
    // the batch it carries
    std::vector<Request*> batch, batched;
    uint64_t hits = 0, est = 0, travs = 0;
};

struct Worker {
    int type = 0;
    tblaeq_ctx* ctx = nullptr;
    tblaeq_batch* b = nullptr;
    DeviceBuffer<double> dQ, dD2;
    DeviceBuffer<uint64_t> dK, dIds, dN, dStatus;
    Pinned<double> hQ, hD2;
    Pinned<uint64_t> hK, hIds, hN, hStatus;
    std::thread th;
    double hitsPerRequest = 0.0;  // range worker: running mean of hits per request (batch sizing)
    // batch size limit after a capacity overflow: half the batch that
    // overflowed, grown by a quarter after each batch that fits (0: none)
    size_t sizeLimit = 0;
    // range worker: result flights, ids sent as 32 bits
    Flight fl[2];
    int cur = 0;
    bool narrow = false;
    // statistics (under the server lock after each batch)
    uint64_t completed = 0, batches = 0, reruns = 0, traversals = 0;
    double busyUs = 0.0;
};

// Copies out[i] = src[off[i] .. off[i+1]) for n requests, on several threads
// when there are many ids.
template <class Id>
void scatter_results(const Id* src, const uint64_t* off, size_t n, std::vector<Request*>& batch, size_t a) {
    const uint64_t total = off[n] - off[0];
    const unsigned threads = total < (4ull << 20) ? 1u : 8u;
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) batch[a + i]->ids.assign(src + off[i], src + off[i + 1]);
    };
    if (threads == 1) {
        work(0, n);
        return;
    }
    // request ranges of about equal id counts
    std::vector<size_t> cut(threads + 1, n);
    cut[0] = 0;
    size_t i = 0;
    for (unsigned t = 1; t < threads; ++t) {
        const uint64_t target = off[0] + total * t / threads;
        while (i < n && off[i] < target) ++i;
        cut[t] = i;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        if (cut[t] < cut[t + 1]) pool.emplace_back(work, cut[t], cut[t + 1]);
    for (auto& th : pool) th.join();
}

}  // namespace

struct tblaeq_server {
    tblaeq_handler* h = nullptr;
    tblaeq_server_config cfg{};
    size_t D = 0;
    uint64_t maxK = 0;
    Worker w[2];
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Request*> queue[2];
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t nextTicket = 1;
    bool stop = false;
    uint64_t budget = 0, inflight = 0, peakInflight = 0;
    // what the JSON front end reports about the served index
    std::string datasetName, policy;
    std::vector<uint64_t> levelOriginal;
    // statistics
    std::vector<double> latUs;
    uint64_t batches = 0, batchedReqs = 0, deferred = 0;
};

namespace {

// Stages requests [a, a+n) of `batch` and launches the engine batch on the
// worker's stream (range results into flight f's device buffers); the status
// words follow into w.hStatus.  Returns Kmax for KNN.
uint64_t launch_batch(tblaeq_server* s, Worker& w, Flight* f, std::vector<Request*>& batch, size_t a, size_t n) {
    hipStream_t st = static_cast<hipStream_t>(tblaeq_ctx_stream(w.ctx));
    uint64_t Kmax = 0;
    for (size_t i = 0; i < n; ++i) {
        Request* r = batch[a + i];
        std::memcpy(w.hQ.p + i * r->q.size(), r->q.data(), r->q.size() * sizeof(double));
        if (w.type == 1) {
            w.hK.p[i] = r->K;
            Kmax = std::max(Kmax, r->K);
        }
    }
    const size_t D = s->D;
    const size_t qw = w.type == 0 ? 2 * D : D;
    hipCheck(hipMemcpyAsync(w.dQ.data(), w.hQ.p, n * qw * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
    int rc;
    if (w.type == 0) {
        rc = tblaeq_batch_range(w.b, w.dQ.data(), n, f->dOff.data(), f->dIds.data(), f->dIds.size(),
                                w.dStatus.data(), nullptr, nullptr);
    } else {
        if (Kmax == 0) Kmax = 1;
        hipCheck(hipMemcpyAsync(w.dK.data(), w.hK.p, n * 8, hipMemcpyHostToDevice, st), "H2D");
        rc = tblaeq_batch_knn(w.b, w.dQ.data(), w.dK.data(), n, Kmax, w.dIds.data(), w.dD2.data(), w.dN.data(),
                              w.dStatus.data(), nullptr, nullptr);
    }
    tbCheck(rc, w.type == 0 ? "serve: range batch" : "serve: KNN batch");
    hipCheck(hipMemcpyAsync(w.hStatus.p, w.dStatus.data(), 32, hipMemcpyDeviceToHost, st), "D2H");
    return Kmax;
}

// Queues the copies of flight f's range results (n requests, f->hits ids) on
// the worker's stream, after the engine batch that wrote them.
void send_range(Worker& w, Flight& f, size_t n) {
    hipStream_t st = static_cast<hipStream_t>(tblaeq_ctx_stream(w.ctx));
    if (w.narrow && f.hits) hipCheck(tblaeq::narrow_ids(st, f.dIds.data(), f.dIds32.data(), f.hits), "narrow ids");
    hipCheck(hipMemcpyAsync(f.hOff.p, f.dOff.data(), (n + 1) * 8, hipMemcpyDeviceToHost, st), "D2H");
    if (w.narrow)
        hipCheck(hipMemcpyAsync(f.hIds32.p, f.dIds32.data(), f.hits * 4, hipMemcpyDeviceToHost, st), "D2H");
    else
        hipCheck(hipMemcpyAsync(f.hIds.p, f.dIds.data(), f.hits * 8, hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipEventRecord(f.copied, st), "event");
}

// Waits for flight f's copies and fills requests [a, a+n) of `batch`.
void receive_range(Worker& w, Flight& f, std::vector<Request*>& batch, size_t a, size_t n) {
    hipCheck(hipEventSynchronize(f.copied), "event sync");
    if (w.narrow) scatter_results(f.hIds32.p, f.hOff.p, n, batch, a);
    else scatter_results(f.hIds.p, f.hOff.p, n, batch, a);
    // running mean of hits per request sizes the next batches
    const double per = static_cast<double>(f.hits) / static_cast<double>(n);
    w.hitsPerRequest = w.hitsPerRequest > 0.0 ? 0.8 * w.hitsPerRequest + 0.2 * per : per;
}

// After a batch overflowed the engine's capacities (status in w.hStatus):
// reruns it in parts, synchronously.
void run_batch(tblaeq_server* s, Worker& w, Flight* f, std::vector<Request*>& batch, size_t a, size_t n);
void split_batch(tblaeq_server* s, Worker& w, Flight* f, std::vector<Request*>& batch, size_t a, size_t n) {
    if (n == 1) {
        batch[a]->status = TBLAEQ_ERR_CAPACITY;
        return;
    }
    // only the hit list overflowed: status[0] still counts every hit, so
    // split once into parts that each fit with a margin; otherwise halve
    ++w.reruns;
    uint64_t parts = 2;
    if (w.type == 0 && w.hStatus.p[1] == 4) {
        const double need = static_cast<double>(w.hStatus.p[0]) * 1.25 / static_cast<double>(s->cfg.hit_cap);
        parts = std::min<uint64_t>(n, std::max<uint64_t>(2, static_cast<uint64_t>(std::ceil(need))));
    }
    // the next batches start at the size that fits, instead of overflowing again
    w.sizeLimit = std::max<size_t>(1, n / parts);
    for (uint64_t k = 0; k < parts; ++k) {
        const size_t b0 = a + n * k / parts, b1 = a + n * (k + 1) / parts;
        if (b1 > b0) run_batch(s, w, f, batch, b0, b1 - b0);
    }
}

// Runs requests [a, a+n) of `batch` on the worker's engine to completion,
// splitting on overflow (range results through flight f).
void run_batch(tblaeq_server* s, Worker& w, Flight* f, std::vector<Request*>& batch, size_t a, size_t n) {
    hipStream_t st = static_cast<hipStream_t>(tblaeq_ctx_stream(w.ctx));
    const uint64_t Kmax = launch_batch(s, w, f, batch, a, n);
    hipCheck(hipStreamSynchronize(st), "sync");
    if (w.hStatus.p[1] != 0) {
        split_batch(s, w, f, batch, a, n);
        return;
    }
    if (w.type == 0) {
        f->hits = w.hStatus.p[0];
        send_range(w, *f, n);
        receive_range(w, *f, batch, a, n);
    } else {
        hipCheck(hipMemcpyAsync(w.hN.p, w.dN.data(), n * 8, hipMemcpyDeviceToHost, st), "D2H");
        hipCheck(hipMemcpyAsync(w.hIds.p, w.dIds.data(), n * Kmax * 8, hipMemcpyDeviceToHost, st), "D2H");
        hipCheck(hipMemcpyAsync(w.hD2.p, w.dD2.data(), n * Kmax * 8, hipMemcpyDeviceToHost, st), "D2H");
        hipCheck(hipStreamSynchronize(st), "sync");
        for (size_t i = 0; i < n; ++i) {
            const uint64_t k = w.hN.p[i];
            batch[a + i]->ids.assign(w.hIds.p + i * Kmax, w.hIds.p + i * Kmax + k);
            batch[a + i]->d2.assign(w.hD2.p + i * Kmax, w.hD2.p + i * Kmax + k);
        }
    }
}

// One request through the per-query traversal (performSingleRangeQuery /
// performSingleKNNQuery with the fine mesh kept, QueryHandler.cu:220-247), as
// the reference's front end runs it: the fine mesh's ids in traversal order,
// per-level output sizes and the query's time.
void run_traversal(tblaeq_server* s, Request* r) {
    std::lock_guard<std::mutex> lock(s->h->plainMu);
    const size_t D = s->D;
    std::vector<double> lo, hi;
    if (r->type == 0) {
        lo.assign(r->q.begin(), r->q.begin() + static_cast<long>(D));
        hi.assign(r->q.begin() + static_cast<long>(D), r->q.end());
    }
    const QueryResult qr = r->type == 0 ? s->h->h->performSingleRangeQuery(hi, lo, true)
                                        : s->h->h->performSingleKNNQuery(r->q, r->K, true);
    if (qr.errorCode || qr.fineMesh.size() != 1) throw std::runtime_error("the traversal returned no fine mesh");
    const SparseGrid* g = qr.fineMesh[0];
    r->ids.assign(g->get_ids_(), g->get_ids_() + g->get_nnz_nums());
    r->levelSizes.assign(qr.levelFineMeshSize[0].begin(), qr.levelFineMeshSize[0].end());
    r->levelOriginal.assign(qr.levelOriginalSize.begin(), qr.levelOriginalSize.end());
    r->queryUs = static_cast<uint64_t>(qr.queryTimeUs[0]);
    r->totalUs = static_cast<uint64_t>(qr.totalTimeUs);
    r->queryCount = qr.queryCount;
    r->policy = qr.memoryPolicy;
}

// Device buffers, events and streams of a worker, then its context (the
// engine has been destroyed).
void release_worker(Worker& w) {
    w.dQ = DeviceBuffer<double>();
    w.dD2 = DeviceBuffer<double>();
    w.dK = DeviceBuffer<uint64_t>();
    w.dIds = DeviceBuffer<uint64_t>();
    w.dN = DeviceBuffer<uint64_t>();
    w.dStatus = DeviceBuffer<uint64_t>();
    for (Flight& f : w.fl) {
        f.dOff = DeviceBuffer<uint64_t>();
        f.dIds = DeviceBuffer<uint64_t>();
        f.dIds32 = DeviceBuffer<uint32_t>();
        if (f.copied) (void)hipEventDestroy(f.copied);
        f.copied = nullptr;
    }
    if (w.ctx) tblaeq_ctx_destroy(w.ctx);
    w.ctx = nullptr;
}

// Marks a batch's requests done (rc for those without their own status) and
// updates the statistics and the in-flight budget.
void complete_batch(tblaeq_server* s, Worker& w, std::vector<Request*>& batch, uint64_t est, uint64_t travs, int rc,
                    const std::string& rcMsg) {
    const auto now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request* r : batch) {
            if (rc && r->status == TBLAEQ_OK) {
                r->status = rc;
                r->error = rcMsg;
            }
            r->finished = now;
            r->done = true;
            s->latUs.push_back(std::chrono::duration<double, std::micro>(now - r->submitted).count());
        }
        for (Request* r : batch)
            if (r->orphan) s->reqs.erase(r->ticket);
        s->inflight -= est;
        s->batches += 1;
        s->batchedReqs += batch.size();
        w.completed += batch.size();
        w.batches += 1;
        w.traversals += travs;
    }
    s->cvDone.notify_all();
    s->cvWork.notify_all();
}

// Completes the range batch in flight f: its ids into its requests.
void finish_flight(tblaeq_server* s, Worker& w, Flight& f) {
    int rc = TBLAEQ_OK;
    std::string rcMsg;
    try {
        receive_range(w, f, f.batched, 0, f.batched.size());
    } catch (...) {
        rc = tblaeq_host::statusOfCurrentException("tblaeq_server batch");
        rcMsg = tblaeq_last_error();
    }
    complete_batch(s, w, f.batch, f.est, f.travs, rc, rcMsg);
    f.batch.clear();
    f.batched.clear();
}

void worker_loop(tblaeq_server* s, Worker& w) {
    if (hipSetDevice(w.ctx->device) != hipSuccess) return;
    const double maxWait = s->cfg.max_wait_us;
    hipStream_t st = static_cast<hipStream_t>(tblaeq_ctx_stream(w.ctx));
    // range: the flight whose results are still on their way (its requests
    // complete while the next batch runs)
    Flight* pending = nullptr;
    auto addBusy = [&](Clock::time_point t0) {
        const double us = std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(s->mu);
        w.busyUs += us;
    };
    for (;;) {
        std::vector<Request*> batch;
        uint64_t est = 0;
        bool drain = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (pending && (s->queue[w.type].empty() || maxWait > 0 || s->stop)) {
                drain = true;  // nothing to overlap with: complete the flight first
            } else {
                s->cvWork.wait(lk, [&] { return s->stop || !s->queue[w.type].empty(); });
                if (s->queue[w.type].empty()) {
                    if (s->stop) return;
                    continue;
                }
                if (maxWait > 0 && s->queue[w.type].size() < s->cfg.max_batch && !s->stop) {
                    s->cvWork.wait_for(lk, std::chrono::duration<double, std::micro>(maxWait),
                                       [&] { return s->stop || s->queue[w.type].size() >= s->cfg.max_batch; });
                }
                auto& q = s->queue[w.type];
                // range batches are sized so that their expected hits fit the
                // hit list (a batch that overflows is split and rerun)
                size_t limit = s->cfg.max_batch;
                if (w.type == 0 && w.hitsPerRequest > 0.0)
                    limit = static_cast<size_t>(std::clamp(
                        0.7 * static_cast<double>(s->cfg.hit_cap) / w.hitsPerRequest, 1.0,
                        static_cast<double>(s->cfg.max_batch)));
                if (w.sizeLimit) limit = std::min(limit, w.sizeLimit);
                while (!q.empty() && batch.size() < limit) {
                    Request* r = q.front();
                    if (!batch.empty() || s->inflight > 0) {
                        if (s->inflight + est + r->est > s->budget) {
                            ++s->deferred;
                            break;  // wait for memory to free up
                        }
                    }
                    q.pop_front();
                    est += r->est;
                    batch.push_back(r);
                }
                if (batch.empty()) {
                    if (pending) {
                        drain = true;  // the budget is held by this worker's own flight
                    } else {
                        // the budget is held by the other worker's batch: wait for it to complete
                        s->cvDone.wait(lk, [&] { return s->inflight == 0; });
                        continue;
                    }
                }
                s->inflight += est;
                s->peakInflight = std::max(s->peakInflight, s->inflight);
            }
        }
        if (drain) {
            const auto t0 = Clock::now();
            finish_flight(s, w, *pending);
            pending = nullptr;
            addBusy(t0);
            continue;
        }
        int rc = TBLAEQ_OK;
        std::string rcMsg;
        const auto busy0 = Clock::now();
        uint64_t travs = 0;
        bool inFlight = false;
        try {
            // traversal requests, and KNN above the batched kernels' K limit,
            // run the per-query traversal; the rest form the engine's batch
            std::vector<Request*> batched;
            for (Request* r : batch) {
                if (!r->traversal && !(w.type == 1 && r->K > s->maxK)) {
                    batched.push_back(r);
                    continue;
                }
                ++travs;
                try {
                    if (r->traversal) run_traversal(s, r);
                    else plainKnn(s->h, r->q.data(), r->K, r->ids, r->d2, nullptr, nullptr);
                } catch (...) {
                    r->status = tblaeq_host::statusOfCurrentException("tblaeq_server request");
                    r->error = tblaeq_last_error();
                }
            }
            if (!batched.empty()) {
                const uint64_t reruns0 = w.reruns;
                if (w.type == 0) {
                    // launch, complete the previous flight while the GPU runs,
                    // then send this batch's results on their way
                    Flight& f = w.fl[w.cur];
                    launch_batch(s, w, &f, batched, 0, batched.size());
                    if (pending) {
                        finish_flight(s, w, *pending);
                        pending = nullptr;
                    }
                    hipCheck(hipStreamSynchronize(st), "sync");
                    if (w.hStatus.p[1] != 0) {
                        split_batch(s, w, &f, batched, 0, batched.size());
                    } else {
                        f.hits = w.hStatus.p[0];
                        send_range(w, f, batched.size());
                        f.batch = batch;
                        f.batched = std::move(batched);
                        f.est = est;
                        f.travs = travs;
                        pending = &f;
                        w.cur ^= 1;
                        inFlight = true;
                    }
                } else {
                    run_batch(s, w, nullptr, batched, 0, batched.size());
                }
                if (w.sizeLimit && w.reruns == reruns0 && batch.size() >= w.sizeLimit) {
                    w.sizeLimit += std::max<size_t>(1, w.sizeLimit / 4);
                    if (w.sizeLimit >= s->cfg.max_batch) w.sizeLimit = 0;
                }
            }
        } catch (...) {
            rc = tblaeq_host::statusOfCurrentException("tblaeq_server batch");
            rcMsg = tblaeq_last_error();
            if (pending) {
                finish_flight(s, w, *pending);
                pending = nullptr;
            }
        }
        if (!inFlight) complete_batch(s, w, batch, est, travs, rc, rcMsg);
        addBusy(busy0);
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p / 100.0 * static_cast<double>(v.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double f = pos - static_cast<double>(i);
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

// Queues a request (admission estimate by kind) and wakes the workers.
int enqueue(tblaeq_server* s, std::unique_ptr<Request> r, uint64_t* ticket) {
    r->est = r->type == 0 ? s->cfg.range_request_bytes : s->cfg.knn_request_bytes;
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            tblaeq::set_error("tblaeq_server_submit: server is stopping");
            return TBLAEQ_ERR_INVALID;
        }
        r->ticket = s->nextTicket++;
        *ticket = r->ticket;
        s->queue[r->type].push_back(r.get());
        s->reqs.emplace(r->ticket, std::move(r));
    }
    s->cvWork.notify_all();
    return TBLAEQ_OK;
}

// ---------------------------------------------------------------- JSON front end

struct Reply {
    int http = 200;
    std::string body;
};

Reply failure(int http, const std::string& msg) {
    tblaeq_json::Writer w;
    w.beginObject();
    w.key("success");
    w.boolean(false);
    w.key("message");
    w.str(msg);
    w.endObject();
    return {http, w.out};
}

// Dataset names the reference's front end accepts (HTTPServer.cu:94-97): one
// or more of [A-Za-z0-9._-].
bool validDatasetName(const std::string& n) {
    if (n.empty()) return false;
    for (const char c : n)
        if (!(std::isalnum(static_cast<unsigned char>(c)) || c == '.' || c == '_' || c == '-')) return false;
    return true;
}

bool numericArray(const tblaeq_json::Value* v, std::vector<double>& out) {
    if (!v || !v->isArray()) return false;
    out.clear();
    for (const auto& e : v->arr) {
        if (!e.isNumber()) return false;
        out.push_back(e.num);
    }
    return true;
}

// The POST /query contract of the reference's front end (HTTPServer.cu:414-553):
// the same request fields and checks, the same response fields.  Requests run
// through this server's queue; by default as the per-query traversal the
// reference runs (result = fine-mesh ids in traversal order), with "exact":
// true on the batched engine (result = the exact answer).
Reply handle_query(tblaeq_server* s, const std::string& text, double timeout_s) {
    using tblaeq_json::Value;
    Value body;
    try {
        body = tblaeq_json::Reader(text).parse();
    } catch (const std::exception& e) {
        return failure(400, std::string("request is not valid JSON: ") + e.what());
    }
    if (!body.isObject()) return failure(400, "the request body must be a JSON object");
    const Value* nameV = body.find("dataset_name");
    if (!nameV || !nameV->isString()) return failure(400, "'dataset_name' is required and must be a string");
    const std::string name = nameV->str;
    if (!validDatasetName(name)) return failure(400, "'dataset_name' may hold only letters, digits, '.', '_' and '-'");
    // an index's dataset name is its directory name up to the first '_'
    // (Utils.cu:6-10); an index with no name answers to any dataset_name
    if (!s->datasetName.empty() && name.substr(0, name.find('_')) != s->datasetName)
        return failure(400, "dataset '" + name + "' is not served here (this server holds '" + s->datasetName + "')");
    const Value* typeV = body.find("query_type");
    if (!typeV || !typeV->isString()) return failure(400, "'query_type' is required and must be 'KNN' or 'RANGE'");
    std::string qtype = typeV->str;
    for (char& c : qtype) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    if (qtype != "KNN" && qtype != "RANGE") return failure(400, "'query_type' must be 'KNN' or 'RANGE'");
    bool exact = false;
    if (const Value* ex = body.find("exact")) {
        if (!ex->isBool()) return failure(400, "'exact' must be a boolean");
        exact = ex->b;
    }
    const size_t D = s->D;
    auto r = std::make_unique<Request>();
    r->traversal = !exact;
    if (qtype == "KNN") {
        r->type = 1;
        const Value* pt = body.find("query_point");
        if (!pt) return failure(400, "'query_point' is required for a KNN query");
        if (!numericArray(pt, r->q)) return failure(400, "'query_point' must be an array of numbers");
        if (r->q.size() != D) return failure(400, "'query_point' must have " + std::to_string(D) + " numbers");
        r->K = 10;
        if (const Value* kv = body.find("k")) {
            if (!kv->isNumber() || !(kv->num > 0) || std::floor(kv->num) != kv->num || kv->num > 9007199254740992.0)
                return failure(400, "'k' must be a positive integer");
            r->K = static_cast<uint64_t>(kv->num);
        }
    } else {
        r->type = 0;
        std::vector<double> lo, hi;
        const Value* up = body.find("upper_bound");
        const Value* low = body.find("lower_bound");
        if (!up || !low) return failure(400, "'upper_bound' and 'lower_bound' are required for a RANGE query");
        if (!numericArray(up, hi) || !numericArray(low, lo))
            return failure(400, "'upper_bound' and 'lower_bound' must be arrays of numbers");
        if (hi.size() != D || lo.size() != D)
            return failure(400, "'upper_bound' and 'lower_bound' must have " + std::to_string(D) + " numbers each");
        for (size_t d = 0; d < D; ++d)
            if (lo[d] > hi[d])
                return failure(400, "empty range: lower_bound[" + std::to_string(d) + "] > upper_bound[" +
                                        std::to_string(d) + "]");
        r->q = lo;
        r->q.insert(r->q.end(), hi.begin(), hi.end());
    }
    uint64_t ticket = 0;
    if (enqueue(s, std::move(r), &ticket) != TBLAEQ_OK) return failure(503, "the server is stopping");
    std::unique_lock<std::mutex> lk(s->mu);
    Request* req = s->reqs.at(ticket).get();
    if (timeout_s < 0) {
        s->cvDone.wait(lk, [&] { return req->done; });
    } else if (!s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return req->done; })) {
        req->orphan = true;
        return failure(504, "the query did not finish within the timeout");
    }
    std::unique_ptr<Request> done = std::move(s->reqs.at(ticket));
    s->reqs.erase(ticket);
    lk.unlock();
    if (done->status != TBLAEQ_OK)
        return failure(500, "the query failed: " + (done->error.empty() ? std::string(tblaeq_status_string(done->status))
                                                                         : done->error));
    const double latencyUs =
        std::floor(std::chrono::duration<double, std::micro>(done->finished - done->submitted).count());
    tblaeq_json::Writer w;
    w.beginObject();
    w.key("success");
    w.boolean(true);
    w.key("dataset_name");
    w.str(name);
    w.key("query_type");
    w.str(qtype);
    w.key("result_count");
    w.integer(done->ids.size());
    w.key("result");
    w.integers(done->ids);
    if (exact && done->type == 1) {
        w.key("squared_distances");
        w.beginArray();
        for (const double v : done->d2) w.number(v);
        w.endArray();
    }
    w.key("memory_policy");
    w.str(exact ? s->policy : done->policy);
    w.key("level_original_sizes");
    w.integers(exact ? s->levelOriginal : done->levelOriginal);
    if (!exact) {
        w.key("level_pruned_sizes");
        w.integers(done->levelSizes);
    }
    // traversal: the reference's timings of the one query; exact: the
    // request's latency through the server (queue, batch, copies)
    const double qUs = exact ? latencyUs : static_cast<double>(done->queryUs);
    const double tUs = exact ? latencyUs : static_cast<double>(done->totalUs);
    const double cnt = exact ? 1.0 : static_cast<double>(done->queryCount);
    w.key("query_runtime");
    w.beginObject();
    w.key("query_time_us");
    w.number(qUs);
    w.key("query_time_ms");
    w.number(qUs / 1000.0);
    w.key("total_time_us");
    w.number(tUs);
    w.key("avg_time_us");
    w.number(cnt > 0 ? tUs / cnt : 0.0);
    w.key("query_count");
    w.number(cnt);
    w.endObject();
    if (exact) {
        w.key("exact");
        w.boolean(true);
    }
    w.endObject();
    return {200, w.out};
}

}  // namespace

extern "C" {

void tblaeq_server_config_init(tblaeq_server_config* c) {
    if (!c) return;
    c->max_batch = 4096;
    c->max_k = 0;
    c->pair_cap = 0;
    c->hit_cap = 0;
    c->hbm_fraction = 0.5;
    c->range_request_bytes = 4ull << 20;
    c->knn_request_bytes = 64ull << 10;
    c->max_wait_us = 0.0;
}

int tblaeq_server_create(tblaeq_handler* h, const tblaeq_server_config* cfgIn, tblaeq_server** out) {
    if (!h || !out) {
        tblaeq::set_error("tblaeq_server_create: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    tblaeq_server_config cfg;
    tblaeq_server_config_init(&cfg);
    if (cfgIn) cfg = *cfgIn;
    if (cfg.max_batch == 0 || cfg.max_batch >= (1ull << 24) || !(cfg.hbm_fraction > 0.0) || cfg.hbm_fraction > 1.0) {
        tblaeq::set_error("tblaeq_server_create: max_batch must be in 1..2^24-1 and hbm_fraction in (0, 1]");
        return TBLAEQ_ERR_INVALID;
    }
    auto s = std::make_unique<tblaeq_server>();
    s->h = h;
    try {
        IndexData& idx = h->h->index();
        if (!h->h->prepared()) h->h->prepareForQuery(LevelPolicy::L0);
        for (size_t l = 0; l < idx.intervals; ++l) idx.uploadLevelValues(l);
        s->D = idx.D;
        s->datasetName = idx.datasetName;
        for (size_t l = 0; l < idx.intervals; ++l) {
            if (l) s->policy += "-";
            s->policy += levelPolicyName(idx.activePolicy[l]);
            s->levelOriginal.push_back(idx.meshSizes[idx.intervals - (l + 1)]);
        }
        std::vector<tblaeq_level> lv(idx.intervals);
        std::vector<double> maxR(idx.intervals, 0.0);
        uint64_t maxEmpty = 0;
        for (size_t l = 0; l < idx.intervals; ++l) {
            lv[l] = idx.levelDesc(l);
            maxEmpty = std::max<uint64_t>(maxEmpty, lv[l].empty_cols);
            for (double r : idx.meshMaxRadius[l]) maxR[l] = std::max(maxR[l], r);
        }
        const uint64_t kLimit = 255 > maxEmpty ? 255 - maxEmpty : 0;
        s->maxK = cfg.max_k ? std::min<uint64_t>(cfg.max_k, kLimit) : kLimit;
        size_t freeB = 0, totalB = 0;
        hipCheck(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        // engine scratch from free HBM when not given: pairs 1/64 of it per
        // engine (16 bytes an entry), at most 256M entries (a KNN batch over a
        // large index keeps tens of thousands of (cluster, query) pairs per
        // query at the level above the last); hits: 1/128 of it, at most 128M
        // (the range worker holds four hit-sized buffers)
        if (!cfg.pair_cap) cfg.pair_cap = std::min<uint64_t>(std::max<uint64_t>(freeB / 64 / 16, 1ull << 20), 1ull << 28);
        if (!cfg.hit_cap) cfg.hit_cap = std::min<uint64_t>(std::max<uint64_t>(freeB / 128 / 8, 1ull << 20), 1ull << 27);
        const int device = idx.ctx->device;
        for (int k = 0; k < 2; ++k) {
            Worker& w = s->w[k];
            w.type = k;
            tbCheck(tblaeq_ctx_create(device, nullptr, &w.ctx), "serve: context");
            tbCheck(tblaeq_batch_create(w.ctx, lv.data(), lv.size(), idx.dCoarsestIds.data(), idx.dCoarsestVals.data(),
                                        idx.coarsestIds.size(), maxR.data(), &w.b),
                    "serve: engine");
            // the KNN engine's candidate records use min(pair_cap, hit_cap) entries
            tbCheck(tblaeq_batch_reserve(w.b, cfg.max_batch, cfg.pair_cap,
                                         k == 0 ? cfg.hit_cap : std::min(cfg.pair_cap, cfg.hit_cap)),
                    "serve: reserve");
            const size_t qw = k == 0 ? 2 * s->D : s->D;
            w.dQ.resize(cfg.max_batch * qw);
            w.hQ.alloc(cfg.max_batch * qw);
            w.dStatus.resize(4);
            w.hStatus.alloc(4);
            if (k == 0) {
                // two result flights; ids cross to the host as 32 bits when the
                // index's ids fit
                w.narrow = idx.N <= (1ull << 32);
                for (Flight& f : w.fl) {
                    f.dOff.resize(cfg.max_batch + 1);
                    f.dIds.resize(cfg.hit_cap);
                    f.hOff.alloc(cfg.max_batch + 1);
                    if (w.narrow) {
                        f.dIds32.resize(cfg.hit_cap);
                        f.hIds32.alloc(cfg.hit_cap);
                    } else {
                        f.hIds.alloc(cfg.hit_cap);
                    }
                    hipCheck(hipEventCreateWithFlags(&f.copied, hipEventDisableTiming), "event");
                }
            } else {
                const size_t rows = cfg.max_batch * std::max<uint64_t>(s->maxK, 1);
                w.dK.resize(cfg.max_batch);
                w.dN.resize(cfg.max_batch);
                w.dIds.resize(rows);
                w.dD2.resize(rows);
                w.hK.alloc(cfg.max_batch);
                w.hN.alloc(cfg.max_batch);
                w.hIds.alloc(rows);
                w.hD2.alloc(rows);
            }
        }
        // launch parameters: those of the handler's batched engine when it has
        // one (tuned with tblaeq_handler_batch_autotune, or set by hand)
        if (h->slots[0] && h->slots[0]->b) {
            static const int kParamIds[] = {TBLAEQ_PARAM_KNN_SEED,           TBLAEQ_PARAM_KNN_WG_PER_CU,
                                            TBLAEQ_PARAM_KNN_MODE,           TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU,
                                            TBLAEQ_PARAM_STEP_BLOCKS_PER_CU, TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU,
                                            TBLAEQ_PARAM_HIT_STAGE,          TBLAEQ_PARAM_RANGE_ROW_QUERIES,
                                            TBLAEQ_PARAM_KNN_ROW_QUERIES,    TBLAEQ_PARAM_KNN_LEAF_REFINE};
            for (int id : kParamIds) {
                uint64_t v = 0;
                tbCheck(tblaeq_batch_get_param(h->slots[0]->b, id, &v), "serve: launch parameter");
                for (int k = 0; k < 2; ++k) tbCheck(tblaeq_batch_set_param(s->w[k].b, id, v), "serve: launch parameter");
            }
        }
        hipCheck(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        s->budget = static_cast<uint64_t>(static_cast<double>(freeB) * cfg.hbm_fraction);
        s->cfg = cfg;
        tblaeq_server* raw = s.get();
        for (int k = 0; k < 2; ++k) raw->w[k].th = std::thread(worker_loop, raw, std::ref(raw->w[k]));
    } catch (...) {
        const int rc = tblaeq_host::statusOfCurrentException("tblaeq_server_create");
        for (auto& w : s->w) {
            if (w.b) tblaeq_batch_destroy(w.b);
            release_worker(w);
        }
        return rc;
    }
    *out = s.release();
    return TBLAEQ_OK;
}

int tblaeq_server_destroy(tblaeq_server* s) {
    if (!s) return TBLAEQ_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cvWork.notify_all();
    s->cvDone.notify_all();
    for (auto& w : s->w)
        if (w.th.joinable()) w.th.join();
    for (auto& w : s->w) {
        if (w.b) tblaeq_batch_destroy(w.b);
        w.b = nullptr;
    }
    for (auto& w : s->w) release_worker(w);
    delete s;
    return TBLAEQ_OK;
}

int tblaeq_server_submit(tblaeq_server* s, int type, const double* q, uint64_t K, uint64_t* ticket) {
    if (!s || !q || !ticket || (type != 0 && type != 1)) {
        tblaeq::set_error("tblaeq_server_submit: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    auto r = std::make_unique<Request>();
    r->type = type;
    r->q.assign(q, q + (type == 0 ? 2 * s->D : s->D));
    r->K = K;
    return enqueue(s, std::move(r), ticket);
}

int tblaeq_server_wait(tblaeq_server* s, uint64_t ticket, double timeout_s) {
    if (!s) return TBLAEQ_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        tblaeq::set_error("tblaeq_server_wait: unknown ticket");
        return TBLAEQ_ERR_INVALID;
    }
    Request* r = it->second.get();
    if (timeout_s < 0) {
        s->cvDone.wait(lk, [&] { return r->done; });
    } else if (!s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; })) {
        return TBLAEQ_ERR_TIMEOUT;
    }
    return r->status;
}

int tblaeq_server_result(tblaeq_server* s, uint64_t ticket, uint64_t* ids, uint64_t cap, double* d2, uint64_t* n,
                         int release) {
    if (!s || !n) return TBLAEQ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || !it->second->done) {
        tblaeq::set_error("tblaeq_server_result: unknown or unfinished ticket");
        return TBLAEQ_ERR_INVALID;
    }
    Request* r = it->second.get();
    *n = r->ids.size();
    const uint64_t c = std::min<uint64_t>(cap, r->ids.size());
    if (ids && c) std::memcpy(ids, r->ids.data(), c * 8);
    if (d2 && c && !r->d2.empty()) std::memcpy(d2, r->d2.data(), c * 8);
    const int st = r->status;
    if (release) s->reqs.erase(it);
    return st;
}

int tblaeq_server_stats(tblaeq_server* s, double* out) {
    if (!s || !out) return TBLAEQ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = static_cast<double>(s->latUs.size());
    out[1] = percentile(s->latUs, 50);
    out[2] = percentile(s->latUs, 99);
    out[3] = static_cast<double>(s->batches);
    out[4] = s->batches ? static_cast<double>(s->batchedReqs) / static_cast<double>(s->batches) : 0.0;
    out[5] = static_cast<double>(s->deferred);
    out[6] = static_cast<double>(s->peakInflight);
    out[7] = static_cast<double>(s->budget);
    return TBLAEQ_OK;
}

int tblaeq_server_query_json(tblaeq_server* s, const char* request, double timeout_s, char** response,
                             int* http_status) {
    if (!s || !request || !response || !http_status) {
        tblaeq::set_error("tblaeq_server_query_json: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    *response = nullptr;
    try {
        const Reply rep = handle_query(s, request, timeout_s);
        char* out = static_cast<char*>(std::malloc(rep.body.size() + 1));
        if (!out) throw std::bad_alloc();
        std::memcpy(out, rep.body.c_str(), rep.body.size() + 1);
        *response = out;
        *http_status = rep.http;
        return TBLAEQ_OK;
    } catch (...) {
        return tblaeq_host::statusOfCurrentException("tblaeq_server_query_json");
    }
}

void tblaeq_free_string(char* p) { std::free(p); }

int tblaeq_server_kind_stats(tblaeq_server* s, int type, double* out) {
    if (!s || !out || (type != 0 && type != 1)) {
        tblaeq::set_error("tblaeq_server_kind_stats: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const Worker& w = s->w[type];
    out[0] = static_cast<double>(w.completed);
    out[1] = static_cast<double>(w.batches);
    out[2] = static_cast<double>(w.reruns);
    out[3] = static_cast<double>(w.traversals);
    out[4] = w.busyUs;
    out[5] = w.hitsPerRequest;
    return TBLAEQ_OK;
}

int tblaeq_server_reset_stats(tblaeq_server* s) {
    if (!s) return TBLAEQ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->latUs.clear();
    s->batches = s->batchedReqs = s->deferred = 0;
    for (auto& w : s->w) {
        w.completed = w.batches = w.reruns = w.traversals = 0;
        w.busyUs = 0.0;
    }
    s->peakInflight = s->inflight;
    return TBLAEQ_OK;
}

}  // extern "C"
