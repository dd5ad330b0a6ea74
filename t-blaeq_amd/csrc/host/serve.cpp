// Serving loop: continuous batching of range and KNN requests (tblaeq_serve.h).
//
// One worker thread per request kind owns a batched engine with its own HIP
// stream, device query / result buffers and pinned staging.  Requests wait in
// a FIFO per kind; a worker admits the waiting requests into its next batch as
// soon as its previous batch has completed, bounded by max_batch and by the
// in-flight HBM budget shared by both workers.  A batch whose scratch would
// overflow the engine's capacities is split in halves and rerun, exactly as
// the handler's batched driver does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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

struct Worker {
    int type = 0;
    tblaeq_ctx* ctx = nullptr;
    tblaeq_batch* b = nullptr;
    DeviceBuffer<double> dQ, dD2;
    DeviceBuffer<uint64_t> dK, dOff, dIds, dN, dStatus;
    Pinned<double> hQ, hD2;
    Pinned<uint64_t> hK, hOff, hIds, hN;
    std::thread th;
};

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
    // statistics
    std::vector<double> latUs;
    uint64_t batches = 0, batchedReqs = 0, deferred = 0;
};

namespace {

// Runs requests [a, a+n) of `batch` on the worker's engine, splitting on overflow.
void run_batch(tblaeq_server* s, Worker& w, std::vector<Request*>& batch, size_t a, size_t n) {
    const size_t D = s->D;
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
    const size_t qw = w.type == 0 ? 2 * D : D;
    hipCheck(hipMemcpyAsync(w.dQ.data(), w.hQ.p, n * qw * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
    int rc;
    if (w.type == 0) {
        rc = tblaeq_batch_range(w.b, w.dQ.data(), n, w.dOff.data(), w.dIds.data(), w.dIds.size(), w.dStatus.data(),
                                nullptr, nullptr);
    } else {
        if (Kmax == 0) Kmax = 1;
        hipCheck(hipMemcpyAsync(w.dK.data(), w.hK.p, n * 8, hipMemcpyHostToDevice, st), "H2D");
        rc = tblaeq_batch_knn(w.b, w.dQ.data(), w.dK.data(), n, Kmax, w.dIds.data(), w.dD2.data(), w.dN.data(),
                              w.dStatus.data(), nullptr, nullptr);
    }
    tbCheck(rc, w.type == 0 ? "serve: range batch" : "serve: KNN batch");
    uint64_t status[4];
    hipCheck(hipMemcpyAsync(status, w.dStatus.data(), 32, hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipStreamSynchronize(st), "sync");
    if (status[1] != 0) {
        if (n == 1) {
            batch[a]->status = TBLAEQ_ERR_CAPACITY;
            return;
        }
        run_batch(s, w, batch, a, n / 2);
        run_batch(s, w, batch, a + n / 2, n - n / 2);
        return;
    }
    if (w.type == 0) {
        const uint64_t hits = status[0];
        hipCheck(hipMemcpyAsync(w.hOff.p, w.dOff.data(), (n + 1) * 8, hipMemcpyDeviceToHost, st), "D2H");
        hipCheck(hipMemcpyAsync(w.hIds.p, w.dIds.data(), hits * 8, hipMemcpyDeviceToHost, st), "D2H");
        hipCheck(hipStreamSynchronize(st), "sync");
        for (size_t i = 0; i < n; ++i)
            batch[a + i]->ids.assign(w.hIds.p + w.hOff.p[i], w.hIds.p + w.hOff.p[i + 1]);
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

void worker_loop(tblaeq_server* s, Worker& w) {
    if (hipSetDevice(w.ctx->device) != hipSuccess) return;
    const double maxWait = s->cfg.max_wait_us;
    for (;;) {
        std::vector<Request*> batch;
        uint64_t est = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
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
            while (!q.empty() && batch.size() < s->cfg.max_batch) {
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
                // the budget is held by the other worker's batch: wait for it to complete
                s->cvDone.wait(lk, [&] { return s->inflight == 0; });
                continue;
            }
            s->inflight += est;
            s->peakInflight = std::max(s->peakInflight, s->inflight);
        }
        int rc = TBLAEQ_OK;
        try {
            if (w.type == 1) {
                // K above the batched kernels' limit: the per-query traversal
                std::vector<Request*> small;
                for (Request* r : batch) {
                    if (r->K <= s->maxK) {
                        small.push_back(r);
                        continue;
                    }
                    try {
                        plainKnn(s->h, r->q.data(), r->K, r->ids, r->d2, nullptr, nullptr);
                    } catch (...) {
                        r->status = tblaeq_host::statusOfCurrentException("tblaeq_server large-K request");
                    }
                }
                if (!small.empty()) run_batch(s, w, small, 0, small.size());
            } else {
                run_batch(s, w, batch, 0, batch.size());
            }
        } catch (...) {
            rc = tblaeq_host::statusOfCurrentException("tblaeq_server batch");
        }
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request* r : batch) {
                if (rc) r->status = rc;
                r->finished = now;
                r->done = true;
                s->latUs.push_back(std::chrono::duration<double, std::micro>(now - r->submitted).count());
            }
            s->inflight -= est;
            s->batches += 1;
            s->batchedReqs += batch.size();
        }
        s->cvDone.notify_all();
        s->cvWork.notify_all();
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
        // engine scratch from free HBM when not given: 1/256 of it per buffer, at most 16M entries
        const uint64_t autoCap = std::min<uint64_t>(std::max<uint64_t>(freeB / 256 / 8, 1ull << 20), 1ull << 24);
        if (!cfg.pair_cap) cfg.pair_cap = autoCap;
        if (!cfg.hit_cap) cfg.hit_cap = autoCap;
        const int device = idx.ctx->device;
        for (int k = 0; k < 2; ++k) {
            Worker& w = s->w[k];
            w.type = k;
            tbCheck(tblaeq_ctx_create(device, nullptr, &w.ctx), "serve: context");
            tbCheck(tblaeq_batch_create(w.ctx, lv.data(), lv.size(), idx.dCoarsestIds.data(), idx.dCoarsestVals.data(),
                                        idx.coarsestIds.size(), maxR.data(), &w.b),
                    "serve: engine");
            tbCheck(tblaeq_batch_reserve(w.b, cfg.max_batch, cfg.pair_cap, cfg.hit_cap), "serve: reserve");
            const size_t qw = k == 0 ? 2 * s->D : s->D;
            w.dQ.resize(cfg.max_batch * qw);
            w.hQ.alloc(cfg.max_batch * qw);
            w.dStatus.resize(4);
            if (k == 0) {
                w.dOff.resize(cfg.max_batch + 1);
                w.dIds.resize(cfg.hit_cap);
                w.hOff.alloc(cfg.max_batch + 1);
                w.hIds.alloc(cfg.hit_cap);
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
        hipCheck(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
        s->budget = static_cast<uint64_t>(static_cast<double>(freeB) * cfg.hbm_fraction);
        s->cfg = cfg;
        tblaeq_server* raw = s.get();
        for (int k = 0; k < 2; ++k) raw->w[k].th = std::thread(worker_loop, raw, std::ref(raw->w[k]));
    } catch (...) {
        const int rc = tblaeq_host::statusOfCurrentException("tblaeq_server_create");
        for (auto& w : s->w) {
            if (w.b) tblaeq_batch_destroy(w.b);
            if (w.ctx) tblaeq_ctx_destroy(w.ctx);
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
    for (auto& w : s->w) {
        // buffers go before the contexts
        w.dQ = DeviceBuffer<double>();
        w.dD2 = DeviceBuffer<double>();
        w.dK = DeviceBuffer<uint64_t>();
        w.dOff = DeviceBuffer<uint64_t>();
        w.dIds = DeviceBuffer<uint64_t>();
        w.dN = DeviceBuffer<uint64_t>();
        w.dStatus = DeviceBuffer<uint64_t>();
        if (w.ctx) tblaeq_ctx_destroy(w.ctx);
        w.ctx = nullptr;
    }
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
    r->est = type == 0 ? s->cfg.range_request_bytes : s->cfg.knn_request_bytes;
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            tblaeq::set_error("tblaeq_server_submit: server is stopping");
            return TBLAEQ_ERR_INVALID;
        }
        r->ticket = s->nextTicket++;
        *ticket = r->ticket;
        s->queue[type].push_back(r.get());
        s->reqs.emplace(r->ticket, std::move(r));
    }
    s->cvWork.notify_all();
    return TBLAEQ_OK;
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

int tblaeq_server_reset_stats(tblaeq_server* s) {
    if (!s) return TBLAEQ_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->latUs.clear();
    s->batches = s->batchedReqs = s->deferred = 0;
    s->peakInflight = s->inflight;
    return TBLAEQ_OK;
}

}  // extern "C"
