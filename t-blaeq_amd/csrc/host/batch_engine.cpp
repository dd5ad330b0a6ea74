// Handler-level driver of the batched engine: device-resident queries and
// results, capacity-overflow recovery by splitting the batch.
#include <algorithm>
#include <cstring>

#include "../../../include/tblaeq_handler.h"
#include "../../../include/tblaeq_learn.h"
#include "../batch_internal.hpp"
#include "../ctx.hpp"
#include "handler_internal.hpp"

using namespace tblaeq_host;

namespace {


template <class F>
int guarded(const char* what, F&& f) {
    try {
        return f();
    } catch (...) {
        return statusOfCurrentException(what);
    }
}

BatchState* stateOf(tblaeq_handler* h) { return h ? h->slots[h->slot].get() : nullptr; }

// TBLAEQ_ERR_INVALID with a message naming the entry point (a null handle, a
// batched engine not initialised, no queries uploaded, or a bad argument).
int invalidArg(const char* fn) {
    tblaeq::set_error(std::string(fn) + ": invalid argument, or the batched engine is not initialised or holds "
                      "no queries (batch_init / batch_set_queries first; at most the reserved query count)");
    return TBLAEQ_ERR_INVALID;
}

// A KNN batch holding queries with K above the batched kernels' limit
// (their LDS selection holds K + empty clusters + 1 <= 256 entries): those
// queries run through the per-query traversal, which takes any K (the
// reference puts no limit on K, KNN.cu:75-96), the rest through the batched
// engine; results, statistics and *total are assembled in the batch's rows.
int runKnnMixed(tblaeq_handler* h, BatchState* st, uint64_t kcap, bool stats, uint64_t* total) {
    IndexData& idx = h->h->index();
    const size_t D = idx.D, L = idx.intervals;
    const uint64_t nq = st->nq, Kmax = st->Kmax;
    hipStream_t s = st->stream();
    std::vector<uint64_t> small, large;
    for (uint64_t i = 0; i < nq; ++i) (st->hK[i] <= kcap ? small : large).push_back(i);
    std::vector<uint64_t> hIds(nq * Kmax, 0), hN(nq, 0), hSurv(stats ? nq * L : 0, 0), hFine(stats ? nq : 0, 0);
    std::vector<double> hD2(nq * Kmax, 0.0);
    st->work[0] = st->work[1] = 0;
    if (!small.empty()) {
        const uint64_t ns = small.size();
        uint64_t ks = 0;
        std::vector<double> q(ns * D);
        std::vector<uint64_t> k(ns);
        for (uint64_t j = 0; j < ns; ++j) {
            std::copy_n(st->hQ.data() + small[j] * D, D, q.data() + j * D);
            k[j] = st->hK[small[j]];
            ks = std::max(ks, k[j]);
        }
        if (ks == 0) ks = 1;
        DeviceBuffer<double> dQ(ns * D), dD2(ns * ks);
        DeviceBuffer<uint64_t> dK(ns), dIds(ns * ks), dN(ns), dSurv(stats ? ns * L : 1), dFine(stats ? ns : 1);
        hipCheck(hipMemcpy(dQ.data(), q.data(), ns * D * 8, hipMemcpyHostToDevice), "H2D");
        hipCheck(hipMemcpy(dK.data(), k.data(), ns * 8, hipMemcpyHostToDevice), "H2D");
        std::vector<std::pair<uint64_t, uint64_t>> todo{{0, ns}};
        while (!todo.empty()) {
            auto [a, n] = todo.back();
            todo.pop_back();
            int rc = tblaeq_batch_knn(st->b, dQ.data() + a * D, dK.data() + a, n, ks, dIds.data() + a * ks,
                                      dD2.data() + a * ks, dN.data() + a, st->dStatus.data(),
                                      stats ? dSurv.data() + a * L : nullptr, stats ? dFine.data() + a : nullptr);
            if (rc) return rc;
            uint64_t status[4];
            hipCheck(hipMemcpyAsync(status, st->dStatus.data(), 32, hipMemcpyDeviceToHost, s), "D2H");
            hipCheck(hipStreamSynchronize(s), "sync");
            if (status[1] != 0) {
                if (n == 1) {
                    tblaeq::set_error("batched KNN: one query exceeds the engine capacities");
                    return TBLAEQ_ERR_CAPACITY;
                }
                todo.push_back({a + n / 2, n - n / 2});
                todo.push_back({a, n / 2});
                continue;
            }
            st->work[0] += status[2];
            st->work[1] += status[3];
        }
        std::vector<uint64_t> ids(ns * ks), n(ns), sv(stats ? ns * L : 0), fn(stats ? ns : 0);
        std::vector<double> d2(ns * ks);
        hipCheck(hipMemcpy(ids.data(), dIds.data(), ns * ks * 8, hipMemcpyDeviceToHost), "D2H");
        hipCheck(hipMemcpy(d2.data(), dD2.data(), ns * ks * 8, hipMemcpyDeviceToHost), "D2H");
        hipCheck(hipMemcpy(n.data(), dN.data(), ns * 8, hipMemcpyDeviceToHost), "D2H");
        if (stats) {
            hipCheck(hipMemcpy(sv.data(), dSurv.data(), ns * L * 8, hipMemcpyDeviceToHost), "D2H");
            hipCheck(hipMemcpy(fn.data(), dFine.data(), ns * 8, hipMemcpyDeviceToHost), "D2H");
        }
        for (uint64_t j = 0; j < ns; ++j) {
            const uint64_t i = small[j];
            hN[i] = n[j];
            std::copy_n(ids.data() + j * ks, n[j], hIds.data() + i * Kmax);
            std::copy_n(d2.data() + j * ks, n[j], hD2.data() + i * Kmax);
            if (stats) {
                std::copy_n(sv.data() + j * L, L, hSurv.data() + i * L);
                hFine[i] = fn[j];
            }
        }
    }
    for (uint64_t i : large) {
        std::vector<uint64_t> ids, lv;
        std::vector<double> d2;
        uint64_t fine = 0;
        plainKnn(h, st->hQ.data() + i * D, st->hK[i], ids, d2, stats ? &lv : nullptr, stats ? &fine : nullptr);
        hN[i] = ids.size();
        std::copy(ids.begin(), ids.end(), hIds.begin() + static_cast<long>(i * Kmax));
        std::copy(d2.begin(), d2.end(), hD2.begin() + static_cast<long>(i * Kmax));
        if (stats) {
            std::copy(lv.begin(), lv.end(), hSurv.begin() + static_cast<long>(i * L));
            hFine[i] = fine;
        }
    }
    hipCheck(hipMemcpy(st->dIds.data(), hIds.data(), nq * Kmax * 8, hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(st->dD2.data(), hD2.data(), nq * Kmax * 8, hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(st->dN.data(), hN.data(), nq * 8, hipMemcpyHostToDevice), "H2D");
    if (stats) {
        hipCheck(hipMemcpy(st->dSurv.data(), hSurv.data(), nq * L * 8, hipMemcpyHostToDevice), "H2D");
        hipCheck(hipMemcpy(st->dFine.data(), hFine.data(), nq * 8, hipMemcpyHostToDevice), "H2D");
    }
    if (total) {
        uint64_t t = 0;
        for (uint64_t v : hN) t += v;
        *total = t;
    }
    return TBLAEQ_OK;
}

}  // namespace

namespace tblaeq_host {

void plainKnn(tblaeq_handler* h, const double* point, uint64_t K, std::vector<uint64_t>& ids,
              std::vector<double>& d2, std::vector<uint64_t>* levelSurv, uint64_t* fine) {
    std::lock_guard<std::mutex> lock(h->plainMu);
    const int D = static_cast<int>(h->h->getDim());
    Query q(1, D, QueryType::POINT);
    std::copy_n(point, D, q.data.data());
    QueryEngine::RunConfig cfg;
    cfg.K = K;
    cfg.refine = true;
    QueryResult r = h->h->performQuery(q, cfg);
    if (r.errorCode) throw std::invalid_argument("query dimension mismatch");
    ids = r.refinedIds[0];
    d2 = r.refinedDist2[0];
    if (levelSurv) levelSurv->assign(r.levelSurvivors[0].begin(), r.levelSurvivors[0].end());
    if (fine) *fine = r.fineMeshSize[0];
}

}  // namespace tblaeq_host

extern "C" {

int tblaeq_handler_batch_init(tblaeq_handler* h, uint64_t maxQueries, uint64_t pairCap, uint64_t hitCap) {
    if (!h || maxQueries == 0) return invalidArg("tblaeq_handler_batch_init");
    return guarded("tblaeq_handler_batch_init", [&]() -> int {
        IndexData& idx = h->h->index();
        if (!h->h->prepared()) h->h->prepareForQuery(LevelPolicy::L0);
        for (size_t l = 0; l < idx.intervals; ++l) idx.uploadLevelValues(l);
        h->slots[h->slot].reset();
        auto st = std::make_unique<BatchState>();
        if (h->slot == 0) {
            st->ctx = idx.ctx;
        } else {
            tbCheck(tblaeq_ctx_create(idx.ctx->device, nullptr, &st->ctx), "batch slot context");
            st->ownsCtx = true;
        }
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&st->hStatus), 8 * sizeof(uint64_t), 0), "hipHostMalloc");
        std::vector<tblaeq_level> lv(idx.intervals);
        std::vector<double> maxR(idx.intervals, 0.0);
        for (size_t l = 0; l < idx.intervals; ++l) {
            lv[l] = idx.levelDesc(l);
            for (double r : idx.meshMaxRadius[l]) maxR[l] = std::max(maxR[l], r);
        }
        int rc = tblaeq_batch_create(st->ctx, lv.data(), lv.size(), idx.dCoarsestIds.data(), idx.dCoarsestVals.data(),
                                     idx.coarsestIds.size(), maxR.data(), &st->b);
        if (rc) return rc;
        rc = tblaeq_batch_reserve(st->b, maxQueries, pairCap, hitCap);
        if (rc) return rc;
        st->maxQ = maxQueries;
        st->pairCap = pairCap;
        st->hitCap = hitCap;
        st->dQ.resize(maxQueries * 2 * idx.D);
        st->dK.resize(maxQueries);
        st->dOff.resize(maxQueries + 1);
        st->dOffTmp.resize(maxQueries + 1);
        st->dStatus.resize(4);
        st->dSurv.resize(maxQueries * idx.intervals);
        st->dFine.resize(maxQueries);
        st->dN.resize(maxQueries);
        st->dIds.resize(std::max<uint64_t>(hitCap, 1));
        h->slots[h->slot] = std::move(st);
        return TBLAEQ_OK;
    });
}

int tblaeq_handler_batch_set_queries(tblaeq_handler* h, int type, const double* queries, const uint64_t* Ks,
                                     uint64_t nq) {
    BatchState* st = stateOf(h);
    if (!st || !queries || nq == 0 || nq > st->maxQ || (type != 0 && type != 1) || (type == 1 && !Ks)) {
        tblaeq::set_error("tblaeq_handler_batch_set_queries: invalid argument (batch_init first; 1 <= nq <= " +
                          std::to_string(st ? st->maxQ : 0) + " queries; type 0 or 1; KNN needs Ks)");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_batch_set_queries", [&]() -> int {
        IndexData& idx = h->h->index();
        const size_t D = idx.D;
        st->type = type;
        st->nq = nq;
        hipCheck(hipMemcpy(st->dQ.data(), queries, nq * (type == 0 ? 2 * D : D) * 8, hipMemcpyHostToDevice), "H2D");
        st->hQ.assign(queries, queries + nq * (type == 0 ? 2 * D : D));
        if (type == 1) st->hK.assign(Ks, Ks + nq);
        if (type == 1) {
            hipCheck(hipMemcpy(st->dK.data(), Ks, nq * 8, hipMemcpyHostToDevice), "H2D");
            st->Kmax = *std::max_element(Ks, Ks + nq);
            if (st->dIds.size() < nq * st->Kmax) st->dIds.resize(nq * st->Kmax);
            if (st->dD2.size() < nq * st->Kmax) st->dD2.resize(std::max<uint64_t>(nq * st->Kmax, 1));
        }
        return TBLAEQ_OK;
    });
}

// Runs the uploaded batch.  Range: one graph launch when no statistics are
// asked for; splits the batch in halves while a capacity is exceeded.
// *total = hits (range) or results (KNN).
int tblaeq_handler_batch_run(tblaeq_handler* h, int stats, uint64_t* total) {
    BatchState* st = stateOf(h);
    if (!st || st->nq == 0) return invalidArg("tblaeq_handler_batch_run");
    return guarded("tblaeq_handler_batch_run", [&]() -> int {
        IndexData& idx = h->h->index();
        const size_t D = idx.D, L = idx.intervals;
        hipStream_t s = st->stream();
        st->work[0] = st->work[1] = 0;
        if (stats) {
            hipCheck(hipMemsetAsync(st->dSurv.data(), 0, st->nq * L * 8, s), "memset");
            hipCheck(hipMemsetAsync(st->dFine.data(), 0, st->nq * 8, s), "memset");
        }
        if (st->type == 1) {
            uint64_t kcap = 0;
            tbCheck(tblaeq_batch_knn_max_k(st->b, &kcap), "batch_knn_max_k");
            if (st->Kmax > kcap) return runKnnMixed(h, st, kcap, stats != 0, total);
            // split in halves while a capacity is exceeded; KNN rows are fixed-size
            std::vector<std::pair<uint64_t, uint64_t>> todo{{0, st->nq}};
            while (!todo.empty()) {
                auto [a, n] = todo.back();
                todo.pop_back();
                int rc = (!stats && st->useGraph)
                             ? tblaeq_batch_knn_graph(st->b, st->dQ.data() + a * D, st->dK.data() + a, n, st->Kmax,
                                                      st->dIds.data() + a * st->Kmax, st->dD2.data() + a * st->Kmax,
                                                      st->dN.data() + a, st->dStatus.data())
                             : tblaeq_batch_knn(st->b, st->dQ.data() + a * D, st->dK.data() + a, n, st->Kmax,
                                                st->dIds.data() + a * st->Kmax, st->dD2.data() + a * st->Kmax,
                                                st->dN.data() + a, st->dStatus.data(),
                                                stats ? st->dSurv.data() + a * L : nullptr,
                                                stats ? st->dFine.data() + a : nullptr);
                if (rc) return rc;
                uint64_t status[4];
                hipCheck(hipMemcpyAsync(status, st->dStatus.data(), 32, hipMemcpyDeviceToHost, s), "D2H");
                hipCheck(hipStreamSynchronize(s), "sync");
                if (status[1] != 0) {
                    if (n == 1) {
                        tblaeq::set_error("batched KNN: one query exceeds the engine capacities");
                        return TBLAEQ_ERR_CAPACITY;
                    }
                    todo.push_back({a + n / 2, n - n / 2});
                    todo.push_back({a, n / 2});
                    continue;
                }
                st->work[0] += status[2];
                st->work[1] += status[3];
            }
            if (total) {
                std::vector<uint64_t> n(st->nq);
                hipCheck(hipMemcpy(n.data(), st->dN.data(), st->nq * 8, hipMemcpyDeviceToHost), "D2H");
                uint64_t t = 0;
                for (auto v : n) t += v;
                *total = t;
            }
            return TBLAEQ_OK;
        }
        if (!stats && st->useGraph) {
            int rc = tblaeq_batch_range_graph(st->b, st->dQ.data(), st->nq, st->dOff.data(), st->dIds.data(),
                                              st->dIds.size(), st->dStatus.data());
            if (rc) return rc;
            uint64_t status[4];
            hipCheck(hipMemcpyAsync(status, st->dStatus.data(), 32, hipMemcpyDeviceToHost, s), "D2H");
            hipCheck(hipStreamSynchronize(s), "sync");
            if (status[1] == 0) {
                st->work[0] = status[2];
                st->work[1] = status[3];
                st->lastTotal = status[0];
                if (total) *total = status[0];
                return TBLAEQ_OK;
            }
            // capacity exceeded: fall through to the splitting path
        }
        // range: [a, a+n) sub-batches
        std::vector<std::pair<uint64_t, uint64_t>> todo{{0, st->nq}};
        uint64_t base = 0;
        std::vector<uint64_t> hostOff(st->nq + 1, 0);
        bool split = false;
        while (!todo.empty()) {
            auto [a, n] = todo.front();
            todo.erase(todo.begin());
            const uint64_t cap = st->dIds.size() - base;
            int rc = tblaeq_batch_range(st->b, st->dQ.data() + a * 2 * D, n, split ? st->dOffTmp.data() : st->dOff.data(),
                                        st->dIds.data() + base, cap, st->dStatus.data(),
                                        stats ? st->dSurv.data() + a * L : nullptr,
                                        stats ? st->dFine.data() + a : nullptr);
            if (rc) return rc;
            uint64_t status[4];
            hipCheck(hipMemcpyAsync(status, st->dStatus.data(), 32, hipMemcpyDeviceToHost, s), "D2H");
            hipCheck(hipStreamSynchronize(s), "sync");
            if (status[1] != 0) {
                if (n == 1) {
                    tblaeq::set_error("batched range: one query exceeds the engine capacities");
                    return TBLAEQ_ERR_CAPACITY;
                }
                if (stats) {
                    hipCheck(hipMemsetAsync(st->dSurv.data() + a * L, 0, n * L * 8, s), "memset");
                    hipCheck(hipMemsetAsync(st->dFine.data() + a, 0, n * 8, s), "memset");
                }
                if (!split) {
                    // restart from scratch in split mode
                    split = true;
                    todo.clear();
                    base = 0;
                    todo.push_back({0, n / 2});
                    todo.push_back({n / 2, n - n / 2});
                    if (stats) {
                        hipCheck(hipMemsetAsync(st->dSurv.data(), 0, st->nq * L * 8, s), "memset");
                        hipCheck(hipMemsetAsync(st->dFine.data(), 0, st->nq * 8, s), "memset");
                    }
                    continue;
                }
                todo.insert(todo.begin(), {a + n / 2, n - n / 2});
                todo.insert(todo.begin(), {a, n / 2});
                continue;
            }
            if (split) {
                std::vector<uint64_t> off(n + 1);
                hipCheck(hipMemcpy(off.data(), st->dOffTmp.data(), (n + 1) * 8, hipMemcpyDeviceToHost), "D2H");
                for (uint64_t i = 0; i <= n; ++i) hostOff[a + i] = base + off[i];
            }
            base += status[0];
            st->work[0] += status[2];
            st->work[1] += status[3];
        }
        if (split) hipCheck(hipMemcpy(st->dOff.data(), hostOff.data(), (st->nq + 1) * 8, hipMemcpyHostToDevice), "H2D");
        st->lastTotal = base;
        if (total) *total = base;
        return TBLAEQ_OK;
    });
}

int tblaeq_handler_batch_results(tblaeq_handler* h, uint64_t* off, uint64_t* ids, double* d2, uint64_t* n) {
    BatchState* st = stateOf(h);
    if (!st || st->nq == 0) return invalidArg("tblaeq_handler_batch_results");
    return guarded("tblaeq_handler_batch_results", [&]() -> int {
        if (st->type == 0) {
            std::vector<uint64_t> o(st->nq + 1);
            hipCheck(hipMemcpy(o.data(), st->dOff.data(), (st->nq + 1) * 8, hipMemcpyDeviceToHost), "D2H");
            if (off) std::memcpy(off, o.data(), (st->nq + 1) * 8);
            if (ids) hipCheck(hipMemcpy(ids, st->dIds.data(), o[st->nq] * 8, hipMemcpyDeviceToHost), "D2H");
        } else {
            if (n) hipCheck(hipMemcpy(n, st->dN.data(), st->nq * 8, hipMemcpyDeviceToHost), "D2H");
            if (ids) hipCheck(hipMemcpy(ids, st->dIds.data(), st->nq * st->Kmax * 8, hipMemcpyDeviceToHost), "D2H");
            if (d2) hipCheck(hipMemcpy(d2, st->dD2.data(), st->nq * st->Kmax * 8, hipMemcpyDeviceToHost), "D2H");
        }
        return TBLAEQ_OK;
    });
}

int tblaeq_handler_batch_stats(tblaeq_handler* h, uint64_t* levelSurv, uint64_t* fine) {
    BatchState* st = stateOf(h);
    if (!st || st->nq == 0) return invalidArg("tblaeq_handler_batch_stats");
    return guarded("tblaeq_handler_batch_stats", [&]() -> int {
        const size_t L = h->h->index().intervals;
        if (levelSurv) hipCheck(hipMemcpy(levelSurv, st->dSurv.data(), st->nq * L * 8, hipMemcpyDeviceToHost), "D2H");
        if (fine) hipCheck(hipMemcpy(fine, st->dFine.data(), st->nq * 8, hipMemcpyDeviceToHost), "D2H");
        return TBLAEQ_OK;
    });
}

int tblaeq_handler_batch_use_graph(tblaeq_handler* h, int on) {
    BatchState* st = stateOf(h);
    if (!st) return invalidArg("tblaeq_handler_batch_use_graph");
    st->useGraph = on != 0;
    return TBLAEQ_OK;
}

int tblaeq_handler_batch_set_param(tblaeq_handler* h, int param, uint64_t value) {
    BatchState* st = stateOf(h);
    if (!st) {
        tblaeq::set_error("tblaeq_handler_batch_set_param: call tblaeq_handler_batch_init first");
        return TBLAEQ_ERR_INVALID;
    }
    return tblaeq_batch_set_param(st->b, param, value);
}

int tblaeq_handler_batch_work(tblaeq_handler* h, uint64_t* work) {
    BatchState* st = stateOf(h);
    if (!st || !work) return invalidArg("tblaeq_handler_batch_work");
    work[0] = st->work[0];
    work[1] = st->work[1];
    return TBLAEQ_OK;
}

int tblaeq_handler_batch_leaf_tests(tblaeq_handler* h, uint64_t* out) {
    BatchState* st = stateOf(h);
    if (!st || !out) return invalidArg("tblaeq_handler_batch_leaf_tests");
    return tblaeq_batch_leaf_tests(st->b, out);
}

int tblaeq_handler_batch_autotune(tblaeq_handler* h, const double* boxes, uint64_t nBoxes, const double* points,
                                  const uint64_t* Ks, uint64_t nPoints, int reps, double* outMs) {
    BatchState* st = stateOf(h);
    if (!st || (nBoxes && !boxes) || (nPoints && (!points || !Ks))) {
        tblaeq::set_error("tblaeq_handler_batch_autotune: invalid argument or batch_init not called");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_batch_autotune", [&]() -> int {
        const size_t D = h->h->index().D;
        DeviceBuffer<double> dB, dP;
        DeviceBuffer<uint64_t> dK;
        uint64_t Kmax = 0;
        if (nBoxes) dB.reset(boxes, nBoxes * 2 * D);
        if (nPoints) {
            dP.reset(points, nPoints * D);
            dK.reset(Ks, nPoints);
            Kmax = *std::max_element(Ks, Ks + nPoints);
        }
        return tblaeq_batch_autotune(st->b, nBoxes ? dB.data() : nullptr, nBoxes, nPoints ? dP.data() : nullptr,
                                     nPoints ? dK.data() : nullptr, nPoints, Kmax, reps, outMs);
    });
}

int tblaeq_handler_batch_get_param(tblaeq_handler* h, int param, uint64_t* value) {
    BatchState* st = stateOf(h);
    if (!st) {
        tblaeq::set_error("tblaeq_handler_batch_get_param: call tblaeq_handler_batch_init first");
        return TBLAEQ_ERR_INVALID;
    }
    return tblaeq_batch_get_param(st->b, param, value);
}

int tblaeq_handler_batch_save_params(tblaeq_handler* h, const char* path) {
    BatchState* st = stateOf(h);
    if (!st) return invalidArg("tblaeq_handler_batch_save_params");
    return tblaeq_batch_save_params(st->b, path);
}

int tblaeq_handler_batch_load_params(tblaeq_handler* h, const char* path) {
    BatchState* st = stateOf(h);
    if (!st) return invalidArg("tblaeq_handler_batch_load_params");
    return tblaeq_batch_load_params(st->b, path);
}

int tblaeq_handler_knn_features(tblaeq_handler* h, const double* points, const uint64_t* Ks, uint64_t n, double vmin,
                                double scale, float* X) {
    BatchState* st = stateOf(h);
    if (!st || !points || !Ks || !X || n == 0) {
        tblaeq::set_error("tblaeq_handler_knn_features: invalid argument or batch_init not called");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_knn_features", [&]() -> int {
        const size_t D = h->h->index().D;
        DeviceBuffer<double> dP(points, n * D);
        DeviceBuffer<uint64_t> dK(Ks, n);
        DeviceBuffer<float> dX(n * 48);
        tbCheck(tblaeq_knn_features(st->b, dP.data(), dK.data(), n, vmin, scale, dX.data()), "features");
        hipStream_t s = st->stream();
        hipCheck(hipMemcpyAsync(X, dX.data(), n * 48 * 4, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipStreamSynchronize(s), "sync");
        return TBLAEQ_OK;
    });
}

// Learned KNN over the uploaded batch: features -> network -> radii -> certified
// batch; queries that fail certification are rerun with the seed method.
int tblaeq_handler_batch_knn_learned(tblaeq_handler* h, tblaeq_mlp* m, double factor, int stats,
                                     uint64_t* fallbacks) {
    BatchState* st = stateOf(h);
    if (!st || !m || st->type != 1 || st->nq == 0 || !(factor > 0.0)) {
        tblaeq::set_error("tblaeq_handler_batch_knn_learned: upload a KNN batch first (and factor > 0)");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_batch_knn_learned", [&]() -> int {
        IndexData& idx = h->h->index();
        const size_t D = idx.D, L = idx.intervals;
        hipStream_t s = st->stream();
        const uint64_t nq = st->nq, Kmax = st->Kmax;
        double vmin = 0, scale = 1;
        tblaeq_mlp_get_norm(m, &vmin, &scale);
        const uint64_t cap = st->maxQ;
        if (st->lX.size() < cap * 48) {
            st->lX.resize(cap * 48);
            st->lR.resize(cap);
            st->lF.resize(cap);
            st->lIdx.resize(cap);
            st->lQ.resize(cap * D);
            st->lK.resize(cap);
            st->lIds.resize(cap * st->Kmax);
            st->lD2.resize(cap * st->Kmax);
            st->lN.resize(cap);
            st->lSurv.resize(cap * L);
            st->lFine.resize(cap);
        }
        if (st->lIds.size() < cap * Kmax) {
            st->lIds.resize(cap * Kmax);
            st->lD2.resize(cap * Kmax);
        }
        float* dX = st->lX.data();
        double* dR = st->lR.data();
        uint8_t* dF = st->lF.data();
        tbCheck(tblaeq_knn_features(st->b, st->dQ.data(), st->dK.data(), nq, vmin, scale, dX), "features");
        hipCheck(hipStreamSynchronize(s), "sync");
        tbCheck(tblaeq_mlp_propose_radius(m, dX, nq, scale, factor, dR), "propose radius");
        hipCheck(hipDeviceSynchronize(), "sync");
        st->work[0] = st->work[1] = 0;
        if (stats) {
            hipCheck(hipMemsetAsync(st->dSurv.data(), 0, nq * L * 8, s), "memset");
            hipCheck(hipMemsetAsync(st->dFine.data(), 0, nq * 8, s), "memset");
        }
        // one engine call over [a, a + n), split in halves while a capacity is exceeded
        auto run = [&](bool seed, const double* Q, const uint64_t* K, uint64_t n0, uint64_t* ids, double* d2,
                       uint64_t* cnt, uint64_t* surv, uint64_t* fine) -> int {
            std::vector<std::pair<uint64_t, uint64_t>> todo{{0, n0}};
            while (!todo.empty()) {
                auto [a, n] = todo.back();
                todo.pop_back();
                uint64_t* sv = stats ? surv + a * L : nullptr;
                uint64_t* fn = stats ? fine + a : nullptr;
                if (seed)
                    tbCheck(tblaeq_batch_knn(st->b, Q + a * D, K + a, n, Kmax, ids + a * Kmax, d2 + a * Kmax, cnt + a,
                                             st->dStatus.data(), sv, fn),
                            "fallback batch");
                else
                    tbCheck(tblaeq_batch_knn_radius(st->b, Q + a * D, K + a, n, Kmax, dR + a, ids + a * Kmax,
                                                    d2 + a * Kmax, cnt + a, dF + a, st->dStatus.data(), sv, fn),
                            "learned batch");
                uint64_t status[4];
                hipCheck(hipMemcpyAsync(status, st->dStatus.data(), 32, hipMemcpyDeviceToHost, s), "D2H");
                hipCheck(hipStreamSynchronize(s), "sync");
                if (status[1] != 0) {
                    if (n == 1) {
                        tblaeq::set_error("learned KNN: one query exceeds the engine capacities");
                        return TBLAEQ_ERR_CAPACITY;
                    }
                    todo.push_back({a + n / 2, n - n / 2});
                    todo.push_back({a, n / 2});
                    continue;
                }
                st->work[0] += status[2];
                st->work[1] += status[3];
            }
            return TBLAEQ_OK;
        };
        int rc = run(false, st->dQ.data(), st->dK.data(), nq, st->dIds.data(), st->dD2.data(), st->dN.data(),
                     st->dSurv.data(), st->dFine.data());
        if (rc) return rc;
        std::vector<uint8_t> fb(nq);
        hipCheck(hipMemcpy(fb.data(), dF, nq, hipMemcpyDeviceToHost), "D2H");
        std::vector<uint64_t> redo;
        for (uint64_t q = 0; q < nq; ++q)
            if (fb[q]) redo.push_back(q);
        if (fallbacks) *fallbacks = redo.size();
        if (!redo.empty()) {
            // rerun the failed queries with the seed method on gathered rows and
            // scatter their result rows back
            const uint64_t n = redo.size();
            tblaeq_ctx* c = st->b->ctx;
            uint64_t* ridx = st->lIdx.data();
            hipCheck(hipMemcpyAsync(ridx, redo.data(), n * 8, hipMemcpyHostToDevice, s), "H2D");
            tblaeq::move_rows(c, false, st->dQ.data(), ridx, n, D, st->lQ.data());
            tblaeq::move_rows(c, false, st->dK.data(), ridx, n, 1, st->lK.data());
            if (stats) {
                hipCheck(hipMemsetAsync(st->lSurv.data(), 0, n * L * 8, s), "memset");
                hipCheck(hipMemsetAsync(st->lFine.data(), 0, n * 8, s), "memset");
            }
            rc = run(true, st->lQ.data(), st->lK.data(), n, st->lIds.data(), st->lD2.data(), st->lN.data(),
                     st->lSurv.data(), st->lFine.data());
            if (rc) return rc;
            tblaeq::move_rows(c, true, st->lIds.data(), ridx, n, Kmax, st->dIds.data());
            tblaeq::move_rows(c, true, st->lD2.data(), ridx, n, Kmax, st->dD2.data());
            tblaeq::move_rows(c, true, st->lN.data(), ridx, n, 1, st->dN.data());
            if (stats) {
                tblaeq::move_rows(c, true, st->lSurv.data(), ridx, n, L, st->dSurv.data());
                tblaeq::move_rows(c, true, st->lFine.data(), ridx, n, 1, st->dFine.data());
            }
            hipCheck(hipGetLastError(), "move rows");
            hipCheck(hipStreamSynchronize(s), "sync");
        }
        return TBLAEQ_OK;
    });
}

int tblaeq_handler_batch_slot(tblaeq_handler* h, int slot) {
    if (!h || slot < 0 || slot >= kBatchSlots) {
        tblaeq::set_error("tblaeq_handler_batch_slot: slot must be 0 or 1");
        return TBLAEQ_ERR_INVALID;
    }
    h->slot = slot;
    return TBLAEQ_OK;
}

int tblaeq_handler_batch_run_slots(tblaeq_handler* h, uint64_t mask, uint64_t* totals) {
    if (!h || mask == 0 || mask >= (1ull << kBatchSlots)) {
        tblaeq::set_error("tblaeq_handler_batch_run_slots: mask must select slots 0 and/or 1");
        return TBLAEQ_ERR_INVALID;
    }
    for (int i = 0; i < kBatchSlots; ++i)
        if (((mask >> i) & 1) && (!h->slots[i] || h->slots[i]->nq == 0)) {
            tblaeq::set_error("tblaeq_handler_batch_run_slots: slot " + std::to_string(i) +
                              " is not initialised or holds no queries");
            return TBLAEQ_ERR_INVALID;
        }
    const int saved = h->slot;
    const int rc = guarded("tblaeq_handler_batch_run_slots", [&]() -> int {
        // launch every selected slot's captured batch on its own stream first
        bool launched[kBatchSlots] = {};
        for (int i = 0; i < kBatchSlots; ++i) {
            if (!((mask >> i) & 1)) continue;
            BatchState* st = h->slots[i].get();
            if (!st->useGraph) continue;
            int r = TBLAEQ_OK;
            if (st->type == 0) {
                r = tblaeq_batch_range_graph(st->b, st->dQ.data(), st->nq, st->dOff.data(), st->dIds.data(),
                                             st->dIds.size(), st->dStatus.data());
            } else {
                uint64_t kcap = 0;
                tbCheck(tblaeq_batch_knn_max_k(st->b, &kcap), "batch_knn_max_k");
                if (st->Kmax > kcap) continue;  // mixed K: the sequential path below
                r = tblaeq_batch_knn_graph(st->b, st->dQ.data(), st->dK.data(), st->nq, st->Kmax, st->dIds.data(),
                                           st->dD2.data(), st->dN.data(), st->dStatus.data());
            }
            if (r) return r;
            hipCheck(hipMemcpyAsync(st->hStatus, st->dStatus.data(), 32, hipMemcpyDeviceToHost, st->stream()), "D2H");
            launched[i] = true;
        }
        // then wait for each; a slot whose capacities overflowed, or that was
        // not launched, runs through batch_run (which splits)
        for (int i = 0; i < kBatchSlots; ++i) {
            if (!((mask >> i) & 1)) continue;
            BatchState* st = h->slots[i].get();
            uint64_t t = 0;
            if (launched[i]) {
                hipCheck(hipStreamSynchronize(st->stream()), "sync");
                const uint64_t* status = st->hStatus;
                if (status[1] == 0) {
                    st->work[0] = status[2];
                    st->work[1] = status[3];
                    if (st->type == 0) {
                        t = status[0];
                        st->lastTotal = t;
                    } else {
                        std::vector<uint64_t> n(st->nq);
                        hipCheck(hipMemcpy(n.data(), st->dN.data(), st->nq * 8, hipMemcpyDeviceToHost), "D2H");
                        for (auto v : n) t += v;
                    }
                    if (totals) totals[i] = t;
                    continue;
                }
            }
            h->slot = i;
            const int r = tblaeq_handler_batch_run(h, 0, &t);
            h->slot = saved;
            if (r) return r;
            if (totals) totals[i] = t;
        }
        return TBLAEQ_OK;
    });
    h->slot = saved;
    return rc;
}

int tblaeq_handler_batch_free(tblaeq_handler* h) {
    if (h) h->slots[h->slot].reset();
    return TBLAEQ_OK;
}

}  // extern "C"
