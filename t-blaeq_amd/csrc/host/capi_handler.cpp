// C facade over QueryHandler for foreign callers (ctypes, cgo, JNI).
#include <algorithm>
#include <cstring>
#include <iostream>

#include "../../../include/tblaeq_handler.h"
#include "../../../include/tblaeq_learn.h"
#include "../ctx.hpp"
#include "../../../include/tblaeq_build.h"
#include "handler_internal.hpp"

using namespace tblaeq_host;

struct tblaeq_result {
    QueryResult r;
    std::vector<std::vector<uint64_t>> ids;  // host copies of fine meshes
};

namespace {
template <class F>
int guarded(const char* what, F&& f) {
    try {
        f();
        return TBLAEQ_OK;
    } catch (...) {
        return statusOfCurrentException(what);
    }
}
std::vector<size_t> toVec(const uint64_t* r, uint64_t n) { return std::vector<size_t>(r, r + n); }
}  // namespace

extern "C" {

int tblaeq_handler_create_random(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, uint64_t height,
                                 const uint64_t* ratios, uint64_t seed, double sigmaDivisor, tblaeq_handler** out) {
    if (!out || !ratios) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_create_random", [&] {
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(N, D, vmin, vmax, isInt != 0, height, toVec(ratios, height - 1), seed,
                                              sigmaDivisor);
        *out = h;
    });
}

int tblaeq_handler_create_synthetic(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, uint64_t height,
                                    const uint64_t* ratios, uint64_t seed, double sigmaDivisor, uint64_t threads,
                                    tblaeq_handler** out) {
    if (!out || !ratios || height < 2) {
        tblaeq::set_error("tblaeq_handler_create_synthetic: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_create_synthetic", [&] {
        const std::vector<size_t> r = toVec(ratios, height - 1);
        std::unique_ptr<IndexData> idx(IndexBuilder::buildSynthetic(
            N, D, vmin, vmax, isInt != 0, "Syn" + std::to_string(N) + "D" + std::to_string(D), height, r, seed,
            sigmaDivisor, static_cast<unsigned>(std::min<uint64_t>(threads, 1024))));
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(std::move(idx));
        *out = h;
    });
}

int tblaeq_handler_data_rows(tblaeq_handler* h, const uint64_t* rows, uint64_t n, double* out) {
    if (!h || (n && (!rows || !out))) {
        tblaeq::set_error("tblaeq_handler_data_rows: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_data_rows", [&] {
        const IndexData& idx = h->h->index();
        if (idx.finestData.empty()) throw UsageError("tblaeq_handler_data_rows: the index holds no generated data");
        const size_t D = idx.D;
        for (uint64_t i = 0; i < n; ++i) {
            if (rows[i] >= idx.N) throw std::out_of_range("tblaeq_handler_data_rows: row " + std::to_string(rows[i]));
            std::memcpy(out + i * D, idx.finestData.data() + rows[i] * D, D * 8);
        }
    });
}

int tblaeq_handler_create_from_data(const double* data, uint64_t N, uint64_t D, uint64_t height,
                                    const uint64_t* ratios, const char* name, tblaeq_handler** out) {
    if (!out || !data || !ratios) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_create_from_data", [&] {
        PointCloud pc;
        pc.size = static_cast<int64_t>(N);
        pc.dim = static_cast<int64_t>(D);
        pc.data.assign(data, data + N * D);
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(true, pc, name ? name : "in-memory", height, toVec(ratios, height - 1));
        *out = h;
    });
}

int tblaeq_handler_build(const double* data, uint64_t N, uint64_t D, uint64_t height, const uint64_t* ratios,
                         const char* name, int forceCPU, tblaeq_handler** out) {
    if (!out || !data || !ratios || N == 0 || D == 0 || height < 2) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_build", [&] {
        PointCloud pc;
        pc.size = static_cast<int64_t>(N);
        pc.dim = static_cast<int64_t>(D);
        pc.data.assign(data, data + N * D);
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(forceCPU != 0, pc, name ? name : "in-memory", height,
                                              toVec(ratios, height - 1));
        *out = h;
    });
}

int tblaeq_handler_build_ex(const void* data, int dtype, uint64_t N, uint64_t D, uint64_t height,
                            const uint64_t* ratios, const char* name, int forceCPU, const tblaeq_kmeans_options* opt,
                            tblaeq_handler** out) {
    if (!out || !data || !ratios || N == 0 || D == 0 || height < 2 || (dtype != 0 && dtype != 1)) {
        tblaeq::set_error("tblaeq_handler_build_ex: invalid argument (data, ratios, N, D > 0, height >= 2, "
                          "dtype 0 (FP64) or 1 (FP32))");
        return TBLAEQ_ERR_INVALID;
    }
    if (forceCPU && opt && opt->seeding != 0) {
        tblaeq::set_error("tblaeq_handler_build_ex: k-means|| seeding runs on the GPU only");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    return guarded("tblaeq_handler_build_ex", [&] {
        const std::vector<size_t> r = toVec(ratios, height - 1);
        std::unique_ptr<IndexData> idx(IndexBuilder::buildPlanned(
            dtype == 0 ? static_cast<const double*>(data) : nullptr,
            dtype == 1 ? static_cast<const float*>(data) : nullptr, N, D, name ? name : "in-memory", forceCPU != 0,
            height, r, opt));
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(std::move(idx));
        *out = h;
    });
}

int tblaeq_handler_build_stats(tblaeq_handler* h, uint64_t* out, uint64_t cap, uint64_t* n) {
    if (!h || !n) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_build_stats", [&] {
        const auto& s = h->h->index().buildStats;
        *n = s.size();
        if (out) std::copy(s.begin(), s.begin() + std::min<size_t>(cap, s.size()), out);
    });
}

int tblaeq_handler_load(const char* path, tblaeq_handler** out) {
    if (!out || !path) return TBLAEQ_ERR_INVALID;
    int rc = guarded("tblaeq_handler_load", [&] {
        auto* h = new tblaeq_handler();
        h->h = std::make_unique<QueryHandler>(std::string(path), true);
        *out = h;
    });
    return rc == TBLAEQ_ERR_IO || rc == TBLAEQ_ERR_INVALID ? TBLAEQ_ERR_IO : rc;
}

int tblaeq_handler_save(tblaeq_handler* h, const char* path) {
    if (!h || !path) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_save", [&] { h->h->saveIndex(path); });
}

int tblaeq_handler_destroy(tblaeq_handler* h) {
    delete h;
    return TBLAEQ_OK;
}

int tblaeq_handler_meta(tblaeq_handler* h, uint64_t* meta, uint64_t* meshSizes, uint64_t cap) {
    if (!h || !meta) return TBLAEQ_ERR_INVALID;
    IndexData& idx = h->h->index();
    meta[0] = idx.D;
    meta[1] = idx.N;
    meta[2] = idx.height;
    meta[3] = idx.intervals;
    for (uint64_t i = 0; i < cap && i < idx.meshSizes.size(); ++i) meshSizes[i] = idx.meshSizes[i];
    return TBLAEQ_OK;
}

int tblaeq_handler_data(tblaeq_handler* h, double* out, uint64_t cap) {
    if (!h || !out) return TBLAEQ_ERR_INVALID;
    const auto& d = h->h->index().finestData;
    if (d.empty() || cap < d.size()) return TBLAEQ_ERR_CAPACITY;
    std::memcpy(out, d.data(), d.size() * 8);
    return TBLAEQ_OK;
}

int tblaeq_handler_prepare(tblaeq_handler* h, const uint8_t* levels, uint64_t n, uint64_t freeMemOverride) {
    if (!h) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_prepare", [&] {
        if (levels) {
            IndexPolicy p;
            for (uint64_t i = 0; i < n; ++i) {
                if (levels[i] > 3) throw std::invalid_argument("policy level out of range");
                p.levels.push_back(static_cast<LevelPolicy>(levels[i]));
            }
            h->h->prepareForQuery(p);
        } else if (freeMemOverride) {
            h->h->prepareForQuery(static_cast<size_t>(freeMemOverride));
        } else {
            h->h->prepareForQuery();
        }
    });
}

int tblaeq_handler_policy(tblaeq_handler* h, uint8_t* out, uint64_t cap) {
    if (!h || !out) return TBLAEQ_ERR_INVALID;
    const auto& p = h->h->getActivePolicy();
    for (uint64_t i = 0; i < cap && i < p.levels.size(); ++i) out[i] = static_cast<uint8_t>(p.levels[i]);
    return TBLAEQ_OK;
}

int tblaeq_handler_recommend(tblaeq_handler* h, uint64_t freeMem, int countWorkBuffers, uint8_t* out, uint64_t cap) {
    if (!h || !out) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_recommend", [&] {
        IndexData& idx = h->h->index();
        idx.computeStats();
        const IndexPolicy p = PolicyScheduler::recommend(idx, freeMem, countWorkBuffers != 0);
        for (uint64_t i = 0; i < cap && i < p.levels.size(); ++i) out[i] = static_cast<uint8_t>(p.levels[i]);
    });
}

int tblaeq_handler_ctx(tblaeq_handler* h, tblaeq_ctx** out) {
    if (!h || !out) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_ctx", [&] {
        h->h->index().ensureCtx();
        *out = h->h->index().ctx;
    });
}

int tblaeq_handler_set_l3_tile_bytes(uint64_t bytes) {
    L3Strategy::forcedTileBytes = bytes;
    return TBLAEQ_OK;
}

int tblaeq_handler_query(tblaeq_handler* h, int type, const double* queries, uint64_t nq, uint64_t K, int refine,
                         int saveFineMesh, tblaeq_result** out) {
    if (!h || !out || !queries || nq == 0 || (type != 0 && type != 1)) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_query", [&] {
        std::lock_guard<std::mutex> lock(h->plainMu);
        const int D = static_cast<int>(h->h->getDim());
        Query q(static_cast<int>(nq), D, type == 0 ? QueryType::RANGE : QueryType::POINT);
        std::memcpy(q.data.data(), queries, q.data.size() * 8);
        QueryEngine::RunConfig cfg;
        cfg.K = K;
        cfg.refine = refine != 0;
        cfg.copyRefined = refine != 2;
        cfg.saveFineMesh = saveFineMesh != 0;
        auto* r = new tblaeq_result{h->h->performQuery(q, cfg), {}};
        if (r->r.errorCode) {
            delete r;
            throw std::invalid_argument("query dimension mismatch");
        }
        *out = r;
    });
}

int tblaeq_handler_query_learned(tblaeq_handler* h, const double* points, uint64_t nq, uint64_t K,
                                 const double* radii, uint64_t* fallbacks, tblaeq_result** out) {
    if (!h || !out || !points || !radii || nq == 0 || K == 0) return TBLAEQ_ERR_INVALID;
    return guarded("tblaeq_handler_query_learned", [&] {
        std::lock_guard<std::mutex> lock(h->plainMu);
        const int D = static_cast<int>(h->h->getDim());
        Query q(static_cast<int>(nq), D, QueryType::POINT);
        std::memcpy(q.data.data(), points, q.data.size() * 8);
        QueryEngine::RunConfig cfg;
        cfg.K = K;
        cfg.refine = true;
        cfg.radii = radii;
        uint64_t fb = 0;
        cfg.fallbacks = &fb;
        auto* r = new tblaeq_result{h->h->performQuery(q, cfg), {}};
        if (r->r.errorCode) {
            delete r;
            throw std::invalid_argument("query dimension mismatch");
        }
        if (fallbacks) *fallbacks = fb;
        *out = r;
    });
}

int tblaeq_handler_query_file(tblaeq_handler* h, const char* path, int type, int save_fine_mesh, int max_queries,
                              uint64_t K, tblaeq_result** out) {
    if (!h || !path || !out || (type != 0 && type != 1)) {
        tblaeq::set_error("tblaeq_handler_query_file: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_handler_query_file", [&] {
        std::lock_guard<std::mutex> lock(h->plainMu);
        auto* r = new tblaeq_result{h->h->performQuery(std::string(path), type == 0 ? QueryType::RANGE : QueryType::POINT,
                                                       save_fine_mesh != 0, max_queries, K),
                                    {}};
        if (r->r.errorCode) {
            delete r;
            throw std::invalid_argument("the query file's dimension differs from the index's");
        }
        *out = r;
    });
}

int tblaeq_dataset_load(const char* path, uint64_t* n, uint64_t* dim, double* out, uint64_t cap) {
    if (!path || !n || !dim) {
        tblaeq::set_error("tblaeq_dataset_load: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_dataset_load", [&] {
        const PointCloud pc = loadFromFile(path);
        *n = static_cast<uint64_t>(pc.size);
        *dim = static_cast<uint64_t>(pc.dim);
        if (out) std::memcpy(out, pc.data.data(), std::min<uint64_t>(cap, pc.data.size()) * 8);
    });
}

int tblaeq_query_file_load(const char* path, int type, uint64_t* n, uint64_t* dim, double* out, uint64_t cap,
                           char* info, uint64_t info_cap) {
    if (!path || !n || !dim || (type != 0 && type != 1)) {
        tblaeq::set_error("tblaeq_query_file_load: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_query_file_load", [&] {
        const Query q = type == 0 ? loadQueryRangeFromFile(path) : loadQueryPointFromFile(path);
        *n = static_cast<uint64_t>(q.length);
        *dim = static_cast<uint64_t>(q.dim);
        if (out) std::memcpy(out, q.data.data(), std::min<uint64_t>(cap, q.data.size()) * 8);
        if (info && info_cap) {
            const size_t c = std::min<size_t>(info_cap - 1, q.queryRangeInfo.size());
            std::memcpy(info, q.queryRangeInfo.data(), c);
            info[c] = '\0';
        }
    });
}

namespace {
// Rejects a missing result or a query index past the queries it holds.
bool resultArgOk(const tblaeq_result* r, uint64_t i, size_t n, const char* fn, const char* what) {
    if (!r) {
        tblaeq::set_error(std::string(fn) + ": null result");
        return false;
    }
    if (i >= n) {
        tblaeq::set_error(std::string(fn) + ": query " + std::to_string(i) + " out of range (" + std::to_string(n) +
                          " " + what + ")");
        return false;
    }
    return true;
}
}  // namespace

int tblaeq_result_destroy(tblaeq_result* r) {
    delete r;
    return TBLAEQ_OK;
}

int tblaeq_result_info(tblaeq_result* r, uint64_t* nq, uint64_t* levels, double* totalUs) {
    if (!resultArgOk(r, 0, 1, "tblaeq_result_info", "results")) return TBLAEQ_ERR_INVALID;
    if (nq) *nq = r->r.fineMeshSize.size();
    if (levels) *levels = r->r.levelFineMeshSize.empty() ? 0 : r->r.levelFineMeshSize[0].size();
    if (totalUs) *totalUs = static_cast<double>(r->r.totalTimeUs);
    return TBLAEQ_OK;
}

int tblaeq_result_query(tblaeq_result* r, uint64_t i, uint64_t* levelSizes, double* us, uint64_t* fineSize,
                        uint64_t* refinedCount) {
    if (!resultArgOk(r, i, r ? r->r.fineMeshSize.size() : 0, "tblaeq_result_query", "queries"))
        return TBLAEQ_ERR_INVALID;
    if (levelSizes)
        for (size_t l = 0; l < r->r.levelFineMeshSize[i].size(); ++l) levelSizes[l] = r->r.levelFineMeshSize[i][l];
    if (us) *us = r->r.queryTimeUsExact[i];
    if (fineSize) *fineSize = r->r.fineMeshSize[i];
    if (refinedCount) *refinedCount = i < r->r.refinedCount.size() ? r->r.refinedCount[i] : 0;
    return TBLAEQ_OK;
}

int tblaeq_result_levels(tblaeq_result* r, uint64_t i, uint64_t* survivors, uint64_t* outputs) {
    if (!resultArgOk(r, i, r ? r->r.fineMeshSize.size() : 0, "tblaeq_result_levels", "queries"))
        return TBLAEQ_ERR_INVALID;
    for (size_t l = 0; l < r->r.levelFineMeshSize[i].size(); ++l) {
        if (survivors) survivors[l] = r->r.levelSurvivors[i][l];
        if (outputs) outputs[l] = r->r.levelFineMeshSize[i][l];
    }
    return TBLAEQ_OK;
}

int tblaeq_result_fine_mesh(tblaeq_result* r, uint64_t i, uint64_t* ids, double* vals) {
    if (!resultArgOk(r, i, r ? r->r.fineMesh.size() : 0, "tblaeq_result_fine_mesh",
                     "fine meshes kept; the query ran with save_fine_mesh off"))
        return TBLAEQ_ERR_INVALID;
    const SparseGrid* g = r->r.fineMesh[i];
    const size_t n = g->get_nnz_nums(), D = g->get_dimensions();
    if (ids) std::memcpy(ids, g->get_ids_(), n * 8);
    if (vals) std::memcpy(vals, g->get_vals_(), n * D * 8);
    return TBLAEQ_OK;
}

int tblaeq_result_refined(tblaeq_result* r, uint64_t i, uint64_t* ids, double* d2) {
    if (!resultArgOk(r, i, r ? r->r.refinedIds.size() : 0, "tblaeq_result_refined",
                     "refined results kept; the queries ran without the exact refine"))
        return TBLAEQ_ERR_INVALID;
    const auto& v = r->r.refinedIds[i];
    if (ids) std::memcpy(ids, v.data(), v.size() * 8);
    if (d2 && i < r->r.refinedDist2.size()) std::memcpy(d2, r->r.refinedDist2[i].data(), r->r.refinedDist2[i].size() * 8);
    return TBLAEQ_OK;
}

int tblaeq_result_save_csv(tblaeq_result* r, const char* path) {
    if (!r || !path) {
        tblaeq::set_error("tblaeq_result_save_csv: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    return guarded("tblaeq_result_save_csv", [&] { saveQueryResult(r->r, path); });
}

}  // extern "C"
