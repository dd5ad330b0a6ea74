// Memory-policy scheduler and the four per-level strategies.
//
// MemoryPolicy.cu:100-316 and strategies/L{0,1,2,3}Strategy.cu.  All four
// strategies produce the same grid for the same level input (same kernels,
// same order); they differ only in what is resident on the device:
//   L0  P values, radii and per-level work buffers resident; no allocation and
//       no host synchronisation inside a query.
//   L1  P values and radii resident; survivor and output buffers allocated per
//       level at their exact size (one host sync per level).
//   L2  radii uploaded per level, P columns of the survivors packed on the host
//       in output order and uploaded (SpMSpV.cu:107-198).
//   L3  grids live on the host and are streamed through the device in tiles
//       sized from free memory (L3Strategy.cu:22-567).
#include <algorithm>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <numeric>

#include "../../../include/tblaeq_host.hpp"

namespace tblaeq_host {

namespace {

hipStream_t streamOf(IndexData& idx) { return static_cast<hipStream_t>(tblaeq_ctx_stream(idx.ctx)); }

void d2h(IndexData& idx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, streamOf(idx)), "hipMemcpyAsync D2H");
    hipCheck(hipStreamSynchronize(streamOf(idx)), "hipStreamSynchronize");
}

void h2d(IndexData& idx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, streamOf(idx)), "hipMemcpyAsync H2D");
}

// Host-known element count of a grid (synchronises when it is still on the device).
size_t hostCount(IndexData& idx, SparseGrid* g) {
    if (g->dCount) {
        uint64_t n = 0;
        d2h(idx, &n, g->dCount, 8);
        g->set_nnz(n);
        g->dCount = nullptr;
    }
    return g->get_nnz_nums();
}

// Common prepare for every policy mix (fixes the reference's mixed-policy gap
// where L1 levels ran without their P values, L0Strategy.cu:34-42).
void prepareAll(IndexData& idx, std::ostream& os) {
    idx.computeStats();
    PolicyScheduler::printReport(idx, idx.activePolicy, os);
    idx.uploadPermanentData();
    for (size_t l = 0; l < idx.intervals; ++l) {
        const LevelPolicy p = idx.activePolicy[l];
        if (p == LevelPolicy::L0 || p == LevelPolicy::L1) idx.uploadLevelValues(l);
        if (p == LevelPolicy::L0) idx.allocLevelWork(l);
    }
    hipCheck(hipDeviceSynchronize(), "prepare");
}

}  // namespace

const char* levelPolicyName(LevelPolicy p) {
    switch (p) {
        case LevelPolicy::L0: return "L0";
        case LevelPolicy::L1: return "L1";
        case LevelPolicy::L2: return "L2";
        case LevelPolicy::L3: return "L3";
    }
    return "?";
}

LevelPolicy IndexPolicy::parseLevel(const std::string& s) {
    if (s == "L0" || s == "l0") return LevelPolicy::L0;
    if (s == "L1" || s == "l1") return LevelPolicy::L1;
    if (s == "L2" || s == "l2") return LevelPolicy::L2;
    if (s == "L3" || s == "l3") return LevelPolicy::L3;
    throw std::invalid_argument("Unknown policy level '" + s + "'. Valid values: L0, L1, L2, L3.");
}

bool IndexPolicy::anyL0() const {
    return std::any_of(levels.begin(), levels.end(), [](LevelPolicy p) { return p == LevelPolicy::L0; });
}
bool IndexPolicy::anyStreaming() const {
    return std::any_of(levels.begin(), levels.end(),
                       [](LevelPolicy p) { return p == LevelPolicy::L2 || p == LevelPolicy::L3; });
}

DeviceMemoryInfo DeviceMemoryInfo::query() {
    DeviceMemoryInfo m;
    hipCheck(hipMemGetInfo(&m.freeMem, &m.totalMem), "hipMemGetInfo");
    return m;
}

// Work buffers the L0 path keeps for level l: survivors (id, offset, D values)
// for every level-l point and the prolongation output (id, D values) for every
// level-(l+1) point.
size_t PolicyScheduler::l0WorkBytes(const IndexData& idx, size_t l) {
    const size_t D = idx.D, cols = idx.pTensors[l]->cols, nnz = idx.pTensors[l]->nnz;
    return cols * (16 + 8 * D) + nnz * (8 + 8 * D);
}

// MemoryPolicy.cu:100-177: budget = free * 0.85 - (coarsest mesh + maps); per
// level, coarsest first: L0 if the L0 cost fits, else L1, else L2 (transient,
// budget unchanged), else L3.
IndexPolicy PolicyScheduler::recommend(const IndexData& idx, size_t freeMem, bool countWorkBuffers) {
    size_t budget = static_cast<size_t>(static_cast<double>(freeMem) * kSafetyFactor);
    const size_t perm = idx.stats.deviceCoarsestMesh + idx.stats.deviceMaps;
    IndexPolicy pol;
    if (budget <= perm) {
        pol.levels.assign(idx.intervals, LevelPolicy::L3);
        return pol;
    }
    budget -= perm;
    pol.levels.resize(idx.intervals);
    for (size_t l = 0; l < idx.intervals; ++l) {
        const size_t nnz = idx.pTensors[l]->nnz, cols = idx.pTensors[l]->cols, D = idx.D;
        const size_t pv = nnz * D * 8, rv = cols * 8, sp = 3 * nnz * 8;
        size_t c0 = pv + rv + sp + pv;
        if (countWorkBuffers) c0 = pv + rv + (cols + 1) * 8 + l0WorkBytes(idx, l);
        const size_t c1 = pv + rv + sp + pv;
        const size_t c2 = 2 * pv + sp;
        if (budget >= c0) {
            pol.levels[l] = LevelPolicy::L0;
            budget -= c0;
        } else if (budget >= c1) {
            pol.levels[l] = LevelPolicy::L1;
            budget -= c1;
        } else if (budget >= c2) {
            pol.levels[l] = LevelPolicy::L2;
        } else {
            pol.levels[l] = LevelPolicy::L3;
        }
    }
    return pol;
}

IndexPolicy PolicyScheduler::recommend(const IndexData& idx) {
    return recommend(idx, DeviceMemoryInfo::query().freeMem, true);
}

void PolicyScheduler::printReport(const IndexData& idx, const IndexPolicy& p, std::ostream& os) {
    os << "=== Per-Level Memory Policy ===\n";
    for (size_t l = 0; l < idx.intervals; ++l) {
        const size_t pv = idx.pTensors[l]->nnz * idx.D * 8;
        os << "  level " << l << "  P vals " << pv << " B  radius " << idx.pTensors[l]->cols * 8 << " B  L0 work "
           << l0WorkBytes(idx, l) << " B  -> " << levelPolicyName(p[l]) << "\n";
    }
}

std::unique_ptr<IQueryStrategy> PolicyScheduler::make(IndexData& idx, const IndexPolicy& policy) {
    idx.activePolicy = policy;
    LevelPolicy rep = LevelPolicy::L3;
    for (auto lp : policy.levels)
        if (static_cast<int>(lp) < static_cast<int>(rep)) rep = lp;
    if (policy.levels.empty()) rep = LevelPolicy::L1;
    switch (rep) {
        case LevelPolicy::L0: return std::make_unique<L0Strategy>(idx);
        case LevelPolicy::L1: return std::make_unique<L1Strategy>(idx);
        case LevelPolicy::L2: return std::make_unique<L2Strategy>(idx);
        default: return std::make_unique<L3Strategy>(idx);
    }
}

std::unique_ptr<IQueryStrategy> PolicyScheduler::makeForLevel(IndexData& idx, size_t l) {
    switch (idx.activePolicy.levels[l]) {
        case LevelPolicy::L0: return std::make_unique<L0Strategy>(idx);
        case LevelPolicy::L1: return std::make_unique<L1Strategy>(idx);
        case LevelPolicy::L2: return std::make_unique<L2Strategy>(idx);
        default: return std::make_unique<L3Strategy>(idx);
    }
}

void freeLevelResult(const LevelResult& r) {
    if (!r.grid) return;
    if (r.onHost) {
        if (r.ownsIds) delete[] r.grid->ids();
        if (r.ownsVals) delete[] r.grid->vals();
    } else {
        if (r.ownsIds && r.grid->ids()) (void)hipFree(r.grid->ids());
        if (r.ownsVals && r.grid->vals()) (void)hipFree(r.grid->vals());
    }
    delete r.grid;
}

// ------------------------------------------------------------------------- L0

void L0Strategy::prepare(std::ostream& os) { prepareAll(idx_, os); }

LevelResult L0Strategy::runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                                 const double* q, size_t K) {
    IndexData& idx = idx_;
    const size_t D = idx.D;
    if (!idx.dLevels[l].resident || idx.work.size() <= l || !idx.work[l].outIds.data())
        throw std::logic_error("L0Strategy: level " + std::to_string(l) + " not prepared for L0");
    if (cur->get_vals_() && !cur->dCount && cur->nnzBound == 0) cur->nnzBound = cur->get_nnz_nums();
    // A device grid without a device count (e.g. an L1/L2 output) or a host grid
    // (L3 output) is first made device-resident with its count.
    DeviceBuffer<uint64_t> tmpIds, tmpCnt;
    DeviceBuffer<double> tmpVals;
    const uint64_t* dCount = cur->dCount;
    const uint64_t* ids = cur->get_ids_();
    const double* vals = cur->get_vals_();
    uint64_t bound = cur->nnzBound ? cur->nnzBound : cur->get_nnz_nums();
    hipPointerAttribute_t attr{};
    const bool onDevice = vals && hipPointerGetAttributes(&attr, vals) == hipSuccess && attr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();
    if (!dCount || !onDevice) {
        const uint64_t n = cur->get_nnz_nums();
        tmpCnt.reset(&n, 1);
        dCount = tmpCnt.data();
        bound = n;
        if (!onDevice) {
            tmpIds.reset(cur->get_ids_(), std::max<size_t>(n, 1));
            tmpVals.reset(cur->get_vals_(), std::max<size_t>(n * D, 1));
            ids = tmpIds.data();
            vals = tmpVals.data();
        }
    }
    const tblaeq_level lv = idx.levelDesc(l);
    auto& W = idx.work[l];
    uint64_t* dSurv = idx.dCounts.data() + 2 * l;
    uint64_t* dOut = dSurv + 1;
    // the query itself was staged in idx.dQuery by QueryEngine (lo|hi or point)
    (void)lo;
    (void)hi;
    (void)q;
    if (t == QueryType::RANGE) {
        tbCheck(tblaeq_range_prune_compact(idx.ctx, &lv, ids, vals, dCount, bound, idx.dQuery.data(),
                                           idx.dQuery.data() + D, W.survIds.data(), W.survVals.data(),
                                           W.survOff.data(), dSurv, dOut),
                "range_prune_compact");
    } else {
        tbCheck(tblaeq_knn_lower_bound_sort_step(idx.ctx, &lv, ids, vals, dCount, bound, idx.dQuery.data(), K,
                                                 W.survIds.data(), W.survVals.data(), W.survOff.data(), dSurv, dOut),
                "knn_lower_bound_sort_step");
    }
    const uint64_t* pIds = W.survIds.data();
    const double* pVals = W.survVals.data();
    const uint64_t* pOff = W.survOff.data();
    const uint64_t* pCount = dSurv;
    if (t != QueryType::RANGE && idx.learnedRadius > 0.0 && l + 1 == idx.intervals && W.radIds.data()) {
        // learned radius: prolong only the survivors within the proposed radius
        // (device copy of the radius staged after the query in idx.dQuery)
        uint64_t* dRad = idx.dCounts.data() + 2 * idx.intervals + 2;
        tbCheck(tblaeq_knn_radius_compact(idx.ctx, &lv, W.survIds.data(), W.survVals.data(), dSurv,
                                          std::min<uint64_t>(bound, lv.cols), idx.dQuery.data(),
                                          idx.dQuery.data() + 2 * D, W.radIds.data(), W.radVals.data(),
                                          W.radOff.data(), dRad, dOut),
                "knn_radius_compact");
        pIds = W.radIds.data();
        pVals = W.radVals.data();
        pOff = W.radOff.data();
        pCount = dRad;
    }
    tbCheck(tblaeq_prolong_refactor(idx.ctx, &lv, pIds, pVals, pOff, pCount, std::min<uint64_t>(bound, lv.cols),
                                    W.outIds.data(), W.outVals.data()),
            "prolong_refactor");
    if (tmpCnt.data()) hipCheck(hipStreamSynchronize(streamOf(idx)), "sync");  // temporaries die here
    auto* g = new SparseGrid(idx.pTensors[l]->rows, D, 0, W.outIds.data(), W.outVals.data());
    g->dCount = dOut;
    g->nnzBound = lv.nnz;
    return {g, false, false, false};
}

// ------------------------------------------------------------------------- L1

void L1Strategy::prepare(std::ostream& os) { prepareAll(idx_, os); }

namespace {

// Moves any grid to the device with a host-known count; returns owned buffers.
struct DeviceGrid {
    DeviceBuffer<uint64_t> ids;
    DeviceBuffer<double> vals;
    const uint64_t* idsPtr = nullptr;
    const double* valsPtr = nullptr;
    size_t n = 0;
};

DeviceGrid toDevice(IndexData& idx, SparseGrid* cur) {
    DeviceGrid g;
    g.n = hostCount(idx, cur);
    hipPointerAttribute_t attr{};
    const bool onDevice = cur->get_vals_() && hipPointerGetAttributes(&attr, cur->get_vals_()) == hipSuccess &&
                          attr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();
    if (onDevice) {
        g.idsPtr = cur->get_ids_();
        g.valsPtr = cur->get_vals_();
    } else {
        g.ids.reset(cur->get_ids_(), std::max<size_t>(g.n, 1));
        g.vals.reset(cur->get_vals_(), std::max<size_t>(g.n * idx.D, 1));
        g.idsPtr = g.ids.data();
        g.valsPtr = g.vals.data();
    }
    return g;
}

// Prune + compact of a device grid into freshly allocated survivor buffers;
// returns survivor and output counts (host).
void pruneDevice(IndexData& idx, size_t l, const tblaeq_level& lv, const DeviceGrid& g, QueryType t, const double* lo,
                 const double* hi, const double* q, size_t K, DeviceBuffer<uint64_t>& sIds,
                 DeviceBuffer<double>& sVals, DeviceBuffer<uint64_t>& sOff, size_t& ns, size_t& nout) {
    const size_t D = idx.D;
    const size_t n = g.n;
    sIds.resize(std::max<size_t>(n, 1));
    sVals.resize(std::max<size_t>(n * D, 1));
    sOff.resize(std::max<size_t>(n, 1));
    DeviceBuffer<uint64_t> cnt(3);
    const uint64_t hn = n;
    h2d(idx, cnt.data() + 2, &hn, 8);
    (void)lo;
    (void)hi;
    (void)q;
    if (t == QueryType::RANGE) {
        tbCheck(tblaeq_range_prune_compact(idx.ctx, &lv, g.idsPtr, g.valsPtr, cnt.data() + 2, n, idx.dQuery.data(),
                                           idx.dQuery.data() + D, sIds.data(), sVals.data(), sOff.data(), cnt.data(),
                                           cnt.data() + 1),
                "range_prune_compact");
    } else {
        tbCheck(tblaeq_knn_lower_bound_sort_step(idx.ctx, &lv, g.idsPtr, g.valsPtr, cnt.data() + 2, n,
                                                 idx.dQuery.data(), K, sIds.data(), sVals.data(), sOff.data(),
                                                 cnt.data(), cnt.data() + 1),
                "knn_lower_bound_sort_step");
    }
    uint64_t hc[2];
    d2h(idx, hc, cnt.data(), 16);
    ns = hc[0];
    nout = hc[1];
    (void)l;
}

}  // namespace

LevelResult L1Strategy::runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                                 const double* q, size_t K) {
    IndexData& idx = idx_;
    if (!idx.dLevels[l].resident) throw std::logic_error("L1Strategy: level values not resident");
    const size_t D = idx.D;
    DeviceGrid g = toDevice(idx, cur);
    const tblaeq_level lv = idx.levelDesc(l);
    DeviceBuffer<uint64_t> sIds, sOff;
    DeviceBuffer<double> sVals;
    size_t ns = 0, nout = 0;
    pruneDevice(idx, l, lv, g, t, lo, hi, q, K, sIds, sVals, sOff, ns, nout);
    if (idx.hostSurv.size() > l) idx.hostSurv[l] = ns;
    uint64_t* oIds = nullptr;
    double* oVals = nullptr;
    hipCheck(hipMalloc(&oIds, std::max<size_t>(nout, 1) * 8), "hipMalloc");
    hipCheck(hipMalloc(&oVals, std::max<size_t>(nout * D, 1) * 8), "hipMalloc");
    DeviceBuffer<uint64_t> dns(1);
    const uint64_t hns = ns;
    h2d(idx, dns.data(), &hns, 8);
    tbCheck(tblaeq_prolong_refactor(idx.ctx, &lv, sIds.data(), sVals.data(), sOff.data(), dns.data(), ns, oIds, oVals),
            "prolong_refactor");
    hipCheck(hipStreamSynchronize(streamOf(idx)), "sync");
    auto* out = new SparseGrid(idx.pTensors[l]->rows, D, nout, oIds, oVals);
    return {out, true, true, false};
}

// ------------------------------------------------------------------------- L2

void L2Strategy::prepare(std::ostream& os) { prepareAll(idx_, os); }

LevelResult L2Strategy::runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                                 const double* q, size_t K) {
    IndexData& idx = idx_;
    const size_t D = idx.D;
    const auto& P = *idx.pTensors[l];
    DeviceGrid g = toDevice(idx, cur);
    DeviceBuffer<double> radius(idx.meshMaxRadius[l].data(), idx.meshMaxRadius[l].size());
    tblaeq_level lv = idx.levelDesc(l);
    lv.radius = radius.data();
    lv.pvals = nullptr;
    DeviceBuffer<uint64_t> sIds, sOff;
    DeviceBuffer<double> sVals;
    size_t ns = 0, nout = 0;
    pruneDevice(idx, l, lv, g, t, lo, hi, q, K, sIds, sVals, sOff, ns, nout);
    if (idx.hostSurv.size() > l) idx.hostSurv[l] = ns;
    radius.free();
    // Pack the survivors' P columns in output order (survivor order, then column order).
    std::vector<uint64_t> hIds(ns);
    d2h(idx, hIds.data(), sIds.data(), ns * 8);
    std::vector<double> packed(std::max<size_t>(nout * D, 1));
    size_t w = 0;
    for (size_t s = 0; s < ns; ++s) {
        const uint64_t c = hIds[s];
        const size_t len = P.colRes[c + 1] - P.colRes[c];
        std::memcpy(&packed[w * D], &P.vals[P.colRes[c] * D], len * D * 8);
        w += len;
    }
    DeviceBuffer<double> dPacked(packed.data(), packed.size());
    uint64_t* oIds = nullptr;
    double* oVals = nullptr;
    hipCheck(hipMalloc(&oIds, std::max<size_t>(nout, 1) * 8), "hipMalloc");
    hipCheck(hipMalloc(&oVals, std::max<size_t>(nout * D, 1) * 8), "hipMalloc");
    DeviceBuffer<uint64_t> dns(1);
    const uint64_t hns = ns;
    h2d(idx, dns.data(), &hns, 8);
    tbCheck(tblaeq_prolong_refactor_packed(idx.ctx, &lv, sIds.data(), sVals.data(), sOff.data(), dns.data(), ns,
                                           dPacked.data(), nullptr, oIds, oVals),
            "prolong_refactor_packed");
    hipCheck(hipStreamSynchronize(streamOf(idx)), "sync");
    auto* out = new SparseGrid(P.rows, D, nout, oIds, oVals);
    return {out, true, true, false};
}

// ------------------------------------------------------------------------- L3

size_t L3Strategy::forcedTileBytes = 0;

void L3Strategy::prepare(std::ostream& os) { prepareAll(idx_, os); }

LevelResult L3Strategy::runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                                 const double* q, size_t K) {
    IndexData& idx = idx_;
    const size_t D = idx.D;
    const auto& P = *idx.pTensors[l];
    const auto& R = idx.meshMaxRadius[l];
    const auto& map = idx.maps[l + 1];
    // Step 0: the input grid on the host.
    const size_t p = hostCount(idx, cur);
    std::vector<uint64_t> hIds(p);
    std::vector<double> hVals(p * D);
    {
        hipPointerAttribute_t attr{};
        const bool onDevice = cur->get_vals_() && hipPointerGetAttributes(&attr, cur->get_vals_()) == hipSuccess &&
                              attr.type == hipMemoryTypeDevice;
        (void)hipGetLastError();
        if (onDevice) {
            d2h(idx, hIds.data(), cur->get_ids_(), p * 8);
            d2h(idx, hVals.data(), cur->get_vals_(), p * D * 8);
        } else if (p) {
            std::memcpy(hIds.data(), cur->get_ids_(), p * 8);
            std::memcpy(hVals.data(), cur->get_vals_(), p * D * 8);
        }
    }
    // Tile size from the device budget (free memory x 0.8, L3Strategy.cu:22-26).
    size_t budget = forcedTileBytes ? forcedTileBytes : tileBudget_;
    if (!budget) budget = static_cast<size_t>(DeviceMemoryInfo::query().freeMem * 0.8);
    const size_t perPoint = 4 * (16 + 8 * D);
    const size_t tile = std::max<size_t>(1, budget / perPoint);

    DeviceBuffer<uint64_t> dColRes(P.colRes.data(), P.colRes.size());
    tblaeq_level lv{};
    lv.D = D;
    lv.cols = P.cols;
    lv.nnz = P.nnz;
    lv.col_res = dColRes.data();
    std::vector<uint8_t> keep(p, 0);
    if (t == QueryType::RANGE) {
        DeviceBuffer<double> dq(2 * D);
        h2d(idx, dq.data(), lo, D * 8);
        h2d(idx, dq.data() + D, hi, D * 8);
        for (size_t b = 0; b < p; b += tile) {
            const size_t n = std::min(tile, p - b);
            std::vector<double> rg(n);
            for (size_t i = 0; i < n; ++i) rg[i] = R[hIds[b + i]];
            DeviceBuffer<uint64_t> tIds(&hIds[b], n), sIds(n), sOff(n), cnt(2);
            DeviceBuffer<double> tVals(&hVals[b * D], n * D), tR(rg.data(), n), sVals(n * D);
            lv.radius = tR.data();
            tbCheck(tblaeq_range_prune_compact_tile(idx.ctx, &lv, tIds.data(), tVals.data(), n, dq.data(),
                                                    dq.data() + D, sIds.data(), sVals.data(), sOff.data(),
                                                    cnt.data(), cnt.data() + 1),
                    "range_prune_compact_tile");
            // mark survivors by position: re-derive from the kept ids in order
            uint64_t hc[2];
            d2h(idx, hc, cnt.data(), 16);
            std::vector<uint64_t> kept(hc[0]);
            d2h(idx, kept.data(), sIds.data(), hc[0] * 8);
            size_t k = 0;
            for (size_t i = 0; i < n && k < kept.size(); ++i)
                if (hIds[b + i] == kept[k]) {
                    keep[b + i] = 1;
                    ++k;
                }
        }
    } else {
        DeviceBuffer<double> dq(q, D), dR(R.data(), R.size());
        std::vector<double> lb(p);
        for (size_t b = 0; b < p; b += tile) {
            const size_t n = std::min(tile, p - b);
            DeviceBuffer<uint64_t> tIds(&hIds[b], n);
            DeviceBuffer<double> tVals(&hVals[b * D], n * D), tLb(n);
            tbCheck(tblaeq_knn_lower_bounds(idx.ctx, D, tIds.data(), tVals.data(), n, dq.data(), dR.data(), tLb.data()),
                    "knn_lower_bounds");
            d2h(idx, &lb[b], tLb.data(), n * 8);
        }
        // STEP (KNN.cu:75-96) over (lb, id) order, as the device selection does.
        std::vector<size_t> ord(p);
        std::iota(ord.begin(), ord.end(), size_t{0});
        std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
            return lb[a] < lb[b] || (lb[a] == lb[b] && hIds[a] < hIds[b]);
        });
        uint64_t count = 0;
        double ub = 0.0;
        size_t i0 = 0;
        for (; i0 < p && count < K; ++i0) {
            const uint64_t id = hIds[ord[i0]];
            count += P.nnzPerCol[id];
            ub = std::max(ub, lb[ord[i0]] + 2.0 * R[id]);
        }
        for (size_t i = 0; i < p; ++i) keep[i] = lb[i] <= ub;
        for (size_t i = 0; i < i0; ++i) keep[ord[i]] = 1;
    }
    // Compact on the host (order preserving) and prolong in survivor chunks.
    std::vector<uint64_t> sId;
    std::vector<double> sVal;
    std::vector<uint64_t> sOff;
    uint64_t nout = 0;
    for (size_t i = 0; i < p; ++i) {
        if (!keep[i]) continue;
        sId.push_back(hIds[i]);
        sVal.insert(sVal.end(), &hVals[i * D], &hVals[i * D] + D);
        sOff.push_back(nout);
        nout += P.colRes[hIds[i] + 1] - P.colRes[hIds[i]];
    }
    const size_t ns = sId.size();
    if (idx.hostSurv.size() > l) idx.hostSurv[l] = ns;
    auto* oIds = new uint64_t[std::max<size_t>(nout, 1)];
    auto* oVals = new double[std::max<size_t>(nout * D, 1)];
    size_t s0 = 0;
    while (s0 < ns) {
        // grow the chunk while its output fits in the tile
        size_t s1 = s0 + 1;
        while (s1 < ns && (((s1 + 1 < ns) ? sOff[s1 + 1] : nout) - sOff[s0]) <= tile) ++s1;
        const uint64_t base = sOff[s0];
        const uint64_t end = s1 < ns ? sOff[s1] : nout;
        const uint64_t m = end - base;
        const uint64_t cn = s1 - s0;
        std::vector<uint64_t> off(cn), pChild(std::max<uint64_t>(m, 1));
        std::vector<double> pVals(std::max<uint64_t>(m * D, 1));
        for (size_t s = s0; s < s1; ++s) {
            const uint64_t c = sId[s], cs = P.colRes[c], len = P.colRes[c + 1] - cs;
            const uint64_t o = sOff[s] - base;
            off[s - s0] = o;
            std::memcpy(&pVals[o * D], &P.vals[cs * D], len * D * 8);
            for (uint64_t j = 0; j < len; ++j) pChild[o + j] = map.empty() ? P.rowIds[cs + j] : map[P.rowIds[cs + j]];
        }
        DeviceBuffer<uint64_t> dId(&sId[s0], cn), dOff(off.data(), cn), dCnt(&cn, 1), dChild(pChild.data(), pChild.size());
        DeviceBuffer<double> dVal(&sVal[s0 * D], cn * D), dP(pVals.data(), pVals.size()), oV(std::max<uint64_t>(m * D, 1));
        DeviceBuffer<uint64_t> oI(std::max<uint64_t>(m, 1));
        tbCheck(tblaeq_prolong_refactor_packed(idx.ctx, &lv, dId.data(), dVal.data(), dOff.data(), dCnt.data(), cn,
                                               dP.data(), dChild.data(), oI.data(), oV.data()),
                "prolong_refactor_packed");
        d2h(idx, oIds + base, oI.data(), m * 8);
        d2h(idx, oVals + base * D, oV.data(), m * D * 8);
        s0 = s1;
    }
    auto* out = new SparseGrid(P.rows, D, nout, oIds, oVals);
    return {out, true, true, true};
}

}  // namespace tblaeq_host
