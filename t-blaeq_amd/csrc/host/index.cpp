// IndexData, on-disk format, synthetic generator, index assembly and text I/O.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <random>
#include <sstream>
#include <thread>

#include "../../../include/tblaeq_build.h"
#include "../../../include/tblaeq_host.hpp"
#include "../ctx.hpp"

namespace tblaeq_host {

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw HipError(std::string(what) + ": " + hipGetErrorString(e),
                       e == hipErrorOutOfMemory ? TBLAEQ_ERR_NOMEM : TBLAEQ_ERR_HIP);
}

void tbCheck(int rc, const char* what) {
    if (rc != TBLAEQ_OK)
        throw HipError(std::string(what) + ": " + tblaeq_status_string(rc) + " (" + tblaeq_last_error() + ")", rc);
}

int statusOfCurrentException(const char* what) noexcept {
    int rc = TBLAEQ_ERR_INTERNAL;
    std::string msg = "unknown exception";
    try {
        throw;
    } catch (const HipError& e) {
        rc = e.status;
        msg = e.what();
    } catch (const IoError& e) {
        rc = TBLAEQ_ERR_IO;
        msg = e.what();
    } catch (const UsageError& e) {
        rc = TBLAEQ_ERR_INVALID;
        msg = e.what();
    } catch (const std::bad_alloc& e) {
        rc = TBLAEQ_ERR_NOMEM;
        msg = std::string("host allocation failed (") + e.what() + ")";
    } catch (const std::logic_error& e) {
        rc = TBLAEQ_ERR_INVALID;
        msg = e.what();
    } catch (const std::exception& e) {
        msg = e.what();
    } catch (...) {
    }
    try {
        tblaeq::set_error(std::string(what) + ": " + msg);
    } catch (...) {
    }
    return rc;
}

// ---------------------------------------------------------------- IndexData

IndexData::~IndexData() {
    dLevels.clear();
    work.clear();
    dCoarsestIds.free();
    dCoarsestVals.free();
    dCounts.free();
    dQuery.free();
    if (hQuery) (void)hipHostFree(hQuery);
    if (ctx) tblaeq_ctx_destroy(ctx);
}

void IndexData::ensureCtx() {
    if (!ctx) tbCheck(tblaeq_ctx_create(0, nullptr, &ctx), "tblaeq_ctx_create");
}

// IndexData.cu:141-194.
void IndexData::computeStats() {
    stats = MemoryStats{};
    for (const auto& t : pTensors) {
        stats.hostPTensors += t->nnz * t->D * 8 + t->cols * 8 + (t->cols + 1) * 8 + t->nnz * 8;
        stats.devicePTensorVals += t->nnz * t->D * 8;
        stats.deviceSpTSpMBufs += 3 * t->rows * 8;
    }
    for (size_t i = 1; i < height; ++i) {
        if (!maps[i].empty()) {
            stats.hostMaps += meshSizes[intervals - i] * 8;
            stats.deviceMaps += meshSizes[intervals - i] * 8;
        }
    }
    for (size_t i = 0; i < intervals; ++i) {
        stats.hostRadius += meshSizes[intervals - i] * 8;
        stats.deviceRadius += meshSizes[intervals - i] * 8;
    }
    stats.hostCoarsestMesh = coarsestIds.size() * 8 + coarsestVals.size() * 8;
    stats.deviceCoarsestMesh = stats.hostCoarsestMesh;
    stats.hostTotal = stats.hostPTensors + stats.hostMaps + stats.hostRadius + stats.hostCoarsestMesh;
    stats.deviceTotal = stats.devicePTensorVals + stats.deviceMaps + stats.deviceRadius + stats.deviceCoarsestMesh;
    stats.grandTotal = stats.hostTotal + stats.deviceTotal;
}

static std::string fmtBytes(size_t b) {
    std::ostringstream o;
    o << std::fixed << std::setprecision(2);
    if (b >= (1ull << 30)) o << b / double(1ull << 30) << " GiB";
    else if (b >= (1ull << 20)) o << b / double(1ull << 20) << " MiB";
    else if (b >= (1ull << 10)) o << b / double(1ull << 10) << " KiB";
    else o << b << " B";
    return o.str();
}

void MemoryStats::print(std::ostream& os) const {
    os << "\nIndex Memory Usage\n[HOST]\n"
       << "    P-tensor arrays : " << fmtBytes(hostPTensors) << "\n"
       << "    Sort maps : " << fmtBytes(hostMaps) << "\n"
       << "    Cluster radius : " << fmtBytes(hostRadius) << "\n"
       << "    Coarsest mesh : " << fmtBytes(hostCoarsestMesh) << "\n"
       << "    HOST TOTAL : " << fmtBytes(hostTotal) << "\n[DEVICE]\n"
       << "    P-tensor vals : " << fmtBytes(devicePTensorVals) << "\n"
       << "    Sort maps : " << fmtBytes(deviceMaps) << "\n"
       << "    Cluster radius : " << fmtBytes(deviceRadius) << "\n"
       << "    Coarsest mesh : " << fmtBytes(deviceCoarsestMesh) << "\n"
       << "    DEVICE TOTAL : " << fmtBytes(deviceTotal) << "\nGRAND TOTAL : " << fmtBytes(grandTotal) << "\n\n";
}

// Permanent device data: CSC column structure, child ids (= map[l+1] composed
// with rowIds, folding the refactor into the index), coarsest mesh.
// IndexData.cu:52-74 uploads maps + coarsest mesh; the column structure is new
// (the reference keeps it on the host and expands index arrays per query).
void IndexData::uploadPermanentData() {
    if (permanentDataOnDevice) return;
    ensureCtx();
    dLevels.resize(intervals);
    for (size_t l = 0; l < intervals; ++l) {
        const auto& P = *pTensors[l];
        auto& L = dLevels[l];
        L.colRes.reset(P.colRes.data(), P.colRes.size());
        L.nnzPerCol.reset(P.nnzPerCol.data(), P.nnzPerCol.size());
        const auto& map = maps[l + 1];
        std::vector<uint64_t> child(P.nnz);
        for (size_t j = 0; j < P.nnz; ++j) {
            const uint64_t r = P.rowIds[j];
            child[j] = map.empty() ? r : map[r];
        }
        L.childIds.reset(child.data(), child.size());
        L.emptyCols = static_cast<uint64_t>(std::count(P.nnzPerCol.begin(), P.nnzPerCol.end(), uint64_t{0}));
    }
    dCoarsestIds.reset(coarsestIds.data(), coarsestIds.size());
    dCoarsestVals.reset(coarsestVals.data(), coarsestVals.size());
    dCounts.resize(2 * intervals + 4);
    std::vector<uint64_t> c(2 * intervals + 4, 0);
    c[2 * intervals] = coarsestIds.size();
    hipCheck(hipMemcpy(dCounts.data(), c.data(), c.size() * 8, hipMemcpyHostToDevice), "counts");
    dQuery.resize(2 * D + 1);
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&hQuery), (2 * D + 1) * sizeof(double), hipHostMallocDefault),
             "hipHostMalloc");
    permanentDataOnDevice = true;
}

void IndexData::uploadLevelValues(size_t l) {
    auto& L = dLevels[l];
    if (L.resident) return;
    L.pvals.reset(pTensors[l]->vals.data(), pTensors[l]->vals.size());
    L.radius.reset(meshMaxRadius[l].data(), meshMaxRadius[l].size());
    L.resident = true;
}

void IndexData::allocLevelWork(size_t l) {
    if (work.size() != intervals) work.resize(intervals);
    auto& W = work[l];
    const size_t cols = pTensors[l]->cols, nnz = pTensors[l]->nnz;
    if (W.survIds.size() == std::max<size_t>(cols, 1)) return;
    W.survIds.resize(std::max<size_t>(cols, 1));
    W.survOff.resize(std::max<size_t>(cols, 1));
    W.survVals.resize(std::max<size_t>(cols * D, 1));
    W.outIds.resize(std::max<size_t>(nnz, 1));
    W.outVals.resize(std::max<size_t>(nnz * D, 1));
    if (l + 1 == intervals) {
        W.radIds.resize(std::max<size_t>(cols, 1));
        W.radOff.resize(std::max<size_t>(cols, 1));
        W.radVals.resize(std::max<size_t>(cols * D, 1));
    }
}

tblaeq_level IndexData::levelDesc(size_t l) const {
    const auto& L = dLevels[l];
    tblaeq_level v{};
    v.D = D;
    v.cols = pTensors[l]->cols;
    v.nnz = pTensors[l]->nnz;
    v.col_res = L.colRes.data();
    v.nnz_per_col = L.nnzPerCol.data();
    v.pvals = L.pvals.data();
    v.radius = L.radius.data();
    v.child_ids = L.childIds.data();
    v.empty_cols = L.emptyCols;
    return v;
}

// ---------------------------------------------------------- IndexSerializer

namespace {
template <class T>
void wr(std::ofstream& f, const T* p, size_t n) {
    f.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
}
template <class T>
void rd(std::ifstream& f, T* p, size_t n, const std::string& path) {
    f.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
    if (!f) throw IoError("IndexSerializer: short read in " + path);
}
}  // namespace

void IndexSerializer::save(const IndexData& d, const std::string& dir) {
    std::filesystem::create_directories(dir);
    const uint8_t aos = d.isAosArch ? 1 : 0;
    {
        const std::string p = dir + "/metadata.bin";
        std::ofstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot write: " + p);
        const uint64_t meta[4] = {d.D, d.N, d.height, d.intervals};
        wr(f, meta, 4);
        wr(f, &aos, 1);
        wr(f, d.ratios.data(), d.intervals);
        const uint64_t msz = d.meshSizes.size();
        wr(f, &msz, 1);
        wr(f, d.meshSizes.data(), msz);
    }
    for (size_t i = 0; i < d.intervals; ++i) {
        const std::string p = dir + "/P_Tensor_" + std::to_string(i) + ".bin";
        std::ofstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot write: " + p);
        const auto& t = *d.pTensors[i];
        const uint64_t hdr[4] = {t.D, t.rows, t.cols, t.nnz};
        const uint8_t a = t.isAos ? 1 : 0;
        wr(f, hdr, 4);
        wr(f, &a, 1);
        wr(f, t.rowIds.data(), t.nnz);
        wr(f, t.nnzPerCol.data(), t.cols);
        wr(f, t.colRes.data(), t.cols + 1);
        wr(f, t.vals.data(), t.nnz * t.D);
    }
    for (size_t i = 0; i < d.intervals; ++i) {
        const std::string p = dir + "/MaxRadius_" + std::to_string(i) + ".bin";
        std::ofstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot write: " + p);
        const uint64_t n = d.meshSizes[d.intervals - i];
        wr(f, &n, 1);
        wr(f, d.meshMaxRadius[i].data(), n);
    }
    for (size_t i = 1; i < d.height; ++i) {
        if (d.maps[i].empty()) continue;
        const std::string p = dir + "/Map_" + std::to_string(i) + ".bin";
        std::ofstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot write: " + p);
        const uint64_t n = d.meshSizes[d.intervals - i];
        wr(f, &n, 1);
        wr(f, d.maps[i].data(), n);
    }
    {
        const std::string p = dir + "/Coreast_Mesh.bin";
        std::ofstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot write: " + p);
        const uint64_t hdr[3] = {d.coarsestRows, d.D, d.coarsestIds.size()};
        wr(f, hdr, 3);
        wr(f, &aos, 1);
        wr(f, d.coarsestIds.data(), d.coarsestIds.size());
        wr(f, d.coarsestVals.data(), d.coarsestVals.size());
    }
}

IndexData* IndexSerializer::load(const std::string& dir) {
    auto d = std::make_unique<IndexData>();
    {
        const std::string p = dir + "/metadata.bin";
        std::ifstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot read: " + p);
        uint64_t meta[4];
        uint8_t aos;
        rd(f, meta, 4, p);
        rd(f, &aos, 1, p);
        d->D = meta[0];
        d->N = meta[1];
        d->height = meta[2];
        d->intervals = meta[3];
        d->isAosArch = aos != 0;
        if (d->height != d->intervals + 1 || d->D == 0 || d->intervals == 0)
            throw IoError("IndexSerializer: inconsistent metadata in " + p);
        d->ratios.resize(d->intervals);
        rd(f, d->ratios.data(), d->intervals, p);
        uint64_t msz;
        rd(f, &msz, 1, p);
        if (msz != d->height) throw IoError("IndexSerializer: mesh size count mismatch in " + p);
        d->meshSizes.resize(msz);
        rd(f, d->meshSizes.data(), msz, p);
    }
    d->pTensors.resize(d->intervals);
    d->meshMaxRadius.resize(d->intervals);
    d->maps.resize(d->height);
    for (size_t i = 0; i < d->intervals; ++i) {
        const std::string p = dir + "/P_Tensor_" + std::to_string(i) + ".bin";
        std::ifstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot read: " + p);
        auto t = std::make_unique<SparseTensorCscFormat>();
        uint64_t hdr[4];
        uint8_t a;
        rd(f, hdr, 4, p);
        rd(f, &a, 1, p);
        t->D = hdr[0];
        t->rows = hdr[1];
        t->cols = hdr[2];
        t->nnz = hdr[3];
        t->isAos = a != 0;
        t->rowIds.resize(t->nnz);
        t->nnzPerCol.resize(t->cols);
        t->colRes.resize(t->cols + 1);
        t->vals.resize(t->nnz * t->D);
        rd(f, t->rowIds.data(), t->nnz, p);
        rd(f, t->nnzPerCol.data(), t->cols, p);
        rd(f, t->colRes.data(), t->cols + 1, p);
        rd(f, t->vals.data(), t->nnz * t->D, p);
        if (t->colRes[t->cols] != t->nnz) throw IoError("IndexSerializer: CSC column pointers in " + p);
        d->pTensors[i] = std::move(t);
    }
    for (size_t i = 0; i < d->intervals; ++i) {
        const std::string p = dir + "/MaxRadius_" + std::to_string(i) + ".bin";
        std::ifstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot read: " + p);
        uint64_t n;
        rd(f, &n, 1, p);
        if (n != d->meshSizes[d->intervals - i]) throw IoError("IndexSerializer: radius count in " + p);
        d->meshMaxRadius[i].resize(n);
        rd(f, d->meshMaxRadius[i].data(), n, p);
    }
    for (size_t i = 1; i < d->height; ++i) {
        const std::string p = dir + "/Map_" + std::to_string(i) + ".bin";
        if (!std::filesystem::exists(p)) continue;
        std::ifstream f(p, std::ios::binary);
        uint64_t n;
        rd(f, &n, 1, p);
        if (n != d->meshSizes[d->intervals - i]) throw IoError("IndexSerializer: map size in " + p);
        d->maps[i].resize(n);
        rd(f, d->maps[i].data(), n, p);
    }
    {
        const std::string p = dir + "/Coreast_Mesh.bin";
        std::ifstream f(p, std::ios::binary);
        if (!f) throw IoError("Cannot read: " + p);
        uint64_t hdr[3];
        uint8_t a;
        rd(f, hdr, 3, p);
        rd(f, &a, 1, p);
        d->coarsestRows = hdr[0];
        d->coarsestIds.resize(hdr[2]);
        d->coarsestVals.resize(hdr[2] * hdr[1]);
        rd(f, d->coarsestIds.data(), hdr[2], p);
        rd(f, d->coarsestVals.data(), hdr[2] * hdr[1], p);
    }
    return d.release();
}

// ------------------------------------------------------------- IndexBuilder

const std::vector<size_t> IndexBuilder::kDefaultRatios = {100, 50, 20};

size_t IndexBuilder::computeCentroidCount(size_t n, size_t ratio) {
    if (ratio == 0 || n / ratio == 0) return 1;
    return n / ratio;
}

// Setup.cuh:87-171 and func.hpp:124-137,202-213 over flat layers.
IndexData* IndexBuilder::assemble(size_t N, size_t D, const std::vector<size_t>& ratios,
                                  std::vector<std::vector<double>>& layers,
                                  const std::vector<std::vector<uint64_t>>& labels) {
    auto idx = std::make_unique<IndexData>();
    idx->D = D;
    idx->N = N;
    idx->height = ratios.size() + 1;
    idx->intervals = ratios.size();
    idx->ratios = ratios;
    idx->pTensors.resize(idx->intervals);
    idx->meshMaxRadius.resize(idx->intervals);
    idx->maps.resize(idx->height);
    for (size_t i = 0; i < idx->intervals; ++i) {
        const size_t fineIdx = idx->intervals - i, pIdx = fineIdx - 1;
        const auto& fine = layers[i];
        const auto& coarse = layers[i + 1];
        const auto& lab = labels[i];
        const size_t nF = lab.size(), nC = coarse.size() / D;
        idx->meshSizes.push_back(nF);
        auto P = std::make_unique<SparseTensorCscFormat>();
        P->D = D;
        P->rows = nF;
        P->cols = nC;
        P->nnz = nF;
        P->nnzPerCol.assign(nC, 0);
        for (uint64_t c : lab) ++P->nnzPerCol[c];
        P->colRes.assign(nC + 1, 0);
        for (size_t c = 0; c < nC; ++c) P->colRes[c + 1] = P->colRes[c] + P->nnzPerCol[c];
        P->rowIds.resize(nF);
        std::iota(P->rowIds.begin(), P->rowIds.end(), uint64_t{0});
        // sort-to-original map: argsort of labels with std::sort on the index
        // vector (the permutation among equal labels is the library's).
        std::vector<size_t> order(nF);
        std::iota(order.begin(), order.end(), size_t{0});
        std::sort(order.begin(), order.end(), [&lab](size_t a, size_t b) { return lab[a] < lab[b]; });
        auto& map = idx->maps[fineIdx];
        map.assign(order.begin(), order.end());
        P->vals.resize(nF * D);
        auto& rad = idx->meshMaxRadius[pIdx];
        rad.assign(nC, 0.0);
        for (size_t c = 0; c < nC; ++c) {
            const double* cv = &coarse[c * D];
            double best = 0.0;
            for (uint64_t pos = P->colRes[c]; pos < P->colRes[c + 1]; ++pos) {
                const double* fv = &fine[map[pos] * D];
                double s = 0.0;
                for (size_t d = 0; d < D; ++d) {
                    P->vals[pos * D + d] = fv[d] / cv[d];
                    const double df = cv[d] - fv[d];
                    const double sq = df * df;
                    s = s + sq;
                }
                best = std::max(best, std::sqrt(s));
            }
            rad[c] = best;
        }
        idx->pTensors[pIdx] = std::move(P);
        if (i + 2 == idx->height) {
            idx->meshSizes.push_back(nC);
            idx->coarsestRows = nC;
            idx->coarsestIds.resize(nC);
            std::iota(idx->coarsestIds.begin(), idx->coarsestIds.end(), uint64_t{0});
            idx->coarsestVals = coarse;
        }
    }
    return idx.release();
}

// RandomKmeans.cuh:176-250: coarsest layer uniform in [valMin, valMax]^D, each
// finer layer assigns child ci to parent ci % nParents and draws it from a
// normal of sigma = ((valMax-valMin)/nParents^(1/D))/sigmaDivisor, clamped.
IndexData* IndexBuilder::buildRandom(size_t N, size_t D, double valMin, double valMax, bool isInt,
                                     const std::string& name, size_t height, const std::vector<size_t>& ratios,
                                     uint64_t seed, double sigmaDivisor) {
    if (valMin <= 0.0) throw std::invalid_argument("RandomKmeans: valMin must be > 0");
    if (valMin >= valMax) throw std::invalid_argument("RandomKmeans: valMin must be < valMax");
    if (sigmaDivisor <= 0.0) throw std::invalid_argument("RandomKmeans: sigmaDivisor must be > 0");
    if (height < 2 || ratios.size() != height - 1) throw std::invalid_argument("buildRandom: bad height/ratios");
    const size_t L = ratios.size();
    std::vector<size_t> sizes(L + 1);
    sizes[0] = N;
    for (size_t i = 0; i < L; ++i) sizes[i + 1] = std::max<size_t>(1, sizes[i] / ratios[i]);
    std::vector<std::vector<double>> layers(L + 1);
    std::vector<std::vector<uint64_t>> labels(L);
    std::mt19937_64 rng(seed);
    auto round = [isInt](double v) { return isInt ? std::round(v) : v; };
    std::uniform_real_distribution<double> uni(valMin, valMax);
    layers[L].resize(sizes[L] * D);
    for (auto& v : layers[L]) v = round(uni(rng));
    for (size_t lvl = L; lvl > 0; --lvl) {
        const size_t nP = sizes[lvl], nC = sizes[lvl - 1];
        const double spacing = (valMax - valMin) / std::pow(static_cast<double>(nP), 1.0 / static_cast<double>(D));
        std::normal_distribution<double> noise(0.0, spacing / sigmaDivisor);
        const auto& par = layers[lvl];
        auto& ch = layers[lvl - 1];
        auto& lab = labels[lvl - 1];
        ch.resize(nC * D);
        lab.resize(nC);
        for (size_t ci = 0; ci < nC; ++ci) {
            const size_t p = ci % nP;
            lab[ci] = p;
            for (size_t d = 0; d < D; ++d) ch[ci * D + d] = round(std::clamp(par[p * D + d] + noise(rng), valMin, valMax));
        }
    }
    IndexData* idx = assemble(N, D, ratios, layers, labels);
    idx->datasetName = name;
    idx->finestData = std::move(layers[0]);
    return idx;
}

// ------------------------------------------------ parallel synthetic hierarchy

namespace {

// SplitMix64 (Steele, Lea and Flood, 2014): decorrelated seeds for the
// per-chunk generators of the parallel builder.
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Runs f(begin, end, t) over [0, n) split into `threads` contiguous ranges.
template <class F>
void parallelRanges(size_t n, unsigned threads, F f) {
    threads = static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(threads, n)));
    std::vector<std::thread> pool;
    const size_t per = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        const size_t b = std::min(n, t * per), e = std::min(n, b + per);
        pool.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto& th : pool) th.join();
}

constexpr size_t kSynthChunk = size_t{1} << 16;  // children per generator stream

}  // namespace

// The generator's hierarchy shape (coarsest layer uniform in [valMin,
// valMax]^D, child ci of a layer under parent ci % nParents, drawn from a
// normal of sigma = ((valMax-valMin)/nParents^(1/D))/sigmaDivisor, clamped,
// rounded for integer data) built on `threads` host threads: every chunk of
// kSynthChunk values or children draws from its own mt19937_64 seeded by
// SplitMix64 of (seed, level, chunk), so the output does not depend on the
// thread count.  No reference counterpart: it is not the reference's
// single-stream sequence (RandomKmeans.cuh:176-250), and each column's rows are
// in ascending point order (a stable argsort of the labels).
IndexData* IndexBuilder::buildSynthetic(size_t N, size_t D, double valMin, double valMax, bool isInt,
                                        const std::string& name, size_t height, const std::vector<size_t>& ratios,
                                        uint64_t seed, double sigmaDivisor, unsigned threads) {
    if (!(valMin > 0.0) || !(valMin < valMax)) throw std::invalid_argument("synthetic: need 0 < valMin < valMax");
    if (!(sigmaDivisor > 0.0)) throw std::invalid_argument("synthetic: sigmaDivisor must be > 0");
    if (height < 2 || ratios.size() != height - 1) throw std::invalid_argument("synthetic: bad height/ratios");
    if (N == 0 || D == 0) throw std::invalid_argument("synthetic: N and D must be positive");
    if (threads == 0) threads = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    const size_t L = ratios.size();
    std::vector<size_t> sizes(L + 1);
    sizes[0] = N;
    for (size_t i = 0; i < L; ++i) sizes[i + 1] = std::max<size_t>(1, sizes[i] / ratios[i]);
    std::vector<std::vector<double>> layers(L + 1);
    std::vector<std::vector<uint64_t>> labels(L);
    auto rngFor = [seed](uint64_t level, uint64_t chunk) {
        return std::mt19937_64(splitmix64(seed ^ splitmix64((level << 40) + chunk + 1)));
    };
    auto round = [isInt](double v) { return isInt ? std::round(v) : v; };
    {
        auto& top = layers[L];
        top.resize(sizes[L] * D);
        const size_t chunks = (top.size() + kSynthChunk - 1) / kSynthChunk;
        parallelRanges(chunks, threads, [&](size_t b, size_t e, unsigned) {
            for (size_t c = b; c < e; ++c) {
                std::mt19937_64 rng = rngFor(L, c);
                std::uniform_real_distribution<double> uni(valMin, valMax);
                const size_t hi = std::min(top.size(), (c + 1) * kSynthChunk);
                for (size_t i = c * kSynthChunk; i < hi; ++i) top[i] = round(uni(rng));
            }
        });
    }
    for (size_t lvl = L; lvl > 0; --lvl) {
        const size_t nP = sizes[lvl], nC = sizes[lvl - 1];
        const double sigma =
            (valMax - valMin) / std::pow(static_cast<double>(nP), 1.0 / static_cast<double>(D)) / sigmaDivisor;
        const auto& par = layers[lvl];
        auto& ch = layers[lvl - 1];
        auto& lab = labels[lvl - 1];
        ch.resize(nC * D);
        lab.resize(nC);
        const size_t chunks = (nC + kSynthChunk - 1) / kSynthChunk;
        parallelRanges(chunks, threads, [&](size_t b, size_t e, unsigned) {
            for (size_t c = b; c < e; ++c) {
                std::mt19937_64 rng = rngFor(lvl - 1, c);
                std::normal_distribution<double> noise(0.0, sigma);
                const size_t hi = std::min(nC, (c + 1) * kSynthChunk);
                for (size_t ci = c * kSynthChunk; ci < hi; ++ci) {
                    const size_t p = ci % nP;
                    lab[ci] = p;
                    for (size_t d = 0; d < D; ++d)
                        ch[ci * D + d] = round(std::clamp(par[p * D + d] + noise(rng), valMin, valMax));
                }
            }
        });
    }
    IndexData* idx = assembleStable(N, D, ratios, layers, labels, threads);
    idx->datasetName = name;
    idx->finestData = std::move(layers[0]);
    return idx;
}

// assemble() on `threads` host threads, with each column's rows in ascending
// point order (a parallel stable counting sort of the labels).
IndexData* IndexBuilder::assembleStable(size_t N, size_t D, const std::vector<size_t>& ratios,
                                        std::vector<std::vector<double>>& layers,
                                        const std::vector<std::vector<uint64_t>>& labels, unsigned threads) {
    auto idx = std::make_unique<IndexData>();
    idx->D = D;
    idx->N = N;
    idx->height = ratios.size() + 1;
    idx->intervals = ratios.size();
    idx->ratios = ratios;
    idx->pTensors.resize(idx->intervals);
    idx->meshMaxRadius.resize(idx->intervals);
    idx->maps.resize(idx->height);
    for (size_t i = 0; i < idx->intervals; ++i) {
        const size_t fineIdx = idx->intervals - i, pIdx = fineIdx - 1;
        const auto& fine = layers[i];
        const auto& coarse = layers[i + 1];
        const auto& lab = labels[i];
        const size_t nF = lab.size(), nC = coarse.size() / D;
        idx->meshSizes.push_back(nF);
        auto P = std::make_unique<SparseTensorCscFormat>();
        P->D = D;
        P->rows = nF;
        P->cols = nC;
        P->nnz = nF;
        // per-thread column counts over contiguous row ranges, then each
        // thread's first slot per column: a stable scatter
        const unsigned T = static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(threads, nF)));
        std::vector<std::vector<uint64_t>> cnt(T, std::vector<uint64_t>(nC, 0));
        parallelRanges(nF, T, [&](size_t b, size_t e, unsigned t) {
            for (size_t r = b; r < e; ++r) ++cnt[t][lab[r]];  // labels come from the builder: < nC
        });
        P->nnzPerCol.assign(nC, 0);
        P->colRes.assign(nC + 1, 0);
        for (size_t c = 0; c < nC; ++c) {
            uint64_t run = P->colRes[c];
            for (unsigned t = 0; t < T; ++t) {
                const uint64_t k = cnt[t][c];
                cnt[t][c] = run;
                run += k;
            }
            P->nnzPerCol[c] = run - P->colRes[c];
            P->colRes[c + 1] = run;
        }
        auto& map = idx->maps[fineIdx];
        map.resize(nF);
        P->rowIds.resize(nF);
        parallelRanges(nF, T, [&](size_t b, size_t e, unsigned t) {
            for (size_t r = b; r < e; ++r) {
                P->rowIds[r] = r;
                map[cnt[t][lab[r]]++] = r;
            }
        });
        cnt.clear();
        P->vals.resize(nF * D);
        auto& rad = idx->meshMaxRadius[pIdx];
        rad.assign(nC, 0.0);
        parallelRanges(nC, threads, [&](size_t b, size_t e, unsigned) {
            for (size_t c = b; c < e; ++c) {
                const double* cv = &coarse[c * D];
                double best = 0.0;
                for (uint64_t pos = P->colRes[c]; pos < P->colRes[c + 1]; ++pos) {
                    const double* fv = &fine[map[pos] * D];
                    double s = 0.0;
                    for (size_t d = 0; d < D; ++d) {
                        P->vals[pos * D + d] = fv[d] / cv[d];
                        const double df = cv[d] - fv[d];
                        s = s + df * df;
                    }
                    best = std::max(best, std::sqrt(s));
                }
                rad[c] = best;
            }
        });
        idx->pTensors[pIdx] = std::move(P);
        if (i + 2 == idx->height) {
            idx->meshSizes.push_back(nC);
            idx->coarsestRows = nC;
            idx->coarsestIds.resize(nC);
            std::iota(idx->coarsestIds.begin(), idx->coarsestIds.end(), uint64_t{0});
            idx->coarsestVals = coarse;
        }
    }
    return idx.release();
}

// IndexBuilder.cu:44-151 on the GPU (the reference's cuVS path when
// forceUseCPU is false, CUDAKmeans.cu:7-195): per level tblaeq_kmeans
// (k-means++ from mt19937_64(42), Lloyd for at most 7 iterations, the CPU
// build's labels and centroids) and tblaeq_build_level (CSC P-tensor, map,
// radius; rows of a column in point order).
IndexData* IndexBuilder::buildGpu(const double* data, size_t N, size_t D, const std::string& name, size_t height,
                                  const std::vector<size_t>& ratios, const tblaeq_kmeans_options* optIn,
                                  const float* data32) {
    if (height < 2 || ratios.size() != height - 1) throw std::invalid_argument("build: bad height/ratios");
    if (!data && !data32) throw std::invalid_argument("build: no data");
    tblaeq_kmeans_options opt;
    tblaeq_kmeans_options_init(&opt);
    if (optIn) opt = *optIn;
    const size_t L = ratios.size();
    auto idx = std::make_unique<IndexData>();
    idx->D = D;
    idx->N = N;
    idx->height = height;
    idx->intervals = L;
    idx->ratios = ratios;
    idx->datasetName = name;
    idx->pTensors.resize(L);
    idx->meshMaxRadius.resize(L);
    idx->maps.resize(height);
    idx->ensureCtx();
    tblaeq_ctx* ctx = idx->ctx;
    DeviceBuffer<double> dx;
    if (data) {
        dx.reset(data, N * D);
    } else {
        DeviceBuffer<float> d32(data32, N * D);
        dx.resize(N * D);
        tbCheck(tblaeq_convert_f32_f64(ctx, d32.data(), N * D, dx.data()), "tblaeq_convert_f32_f64");
        hipCheck(hipStreamSynchronize(static_cast<hipStream_t>(tblaeq_ctx_stream(ctx))), "sync");
    }
    size_t n = N;
    for (size_t i = 0; i < L; ++i) {
        const size_t fineIdx = L - i, pIdx = fineIdx - 1;
        const size_t k = computeCentroidCount(n, ratios[i]);
        DeviceBuffer<double> dc;
        DeviceBuffer<uint64_t> dlab, dColRes, dNnz, dMap;
        DeviceBuffer<double> dVals, dRad;
        dc.resize(k * D);
        dlab.resize(n);
        dColRes.resize(k + 1);
        dNnz.resize(k);
        dMap.resize(n);
        dVals.resize(n * D);
        dRad.resize(k);
        uint64_t st[6] = {0, 0, 0, 0, 0, 0};
        tblaeq::progress("GPU build: level %zu of %zu, %zu points into %zu clusters", i + 1, L, n, k);
        tbCheck(tblaeq_kmeans_ex(ctx, dx.data(), n, D, k, &opt, dc.data(), dlab.data(), st), "tblaeq_kmeans");
        tbCheck(tblaeq_build_level(ctx, dx.data(), n, D, dc.data(), k, dlab.data(), dColRes.data(), dNnz.data(),
                                   dMap.data(), dVals.data(), dRad.data()),
                "tblaeq_build_level");
        idx->buildStats.insert(idx->buildStats.end(), st, st + 6);
        idx->meshSizes.push_back(n);
        auto P = std::make_unique<SparseTensorCscFormat>();
        P->D = D;
        P->rows = n;
        P->cols = k;
        P->nnz = n;
        P->rowIds.resize(n);
        std::iota(P->rowIds.begin(), P->rowIds.end(), uint64_t{0});
        P->colRes.resize(k + 1);
        P->nnzPerCol.resize(k);
        P->vals.resize(n * D);
        idx->maps[fineIdx].resize(n);
        idx->meshMaxRadius[pIdx].resize(k);
        dColRes.download(P->colRes.data(), k + 1);
        dNnz.download(P->nnzPerCol.data(), k);
        dVals.download(P->vals.data(), n * D);
        dMap.download(idx->maps[fineIdx].data(), n);
        dRad.download(idx->meshMaxRadius[pIdx].data(), k);
        idx->pTensors[pIdx] = std::move(P);
        if (i + 1 == L) {
            idx->meshSizes.push_back(k);
            idx->coarsestRows = k;
            idx->coarsestIds.resize(k);
            std::iota(idx->coarsestIds.begin(), idx->coarsestIds.end(), uint64_t{0});
            idx->coarsestVals.resize(k * D);
            dc.download(idx->coarsestVals.data(), k * D);
        }
        dx = std::move(dc);
        n = k;
    }
    return idx.release();
}

// IndexBuilder.cu:44-151 with Lloyd k-means + k-means++ (CPUKmeans.cuh:67-205)
// on the host, or on the GPU (buildGpu) when forceUseCPU is false.
IndexData* IndexBuilder::build(const double* data, size_t N, size_t D, const std::string& name, bool forceUseCPU,
                               size_t height, const std::vector<size_t>& ratios, const tblaeq_kmeans_options* optIn) {
    if (!forceUseCPU) return buildGpu(data, N, D, name, height, ratios, optIn);
    if (height < 2 || ratios.size() != height - 1) throw std::invalid_argument("build: bad height/ratios");
    tblaeq_kmeans_options opt;
    tblaeq_kmeans_options_init(&opt);
    if (optIn) opt = *optIn;
    if (opt.seeding != 0) throw std::invalid_argument("build: k-means|| seeding runs on the GPU only");
    const size_t maxIter = opt.max_iter;
    const size_t L = ratios.size();
    std::vector<uint64_t> stats;
    std::vector<std::vector<double>> layers(L + 1);
    std::vector<std::vector<uint64_t>> labels(L);
    layers[0].assign(data, data + N * D);
    for (size_t i = 0; i < L; ++i) {
        const auto& x = layers[i];
        const size_t n = x.size() / D;
        const size_t k = computeCentroidCount(n, ratios[i]);
        auto sq = [D](const double* a, const double* b) {
            double s = 0.0;
            for (size_t d = 0; d < D; ++d) {
                const double df = a[d] - b[d];
                const double p = df * df;
                s = s + p;
            }
            return s;
        };
        std::vector<double> c;
        c.reserve(k * D);
        std::mt19937_64 rng(opt.seed);
        std::uniform_int_distribution<size_t> uni(0, n - 1);
        const size_t first = uni(rng);
        c.insert(c.end(), &x[first * D], &x[first * D] + D);
        std::vector<double> minD(n, std::numeric_limits<double>::max());
        for (size_t cc = 1; cc < k; ++cc) {
            const double* last = &c[(cc - 1) * D];
            double total = 0.0;
            for (size_t j = 0; j < n; ++j) {
                const double dd = sq(&x[j * D], last);
                if (dd < minD[j]) minD[j] = dd;
                total += minD[j];
            }
            std::uniform_real_distribution<double> dist(0.0, total);
            const double thr = dist(rng);
            double cum = 0.0;
            size_t chosen = n - 1;
            for (size_t j = 0; j < n; ++j) {
                cum += minD[j];
                if (cum >= thr) {
                    chosen = j;
                    break;
                }
            }
            c.insert(c.end(), &x[chosen * D], &x[chosen * D] + D);
        }
        auto& lab = labels[i];
        lab.assign(n, 0);
        size_t iters = 0;
        for (size_t it = 0; it < maxIter; ++it) {
            ++iters;
            bool changed = false;
            for (size_t j = 0; j < n; ++j) {
                double best = std::numeric_limits<double>::max();
                uint64_t bc = 0;
                for (size_t q = 0; q < k; ++q) {
                    const double dd = sq(&x[j * D], &c[q * D]);
                    if (dd < best) {
                        best = dd;
                        bc = q;
                    }
                }
                if (lab[j] != bc) {
                    lab[j] = bc;
                    changed = true;
                }
            }
            std::vector<double> sums(k * D, 0.0);
            std::vector<size_t> cnt(k, 0);
            for (size_t j = 0; j < n; ++j) {
                ++cnt[lab[j]];
                for (size_t d = 0; d < D; ++d) sums[lab[j] * D + d] += x[j * D + d];
            }
            for (size_t q = 0; q < k; ++q)
                if (cnt[q])
                    for (size_t d = 0; d < D; ++d) c[q * D + d] = sums[q * D + d] / static_cast<double>(cnt[q]);
            if (!changed) break;
        }
        // inertia: squared distances to the final centroids, summed in point order
        double inertia = 0.0;
        for (size_t j = 0; j < n; ++j) inertia += sq(&x[j * D], &c[lab[j] * D]);
        uint64_t bits = 0;
        std::memcpy(&bits, &inertia, sizeof bits);
        const uint64_t st[6] = {iters, 0, 0, 0, bits, 0};
        stats.insert(stats.end(), st, st + 6);
        layers[i + 1] = std::move(c);
    }
    IndexData* idx = assemble(N, D, ratios, layers, labels);
    idx->datasetName = name;
    idx->buildStats = std::move(stats);
    return idx;
}

// ------------------------------------------------------------------ text I/O

// Reads `count dim` and then `count` rows of numbers from a text file (the
// dataset format of File.cu:8-50; range query files hold 2*dim numbers per
// row, lower corner then upper corner, File.cu:139-216).  Extra numbers on a
// row are ignored and rows past `count` are not read, as in the reference; a
// short row, a missing row, a bad header or an unreadable file is an IoError.
static std::vector<double> readRows(const std::string& path, const char* what, int64_t* count, int64_t* dim,
                                    int64_t perDim) {
    std::ifstream f(path);
    if (!f) throw IoError(std::string(what) + ": '" + path + "' could not be opened");
    std::string line;
    if (!std::getline(f, line)) throw IoError(std::string(what) + ": '" + path + "' has no header line");
    {
        std::istringstream iss(line);
        if (!(iss >> *count >> *dim) || *count <= 0 || *dim <= 0)
            throw IoError(std::string(what) + ": header of '" + path + "' is not two positive integers");
    }
    const int64_t cols = *dim * perDim;
    std::vector<double> out(static_cast<size_t>(*count * cols));
    int64_t row = 0;
    for (; row < *count && std::getline(f, line); ++row) {
        std::istringstream iss(line);
        int64_t col = 0;
        for (double v; col < cols && iss >> v; ++col) out[static_cast<size_t>(row * cols + col)] = v;
        if (col != cols)
            throw IoError(std::string(what) + ": line " + std::to_string(row + 2) + " of '" + path + "' holds " +
                          std::to_string(col) + " numbers, " + std::to_string(cols) + " expected");
    }
    if (row != *count)
        throw IoError(std::string(what) + ": '" + path + "' ends after " + std::to_string(row) + " of " +
                      std::to_string(*count) + " rows");
    return out;
}

PointCloud loadFromFile(const std::string& path) {
    PointCloud pc;
    pc.data = readRows(path, "loadFromFile", &pc.size, &pc.dim, 1);
    return pc;
}

static Query loadQuery(const std::string& path, QueryType t) {
    int64_t n = 0, dim = 0;
    std::vector<double> rows = readRows(path, "loadQueryFromFile", &n, &dim, t == QueryType::RANGE ? 2 : 1);
    if (n > INT_MAX || dim > INT_MAX) throw IoError("loadQueryFromFile: '" + path + "' is too large");
    Query q(static_cast<int>(n), static_cast<int>(dim), t);
    q.data = std::move(rows);
    return q;
}

// Label of a range query file for the CSV's "Query Parameter" column
// (Utils.cu:18-35): the text between the path's last '_' and the next '.';
// "P-E" (integers) becomes "P%^E=C%" with C = (P/100)^E * 100 to two
// decimals, anything else gets a '%' appended.
std::string rangeInfoOfPath(const std::string& path) {
    const size_t us = path.rfind('_');
    const size_t b = us == std::string::npos ? 0 : us + 1;
    const size_t dot = path.find('.', b);
    const std::string seg = path.substr(b, dot == std::string::npos ? std::string::npos : dot - b);
    const size_t dash = seg.find('-');
    if (dash == std::string::npos) return seg + "%";
    const int pct = std::stoi(seg.substr(0, dash));
    const int exps = std::stoi(seg.substr(dash + 1));
    std::ostringstream o;
    o << pct << "%^" << exps << "=" << std::fixed << std::setprecision(2)
      << std::pow(pct / 100.0, exps) * 100.0 << "%";
    return o.str();
}

Query loadQueryPointFromFile(const std::string& path) { return loadQuery(path, QueryType::POINT); }
Query loadQueryRangeFromFile(const std::string& path) {
    Query q = loadQuery(path, QueryType::RANGE);
    q.queryRangeInfo = rangeInfoOfPath(path);
    return q;
}

void saveQueryToFile(const Query& q, const std::string& path) {
    std::ofstream f(path);
    if (!f) throw IoError("saveQueryToFile: '" + path + "' could not be created");
    f << q.length << ' ' << q.dim << '\n';
    f << std::setprecision(17);
    const int cols = q.type == QueryType::RANGE ? 2 * q.dim : q.dim;
    for (int i = 0; i < q.length; ++i) {
        for (int j = 0; j < cols; ++j) {
            if (j) f << ' ';
            f << q.data[static_cast<size_t>(i) * cols + j];
        }
        f << '\n';
    }
}

std::vector<double> Query::getQueryPoint(int i) const {
    if (type != QueryType::POINT) throw UsageError("getQueryPoint: the query batch does not hold point (KNN) queries");
    if (i < 0 || i >= length) throw std::out_of_range("getQueryPoint: query " + std::to_string(i) + " of " + std::to_string(length));
    return std::vector<double>(data.begin() + static_cast<long>(i) * dim, data.begin() + static_cast<long>(i + 1) * dim);
}

std::pair<std::vector<double>, std::vector<double>> Query::getQueryRange(int i) const {
    if (type != QueryType::RANGE) throw UsageError("getQueryRange: the query batch does not hold range queries");
    if (i < 0 || i >= length) throw std::out_of_range("getQueryRange: query " + std::to_string(i) + " of " + std::to_string(length));
    const size_t off = static_cast<size_t>(i) * 2 * dim;
    return {std::vector<double>(data.begin() + off, data.begin() + off + dim),
            std::vector<double>(data.begin() + off + dim, data.begin() + off + 2 * dim)};
}

void Query::setQueryPoint(int i, const std::vector<double>& p) {
    if (type != QueryType::POINT) throw UsageError("setQueryPoint: the query batch does not hold point (KNN) queries");
    if (i < 0 || i >= length) throw std::out_of_range("setQueryPoint: query " + std::to_string(i) + " of " + std::to_string(length));
    if (static_cast<int>(p.size()) != dim) throw std::invalid_argument("setQueryPoint: dimension mismatch");
    std::copy(p.begin(), p.end(), data.begin() + static_cast<long>(i) * dim);
}

void Query::setQueryRange(int i, const std::vector<double>& lo, const std::vector<double>& hi) {
    if (type != QueryType::RANGE) throw UsageError("setQueryRange: the query batch does not hold range queries");
    if (i < 0 || i >= length) throw std::out_of_range("setQueryRange: query " + std::to_string(i) + " of " + std::to_string(length));
    if (static_cast<int>(lo.size()) != dim || static_cast<int>(hi.size()) != dim)
        throw std::invalid_argument("setQueryRange: dimension mismatch");
    const size_t off = static_cast<size_t>(i) * 2 * dim;
    std::copy(lo.begin(), lo.end(), data.begin() + off);
    std::copy(hi.begin(), hi.end(), data.begin() + off + dim);
}

}  // namespace tblaeq_host
