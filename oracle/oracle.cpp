// CPU oracle for the T-BLAEQ query path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Built with -ffp-contract=off: every fused multiply-add below is an explicit std::fma.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Csc {
    uint64_t D = 0, rows = 0, cols = 0, nnz = 0;
    std::vector<uint64_t> rowIds, nnzPerCol, colRes;
    std::vector<double> vals;
};

}  // namespace

struct OIndex {
    uint64_t D = 0, N = 0, height = 0, intervals = 0;
    std::vector<uint64_t> ratios, meshSizes;       // meshSizes finest first
    std::vector<Csc> P;                            // P[l], l = 0 is the coarsest interval
    std::vector<std::vector<double>> radius;       // radius[l], indexed by level-l point id
    std::vector<std::vector<uint64_t>> maps;       // maps[i], i = 1..height-1 (maps[0] empty)
    uint64_t cRows = 0;                            // coarsest mesh
    std::vector<uint64_t> cIds;
    std::vector<double> cVals;
    std::vector<double> finestData;                // generator output (finest layer), if generated
    std::vector<std::vector<double>> genLayers;    // generator layers, finest first (if generated)
    std::vector<std::vector<uint64_t>> genLabels;
};

namespace {

struct Grid {
    std::vector<uint64_t> ids;
    std::vector<double> vals;
    size_t size() const { return ids.size(); }
};

}  // namespace

struct OResult {
    uint64_t D = 0;
    Grid fine;
    std::vector<uint64_t> cand, surv, out;
    std::vector<std::vector<uint64_t>> survIds;
};

namespace {

// QueryHandler.cu:16-45 -- truncate the plan when a level would collapse to one centroid.
std::vector<uint64_t> adjustPlan(uint64_t N, const std::vector<uint64_t>& ratios) {
    uint64_t n = N;
    size_t keep = ratios.size();
    for (size_t i = 0; i < ratios.size(); ++i) {
        const uint64_t c = (ratios[i] == 0 || n / ratios[i] == 0) ? 1 : n / ratios[i];
        if (c == 1) {
            keep = i + 1;
            break;
        }
        n = c;
    }
    return std::vector<uint64_t>(ratios.begin(), ratios.begin() + keep);
}

// Build the index arrays of one interval from (fine layer, coarse layer, labels).
// Setup.cuh:87-146 (P values = fine/coarse component ratio, CSC by label),
// func.hpp:124-137 (map = argsort of labels with std::sort), Setup.cuh:148-171 +
// func.hpp:202-213 (radius = max Euclidean distance, plain mul/add then sqrt).
void buildInterval(OIndex& ix, size_t pIdx, size_t fineIdx, const std::vector<double>& fine,
                   const std::vector<double>& coarse, const std::vector<uint64_t>& labels, uint64_t D) {
    const uint64_t nFine = labels.size();
    const uint64_t nCoarse = coarse.size() / D;
    Csc& P = ix.P[pIdx];
    P.D = D;
    P.rows = nFine;
    P.cols = nCoarse;
    P.nnz = nFine;
    P.nnzPerCol.assign(nCoarse, 0);
    for (uint64_t l : labels) ++P.nnzPerCol[l];
    P.colRes.assign(nCoarse + 1, 0);
    for (uint64_t c = 0; c < nCoarse; ++c) P.colRes[c + 1] = P.colRes[c] + P.nnzPerCol[c];
    P.rowIds.resize(nFine);
    std::iota(P.rowIds.begin(), P.rowIds.end(), uint64_t{0});

    std::vector<size_t> order(nFine);
    std::iota(order.begin(), order.end(), size_t{0});
    std::sort(order.begin(), order.end(), [&labels](size_t a, size_t b) { return labels[a] < labels[b]; });
    auto& map = ix.maps[fineIdx];
    map.assign(order.begin(), order.end());

    P.vals.resize(nFine * D);
    for (uint64_t c = 0; c < nCoarse; ++c) {
        for (uint64_t pos = P.colRes[c]; pos < P.colRes[c + 1]; ++pos) {
            const uint64_t src = map[pos];
            for (uint64_t d = 0; d < D; ++d) P.vals[pos * D + d] = fine[src * D + d] / coarse[c * D + d];
        }
    }
    auto& r = ix.radius[pIdx];
    r.assign(nCoarse, 0.0);
    for (uint64_t c = 0; c < nCoarse; ++c) {
        double best = 0.0;
        for (uint64_t pos = P.colRes[c]; pos < P.colRes[c + 1]; ++pos) {
            const uint64_t src = map[pos];
            double s = 0.0;
            for (uint64_t d = 0; d < D; ++d) {
                const double df = coarse[c * D + d] - fine[src * D + d];
                const double sq = df * df;  // contraction is off: separate rounding as in func.hpp:209-210
                s = s + sq;
            }
            best = std::max(best, std::sqrt(s));
        }
        r[c] = best;
    }
}

OIndex* assemble(uint64_t N, uint64_t D, const std::vector<uint64_t>& ratios,
                 const std::vector<std::vector<double>>& layers, const std::vector<std::vector<uint64_t>>& labels) {
    auto* ix = new OIndex();
    ix->D = D;
    ix->N = N;
    ix->height = ratios.size() + 1;
    ix->intervals = ratios.size();
    ix->ratios = ratios;
    ix->P.resize(ix->intervals);
    ix->radius.resize(ix->intervals);
    ix->maps.resize(ix->height);
    // IndexBuilder.cu:196-226: interval i goes fine (layer i) -> coarse (layer i+1).
    for (size_t i = 0; i < ix->intervals; ++i) {
        const size_t fineIdx = ix->intervals - i;
        const size_t pIdx = fineIdx - 1;
        ix->meshSizes.push_back(layers[i].size() / D);
        buildInterval(*ix, pIdx, fineIdx, layers[i], layers[i + 1], labels[i], D);
        if (i + 2 == ix->height) {
            const uint64_t nc = layers[i + 1].size() / D;
            ix->meshSizes.push_back(nc);
            ix->cRows = nc;
            ix->cIds.resize(nc);
            std::iota(ix->cIds.begin(), ix->cIds.end(), uint64_t{0});
            ix->cVals = layers[i + 1];
        }
    }
    return ix;
}

// ---- traversal primitives -------------------------------------------------

// RangePruning.cu:5-25: squared distance from a centroid to an axis-aligned box.
inline double rangeDist2(const double* c, const double* lo, const double* hi, uint64_t D) {
    double dist = 0.0;
    for (uint64_t d = 0; d < D; ++d) {
        double diff = 0.0;
        if (c[d] < lo[d]) diff = lo[d] - c[d];
        else if (c[d] > hi[d]) diff = c[d] - hi[d];
        dist = std::fma(diff, diff, dist);
    }
    return dist;
}

inline double sqrtRd(double x) {
    double s = std::sqrt(x);
    if (std::fma(s, s, -x) > 0.0) s = std::nextafter(s, -std::numeric_limits<double>::infinity());
    return s;
}

// KNN.cu:9-24: lower bound = sqrt_rd(sum (q-c)^2) - r.
inline double knnLb(const double* c, const double* q, double r, uint64_t D) {
    double dist = 0.0;
    for (uint64_t d = 0; d < D; ++d) {
        const double diff = q[d] - c[d];
        dist = std::fma(diff, diff, dist);
    }
    return sqrtRd(dist) - r;
}

// GridCompact.cuh:41-100 -- order-preserving compaction of (id, vals).
Grid compact(const Grid& g, const std::vector<uint8_t>& mask, uint64_t D) {
    Grid o;
    for (size_t i = 0; i < g.size(); ++i) {
        if (!mask[i]) continue;
        o.ids.push_back(g.ids[i]);
        o.vals.insert(o.vals.end(), g.vals.begin() + i * D, g.vals.begin() + (i + 1) * D);
    }
    return o;
}

// SpMSpV.cu:200-254 + SpMSpV.cuh:47-67 + Refactor.cu:10-16: for every survivor
// (grid order) expand its CSC column (column order): y = P[j] * x, id = map[rowIds[j]].
Grid prolong(const OIndex& ix, size_t l, const Grid& s) {
    const Csc& P = ix.P[l];
    const auto& map = ix.maps[l + 1];
    const uint64_t D = ix.D;
    Grid o;
    uint64_t n = 0;
    for (size_t i = 0; i < s.size(); ++i) n += P.colRes[s.ids[i] + 1] - P.colRes[s.ids[i]];
    o.ids.resize(n);
    o.vals.resize(n * D);
    uint64_t w = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const uint64_t c = s.ids[i];
        for (uint64_t j = P.colRes[c]; j < P.colRes[c + 1]; ++j, ++w) {
            o.ids[w] = map[P.rowIds[j]];
            for (uint64_t d = 0; d < D; ++d) o.vals[w * D + d] = P.vals[j * D + d] * s.vals[i * D + d];
        }
    }
    return o;
}

// KNN STEP in closed form.  KNN.cu:75-96 walks the clusters sorted by lower bound:
// while the running child count is below K it takes the cluster and raises
// ub = max(ub, lb + 2r); afterwards it takes clusters while lb <= ub.  Because the
// walk is over ascending lb and every taken cluster has lb <= lb + 2r <= ub, the
// selected set equals {lb <= ub} with ub from the first phase.  Ties of equal lb are
// ordered by point id.
std::vector<uint8_t> knnStep(const OIndex& ix, size_t l, const Grid& g, const double* q, uint64_t K) {
    const uint64_t D = ix.D;
    const size_t p = g.size();
    std::vector<double> lb(p);
    for (size_t i = 0; i < p; ++i) lb[i] = knnLb(&g.vals[i * D], q, ix.radius[l][g.ids[i]], D);
    std::vector<size_t> ord(p);
    std::iota(ord.begin(), ord.end(), size_t{0});
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        if (lb[a] != lb[b]) return lb[a] < lb[b];
        return g.ids[a] < g.ids[b];
    });
    const auto& w = ix.P[l].nnzPerCol;
    uint64_t count = 0;
    double ub = 0.0;
    size_t i = 0;
    for (; i < p && count < K; ++i) {
        const uint64_t id = g.ids[ord[i]];
        count += w[id];
        ub = std::max(ub, lb[ord[i]] + 2.0 * ix.radius[l][id]);
    }
    std::vector<uint8_t> mask(p, 0);
    for (size_t j = 0; j < p; ++j) mask[j] = (lb[j] <= ub) ? 1 : 0;
    for (size_t j = 0; j < i; ++j) mask[ord[j]] = 1;
    return mask;
}

OResult* traverse(const OIndex& ix, int type, const double* lo, const double* hi, const double* q, uint64_t K) {
    auto* r = new OResult();
    const uint64_t D = ix.D;
    r->D = D;
    Grid g;
    g.ids = ix.cIds;
    g.vals = ix.cVals;
    for (size_t l = 0; l < ix.intervals; ++l) {
        const size_t p = g.size();
        std::vector<uint8_t> mask(p, 0);
        if (type == 0) {
            for (size_t i = 0; i < p; ++i) {
                const double rr = ix.radius[l][g.ids[i]];
                mask[i] = rangeDist2(&g.vals[i * D], lo, hi, D) <= rr * rr;
            }
        } else {
            mask = knnStep(ix, l, g, q, K);
        }
        Grid s = compact(g, mask, D);
        r->cand.push_back(p);
        r->surv.push_back(s.size());
        r->survIds.push_back(s.ids);
        g = prolong(ix, l, s);
        r->out.push_back(g.size());
    }
    r->fine = std::move(g);
    return r;
}

inline double pointDist2(const double* v, const double* q, uint64_t D) {
    double acc = 0.0;
    for (uint64_t d = 0; d < D; ++d) {
        const double diff = q[d] - v[d];
        acc = std::fma(diff, diff, acc);
    }
    return acc;
}

inline bool inBox(const double* v, const double* lo, const double* hi, uint64_t D) {
    for (uint64_t d = 0; d < D; ++d)
        if (!(v[d] >= lo[d] && v[d] <= hi[d])) return false;
    return true;
}

uint64_t topK(const std::vector<std::pair<double, uint64_t>>& in, uint64_t K, uint64_t* ids, double* d2) {
    std::vector<std::pair<double, uint64_t>> v = in;
    const size_t k = std::min<size_t>(K, v.size());
    std::partial_sort(v.begin(), v.begin() + k, v.end());
    for (size_t i = 0; i < k; ++i) {
        ids[i] = v[i].second;
        d2[i] = v[i].first;
    }
    return k;
}

template <class T>
void wr(std::ofstream& f, const T* p, size_t n) {
    f.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
}
template <class T>
void rd(std::ifstream& f, T* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
    if (!f) throw std::runtime_error("short read");
}

// CPUKmeans.cuh:122-145 (k-means++) and :152-205 (Lloyd).
struct KM {
    const double* x;
    uint64_t N, D;
    static double sq(const double* a, const double* b, uint64_t D) {
        double s = 0.0;
        for (uint64_t d = 0; d < D; ++d) {
            const double df = a[d] - b[d];
            const double p = df * df;
            s = s + p;
        }
        return s;
    }
    std::vector<double> seed(uint64_t k, uint64_t streamSeed = 42) const {
        std::vector<double> c;
        c.reserve(k * D);
        std::mt19937_64 rng(streamSeed);
        std::uniform_int_distribution<size_t> uni(0, N - 1);
        const size_t first = uni(rng);
        c.insert(c.end(), x + first * D, x + first * D + D);
        std::vector<double> minD(N, std::numeric_limits<double>::max());
        for (uint64_t cc = 1; cc < k; ++cc) {
            const double* last = &c[(cc - 1) * D];
            double total = 0.0;
            for (uint64_t i = 0; i < N; ++i) {
                const double dd = sq(x + i * D, last, D);
                if (dd < minD[i]) minD[i] = dd;
                total += minD[i];
            }
            std::uniform_real_distribution<double> dist(0.0, total);
            const double thr = dist(rng);
            double cum = 0.0;
            size_t chosen = N - 1;
            for (uint64_t i = 0; i < N; ++i) {
                cum += minD[i];
                if (cum >= thr) {
                    chosen = i;
                    break;
                }
            }
            c.insert(c.end(), x + chosen * D, x + chosen * D + D);
        }
        return c;
    }
    double assign(const double* c, uint64_t k, uint64_t* lab, bool* changed) const {
        double inertia = 0.0;
        bool ch = false;
        for (uint64_t i = 0; i < N; ++i) {
            double best = std::numeric_limits<double>::max();
            uint64_t bc = 0;
            for (uint64_t j = 0; j < k; ++j) {
                const double dd = sq(x + i * D, c + j * D, D);
                if (dd < best) {
                    best = dd;
                    bc = j;
                }
            }
            if (lab[i] != bc) {
                lab[i] = bc;
                ch = true;
            }
            inertia += best;
        }
        if (changed) *changed = ch;
        return inertia;
    }
    void update(double* c, uint64_t k, const uint64_t* lab) const {
        std::vector<double> sums(k * D, 0.0);
        std::vector<uint64_t> cnt(k, 0);
        for (uint64_t i = 0; i < N; ++i) {
            ++cnt[lab[i]];
            for (uint64_t d = 0; d < D; ++d) sums[lab[i] * D + d] += x[i * D + d];
        }
        for (uint64_t j = 0; j < k; ++j)
            if (cnt[j] > 0)
                for (uint64_t d = 0; d < D; ++d) c[j * D + d] = sums[j * D + d] / static_cast<double>(cnt[j]);
    }
};

}  // namespace

extern "C" {

// RandomKmeans.cuh:176-250 restated over flat arrays, then IndexBuilder.cu:153-234.
OIndex* oracle_gen_index(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, const uint64_t* ratiosIn,
                         uint64_t nratios, uint64_t seed, double sigmaDivisor) {
    if (!(vmin > 0.0) || !(vmin < vmax) || !(sigmaDivisor > 0.0) || D == 0 || N == 0) return nullptr;
    const std::vector<uint64_t> ratios = adjustPlan(N, std::vector<uint64_t>(ratiosIn, ratiosIn + nratios));
    const size_t L = ratios.size();
    std::vector<uint64_t> sizes(L + 1);
    sizes[0] = N;
    for (size_t i = 0; i < L; ++i) sizes[i + 1] = std::max<uint64_t>(1, sizes[i] / ratios[i]);
    std::vector<std::vector<double>> layers(L + 1);
    std::vector<std::vector<uint64_t>> labels(L);
    std::mt19937_64 rng(seed);
    auto rnd = [isInt](double v) { return isInt ? std::round(v) : v; };
    std::uniform_real_distribution<double> uni(vmin, vmax);
    layers[L].resize(sizes[L] * D);
    for (uint64_t i = 0; i < sizes[L] * D; ++i) layers[L][i] = rnd(uni(rng));
    for (size_t lvl = L; lvl > 0; --lvl) {
        const uint64_t nP = sizes[lvl], nC = sizes[lvl - 1];
        const double spacing = (vmax - vmin) / std::pow(static_cast<double>(nP), 1.0 / static_cast<double>(D));
        std::normal_distribution<double> nd(0.0, spacing / sigmaDivisor);
        auto& ch = layers[lvl - 1];
        auto& lab = labels[lvl - 1];
        ch.resize(nC * D);
        lab.resize(nC);
        const auto& par = layers[lvl];
        for (uint64_t ci = 0; ci < nC; ++ci) {
            const uint64_t p = ci % nP;
            lab[ci] = p;
            for (uint64_t d = 0; d < D; ++d) {
                double v = par[p * D + d] + nd(rng);
                v = std::clamp(v, vmin, vmax);
                ch[ci * D + d] = rnd(v);
            }
        }
    }
    OIndex* ix = assemble(N, D, ratios, layers, labels);
    ix->finestData = layers[0];
    ix->genLayers = std::move(layers);
    ix->genLabels = std::move(labels);
    return ix;
}

// IndexBuilder.cu:44-151 with the CPU backend (CPUKmeans).
OIndex* oracle_build_index(const double* data, uint64_t N, uint64_t D, const uint64_t* ratiosIn, uint64_t nratios,
                           uint64_t maxIter) {
    const std::vector<uint64_t> ratios = adjustPlan(N, std::vector<uint64_t>(ratiosIn, ratiosIn + nratios));
    const size_t L = ratios.size();
    std::vector<std::vector<double>> layers(L + 1);
    std::vector<std::vector<uint64_t>> labels(L);
    layers[0].assign(data, data + N * D);
    for (size_t i = 0; i < L; ++i) {
        const uint64_t n = layers[i].size() / D;
        const uint64_t k = (ratios[i] == 0 || n / ratios[i] == 0) ? 1 : n / ratios[i];
        labels[i].assign(n, 0);
        layers[i + 1].resize(k * D);
        oracle_kmeans(layers[i].data(), n, D, k, maxIter, labels[i].data(), layers[i + 1].data());
    }
    return assemble(N, D, ratios, layers, labels);
}

// IndexSerializer.cu:34-259 (native-endian, size_t fields, 1-byte bool).
OIndex* oracle_load_index(const char* dirc) {
    try {
        const std::string dir(dirc);
        auto* ix = new OIndex();
        {
            std::ifstream f(dir + "/metadata.bin", std::ios::binary);
            if (!f) throw std::runtime_error("metadata");
            uint8_t aos;
            rd(f, &ix->D, 1);
            rd(f, &ix->N, 1);
            rd(f, &ix->height, 1);
            rd(f, &ix->intervals, 1);
            rd(f, &aos, 1);
            ix->ratios.resize(ix->intervals);
            rd(f, ix->ratios.data(), ix->intervals);
            uint64_t msz;
            rd(f, &msz, 1);
            ix->meshSizes.resize(msz);
            rd(f, ix->meshSizes.data(), msz);
        }
        ix->P.resize(ix->intervals);
        ix->radius.resize(ix->intervals);
        ix->maps.resize(ix->height);
        for (size_t i = 0; i < ix->intervals; ++i) {
            std::ifstream f(dir + "/P_Tensor_" + std::to_string(i) + ".bin", std::ios::binary);
            if (!f) throw std::runtime_error("P");
            Csc& P = ix->P[i];
            uint8_t aos;
            rd(f, &P.D, 1);
            rd(f, &P.rows, 1);
            rd(f, &P.cols, 1);
            rd(f, &P.nnz, 1);
            rd(f, &aos, 1);
            P.rowIds.resize(P.nnz);
            P.nnzPerCol.resize(P.cols);
            P.colRes.resize(P.cols + 1);
            P.vals.resize(P.nnz * P.D);
            rd(f, P.rowIds.data(), P.nnz);
            rd(f, P.nnzPerCol.data(), P.cols);
            rd(f, P.colRes.data(), P.cols + 1);
            rd(f, P.vals.data(), P.nnz * P.D);
        }
        for (size_t i = 0; i < ix->intervals; ++i) {
            std::ifstream f(dir + "/MaxRadius_" + std::to_string(i) + ".bin", std::ios::binary);
            if (!f) throw std::runtime_error("R");
            uint64_t n;
            rd(f, &n, 1);
            ix->radius[i].resize(n);
            rd(f, ix->radius[i].data(), n);
        }
        for (size_t i = 1; i < ix->height; ++i) {
            std::ifstream f(dir + "/Map_" + std::to_string(i) + ".bin", std::ios::binary);
            if (!f) continue;
            uint64_t n;
            rd(f, &n, 1);
            ix->maps[i].resize(n);
            rd(f, ix->maps[i].data(), n);
        }
        {
            std::ifstream f(dir + "/Coreast_Mesh.bin", std::ios::binary);
            if (!f) throw std::runtime_error("C");
            uint64_t dim, nnz;
            uint8_t aos;
            rd(f, &ix->cRows, 1);
            rd(f, &dim, 1);
            rd(f, &nnz, 1);
            rd(f, &aos, 1);
            ix->cIds.resize(nnz);
            ix->cVals.resize(nnz * dim);
            rd(f, ix->cIds.data(), nnz);
            rd(f, ix->cVals.data(), nnz * dim);
        }
        return ix;
    } catch (...) {
        return nullptr;
    }
}

int oracle_save_index(const OIndex* ix, const char* dirc) {
    try {
        const std::string dir(dirc);
        std::filesystem::create_directories(dir);
        const uint8_t aos = 1;
        {
            std::ofstream f(dir + "/metadata.bin", std::ios::binary);
            wr(f, &ix->D, 1);
            wr(f, &ix->N, 1);
            wr(f, &ix->height, 1);
            wr(f, &ix->intervals, 1);
            wr(f, &aos, 1);
            wr(f, ix->ratios.data(), ix->intervals);
            const uint64_t msz = ix->meshSizes.size();
            wr(f, &msz, 1);
            wr(f, ix->meshSizes.data(), msz);
        }
        for (size_t i = 0; i < ix->intervals; ++i) {
            std::ofstream f(dir + "/P_Tensor_" + std::to_string(i) + ".bin", std::ios::binary);
            const Csc& P = ix->P[i];
            wr(f, &P.D, 1);
            wr(f, &P.rows, 1);
            wr(f, &P.cols, 1);
            wr(f, &P.nnz, 1);
            wr(f, &aos, 1);
            wr(f, P.rowIds.data(), P.nnz);
            wr(f, P.nnzPerCol.data(), P.cols);
            wr(f, P.colRes.data(), P.cols + 1);
            wr(f, P.vals.data(), P.nnz * P.D);
        }
        for (size_t i = 0; i < ix->intervals; ++i) {
            std::ofstream f(dir + "/MaxRadius_" + std::to_string(i) + ".bin", std::ios::binary);
            const uint64_t n = ix->radius[i].size();
            wr(f, &n, 1);
            wr(f, ix->radius[i].data(), n);
        }
        for (size_t i = 1; i < ix->height; ++i) {
            if (ix->maps[i].empty()) continue;
            std::ofstream f(dir + "/Map_" + std::to_string(i) + ".bin", std::ios::binary);
            const uint64_t n = ix->maps[i].size();
            wr(f, &n, 1);
            wr(f, ix->maps[i].data(), n);
        }
        {
            std::ofstream f(dir + "/Coreast_Mesh.bin", std::ios::binary);
            const uint64_t nnz = ix->cIds.size();
            wr(f, &ix->cRows, 1);
            wr(f, &ix->D, 1);
            wr(f, &nnz, 1);
            wr(f, &aos, 1);
            wr(f, ix->cIds.data(), nnz);
            wr(f, ix->cVals.data(), nnz * ix->D);
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

void oracle_free_index(OIndex* ix) { delete ix; }

uint64_t oracle_index_meta(const OIndex* ix, uint64_t* meta, uint64_t* ms, uint64_t cap) {
    meta[0] = ix->D;
    meta[1] = ix->N;
    meta[2] = ix->height;
    meta[3] = ix->intervals;
    const uint64_t n = std::min<uint64_t>(cap, ix->meshSizes.size());
    for (uint64_t i = 0; i < n; ++i) ms[i] = ix->meshSizes[i];
    return ix->meshSizes.size();
}

uint64_t oracle_index_data(const OIndex* ix, double* out, uint64_t cap) {
    if (ix->finestData.empty() || cap < ix->finestData.size()) return 0;
    std::memcpy(out, ix->finestData.data(), ix->finestData.size() * sizeof(double));
    return ix->finestData.size();
}

uint64_t oracle_index_layer(const OIndex* ix, uint64_t i, double* out) {
    if (i >= ix->genLayers.size()) return 0;
    if (out) std::memcpy(out, ix->genLayers[i].data(), ix->genLayers[i].size() * sizeof(double));
    return ix->genLayers[i].size();
}

uint64_t oracle_index_labels(const OIndex* ix, uint64_t i, uint64_t* out) {
    if (i >= ix->genLabels.size()) return 0;
    if (out) std::memcpy(out, ix->genLabels[i].data(), ix->genLabels[i].size() * sizeof(uint64_t));
    return ix->genLabels[i].size();
}

// Index arrays of interval l: which = 0 colRes, 1 nnzPerCol, 2 rowIds, 3 map[l+1] (u64);
// 4 P values, 5 radius (f64).  Returns the element count, copies when out != NULL.
uint64_t oracle_index_array(const OIndex* ix, uint64_t l, int which, void* out) {
    const auto& P = ix->P[l];
    const void* src = nullptr;
    uint64_t n = 0;
    switch (which) {
        case 0: src = P.colRes.data(); n = P.colRes.size(); break;
        case 1: src = P.nnzPerCol.data(); n = P.nnzPerCol.size(); break;
        case 2: src = P.rowIds.data(); n = P.rowIds.size(); break;
        case 3: src = ix->maps[l + 1].data(); n = ix->maps[l + 1].size(); break;
        case 4: src = P.vals.data(); n = P.vals.size(); break;
        case 5: src = ix->radius[l].data(); n = ix->radius[l].size(); break;
        default: return 0;
    }
    if (out) std::memcpy(out, src, n * 8);
    return n;
}

void oracle_index_scale_radius(OIndex* ix, double factor) {
    for (auto& r : ix->radius)
        for (auto& v : r) v = v * factor;
}

OResult* oracle_range(const OIndex* ix, const double* lo, const double* hi) {
    return traverse(*ix, 0, lo, hi, nullptr, 0);
}
OResult* oracle_knn(const OIndex* ix, const double* q, uint64_t K) { return traverse(*ix, 1, nullptr, nullptr, q, K); }
void oracle_free_result(OResult* r) { delete r; }
uint64_t oracle_result_size(const OResult* r) { return r->fine.size(); }
void oracle_result_copy(const OResult* r, uint64_t* ids, double* vals) {
    if (ids) std::memcpy(ids, r->fine.ids.data(), r->fine.ids.size() * sizeof(uint64_t));
    if (vals) std::memcpy(vals, r->fine.vals.data(), r->fine.vals.size() * sizeof(double));
}
void oracle_result_levels(const OResult* r, uint64_t* cand, uint64_t* surv, uint64_t* out) {
    for (size_t l = 0; l < r->cand.size(); ++l) {
        cand[l] = r->cand[l];
        surv[l] = r->surv[l];
        out[l] = r->out[l];
    }
}
uint64_t oracle_result_survivors(const OResult* r, uint64_t l, uint64_t* ids, uint64_t cap) {
    const auto& v = r->survIds[l];
    const uint64_t n = std::min<uint64_t>(cap, v.size());
    std::memcpy(ids, v.data(), n * sizeof(uint64_t));
    return v.size();
}

uint64_t oracle_refine_range(const OResult* r, const double* lo, const double* hi, uint64_t* out, uint64_t cap) {
    std::vector<uint64_t> v;
    const uint64_t D = r->D;
    for (size_t i = 0; i < r->fine.size(); ++i)
        if (inBox(&r->fine.vals[i * D], lo, hi, D)) v.push_back(r->fine.ids[i]);
    std::sort(v.begin(), v.end());
    const uint64_t n = std::min<uint64_t>(cap, v.size());
    if (out) std::memcpy(out, v.data(), n * sizeof(uint64_t));
    return v.size();
}

uint64_t oracle_refine_knn(const OResult* r, const double* q, uint64_t K, uint64_t* ids, double* d2) {
    const uint64_t D = r->D;
    std::vector<std::pair<double, uint64_t>> v(r->fine.size());
    for (size_t i = 0; i < r->fine.size(); ++i) v[i] = {pointDist2(&r->fine.vals[i * D], q, D), r->fine.ids[i]};
    return topK(v, K, ids, d2);
}

void oracle_recon_finest(const OIndex* ix, double* out) {
    const uint64_t D = ix->D;
    Grid g;
    g.ids = ix->cIds;
    g.vals = ix->cVals;
    for (size_t l = 0; l < ix->intervals; ++l) g = prolong(*ix, l, g);
    for (size_t i = 0; i < g.size(); ++i)
        std::memcpy(out + g.ids[i] * D, &g.vals[i * D], D * sizeof(double));
}

// Reconstructed coordinates of every point of level `level` (0 = coarsest
// mesh, `intervals` = finest), by level id: the coarsest mesh prolonged level by
// level over all clusters (SpMSpV.cuh:47-67, Refactor.cu:10-16), the values
// every traversal sees at that level.  Returns the point count.
uint64_t oracle_recon_level(const OIndex* ix, uint64_t level, double* out) {
    const uint64_t D = ix->D;
    Grid g;
    g.ids = ix->cIds;
    g.vals = ix->cVals;
    for (size_t l = 0; l < level && l < ix->intervals; ++l) g = prolong(*ix, l, g);
    if (out)
        for (size_t i = 0; i < g.size(); ++i) std::memcpy(out + g.ids[i] * D, &g.vals[i * D], D * sizeof(double));
    return g.size();
}

uint64_t oracle_brute_range(const double* pts, uint64_t N, uint64_t D, const double* lo, const double* hi,
                            uint64_t* out, uint64_t cap) {
    uint64_t n = 0;
    for (uint64_t i = 0; i < N; ++i)
        if (inBox(pts + i * D, lo, hi, D)) {
            if (out && n < cap) out[n] = i;
            ++n;
        }
    return n;
}

uint64_t oracle_brute_knn(const double* pts, uint64_t N, uint64_t D, const double* q, uint64_t K, uint64_t* ids,
                          double* d2) {
    std::vector<std::pair<double, uint64_t>> v(N);
    for (uint64_t i = 0; i < N; ++i) v[i] = {pointDist2(pts + i * D, q, D), i};
    return topK(v, K, ids, d2);
}

double oracle_sqrt_rd(double x) { return sqrtRd(x); }
double oracle_range_dist2(const double* c, const double* lo, const double* hi, uint64_t D) {
    return rangeDist2(c, lo, hi, D);
}
double oracle_knn_lb(const double* c, const double* q, double r, uint64_t D) { return knnLb(c, q, r, D); }

// KNN.cu:75-96, literally: sequential walk over the sorted list.
uint64_t oracle_step_sequential(const double* lb, const uint64_t* w, const double* r, uint64_t p, uint64_t K,
                                uint8_t* sel) {
    uint64_t count = 0, n = 0;
    double cur = 0.0;
    for (uint64_t i = 0; i < p; ++i) sel[i] = 0;
    for (uint64_t i = 0; i < p; ++i) {
        if (count >= K) {
            if (lb[i] > cur) break;
            count += w[i];
            sel[i] = 1;
            ++n;
        } else {
            count += w[i];
            cur = std::max(cur, lb[i] + 2.0 * r[i]);
            sel[i] = 1;
            ++n;
        }
    }
    return n;
}

// MemoryPolicy.cu:100-177.
void oracle_policy_recommend(const uint64_t* nnz, const uint64_t* cols, uint64_t L, uint64_t D, uint64_t coarsest,
                             uint64_t maps, uint64_t freeMem, uint8_t* out) {
    uint64_t budget = static_cast<uint64_t>(static_cast<double>(freeMem) * 0.85);
    const uint64_t perm = coarsest + maps;
    if (budget <= perm) {
        for (uint64_t l = 0; l < L; ++l) out[l] = 3;
        return;
    }
    budget -= perm;
    for (uint64_t l = 0; l < L; ++l) {
        const uint64_t pv = nnz[l] * D * 8, rv = cols[l] * 8, sp = 3 * nnz[l] * 8;
        const uint64_t c0 = pv + rv + sp + pv, c1 = pv + rv + sp + pv, c2 = 2 * pv + sp;
        if (budget >= c0) {
            out[l] = 0;
            budget -= c0;
        } else if (budget >= c1) {
            out[l] = 1;
            budget -= c1;
        } else if (budget >= c2) {
            out[l] = 2;
        } else {
            out[l] = 3;
        }
    }
}

double oracle_cpu_batch(const OIndex* ix, int type, const double* qs, uint64_t nq, uint64_t K, int nthreads,
                        double* us, uint64_t* cnt, int* threadsUsed) {
    const uint64_t D = ix->D;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    if (threadsUsed) *threadsUsed = omp_get_max_threads();
#else
    if (threadsUsed) *threadsUsed = 1;
#endif
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < static_cast<int64_t>(nq); ++i) {
        const auto a = std::chrono::steady_clock::now();
        uint64_t c = 0;
        if (type == 0) {
            const double* lo = qs + i * 2 * D;
            OResult* r = traverse(*ix, 0, lo, lo + D, nullptr, 0);
            c = oracle_refine_range(r, lo, lo + D, nullptr, 0);
            delete r;
        } else {
            const double* q = qs + i * D;
            OResult* r = traverse(*ix, 1, nullptr, nullptr, q, K);
            std::vector<uint64_t> ids(K);
            std::vector<double> d2(K);
            c = oracle_refine_knn(r, q, K, ids.data(), d2.data());
            delete r;
        }
        const auto b = std::chrono::steady_clock::now();
        us[i] = std::chrono::duration<double, std::micro>(b - a).count();
        cnt[i] = c;
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

uint64_t oracle_kmeans(const double* data, uint64_t N, uint64_t D, uint64_t k, uint64_t maxIter, uint64_t* labels,
                       double* centroids) {
    return oracle_kmeans_seeded(data, N, D, k, maxIter, 42, labels, centroids);
}

// oracle_kmeans with the random stream's seed as a parameter (the CPU build
// seeds mt19937_64 with 42; the index-build configuration asks for 7).
uint64_t oracle_kmeans_seeded(const double* data, uint64_t N, uint64_t D, uint64_t k, uint64_t maxIter,
                              uint64_t streamSeed, uint64_t* labels, double* centroids) {
    KM km{data, N, D};
    std::vector<double> c = km.seed(k, streamSeed);
    for (uint64_t i = 0; i < N; ++i) labels[i] = 0;
    uint64_t it = 0;
    for (; it < maxIter; ++it) {
        bool changed = false;
        km.assign(c.data(), k, labels, &changed);
        km.update(c.data(), k, labels);
        if (!changed) {
            ++it;
            break;
        }
    }
    std::memcpy(centroids, c.data(), k * D * sizeof(double));
    return it;
}

double oracle_kmeans_assign(const double* data, uint64_t N, uint64_t D, const double* cent, uint64_t k,
                            uint64_t* labels) {
    KM km{data, N, D};
    for (uint64_t i = 0; i < N; ++i) labels[i] = ~uint64_t{0};
    return km.assign(cent, k, labels, nullptr);
}

}  // extern "C"
