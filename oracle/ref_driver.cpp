// Thin C wrapper compiled together with the reference's own pure-C++ headers
// (src/Kmeans/RandomKmeans.cuh, src/Kmeans/CPUKmeans.cuh, src/func.hpp) taken
// from the reference tree at build time (-I<reference root>).  It produces
// oracle/_ref/libref.so, used only by tests/ to pin the oracle's generator,
// argsort, radius distance and CPU k-means against the reference code itself.
// TEST INFRASTRUCTURE ONLY; nothing here is shipped or copied.
#include <cstdint>
#include <cstring>
#include <vector>

#include "src/Kmeans/CPUKmeans.cuh"
#include "src/Kmeans/RandomKmeans.cuh"
#include "src/func.hpp"

namespace {
struct Gen {
    std::vector<std::vector<double>> layers;  // finest first, flat
    std::vector<std::vector<size_t>> labels;
};
}  // namespace

extern "C" {

void* ref_gen_create(uint64_t N, uint64_t D, const uint64_t* ratios, uint64_t nr, double vmin, double vmax,
                     int isInt, uint64_t seed, double sigmaDivisor) {
    std::vector<size_t> r(ratios, ratios + nr);
    RandomKmeans::Config cfg(seed, sigmaDivisor);
    RandomKmeans km(N, D, r, vmin, vmax, isInt != 0, cfg);
    auto* g = new Gen();
    for (size_t i = 0; i < nr; ++i) {
        const size_t n = km.getdatas().size();
        const size_t k = (r[i] == 0 || n / r[i] == 0) ? 1 : n / r[i];
        km.run(k, 0);
        std::vector<double> flat;
        for (const auto& p : km.getdatas()) flat.insert(flat.end(), p.begin(), p.end());
        g->layers.push_back(std::move(flat));
        g->labels.push_back(km.getLabels());
        if (i + 1 == nr) {
            std::vector<double> c;
            for (const auto& p : km.getCentroids()) c.insert(c.end(), p.begin(), p.end());
            g->layers.push_back(std::move(c));
        }
        km.reset();
    }
    return g;
}

uint64_t ref_gen_layer(void* h, uint64_t i, double* out) {
    auto* g = static_cast<Gen*>(h);
    if (out) std::memcpy(out, g->layers[i].data(), g->layers[i].size() * 8);
    return g->layers[i].size();
}

uint64_t ref_gen_labels(void* h, uint64_t i, uint64_t* out) {
    auto* g = static_cast<Gen*>(h);
    if (out)
        for (size_t j = 0; j < g->labels[i].size(); ++j) out[j] = g->labels[i][j];
    return g->labels[i].size();
}

void ref_gen_free(void* h) { delete static_cast<Gen*>(h); }

void ref_argsort(const uint64_t* labels, uint64_t n, uint64_t* out) {
    std::vector<size_t> l(labels, labels + n);
    const auto idx = Sort::Sorted_Layer_With_Original_idxs(l);
    for (size_t i = 0; i < n; ++i) out[i] = idx[i];
}

double ref_euclidean(const double* a, const double* b, uint64_t D) {
    return dist::euclidean(std::vector<double>(a, a + D), std::vector<double>(b, b + D));
}

void ref_kmeans(const double* data, uint64_t N, uint64_t D, uint64_t k, uint64_t maxIter, uint64_t* labels,
                double* centroids) {
    CPUKmeans km(data, N, D, true);
    km.run(k, maxIter);
    for (size_t i = 0; i < N; ++i) labels[i] = km.getLabels()[i];
    for (size_t c = 0; c < k; ++c)
        for (size_t d = 0; d < D; ++d) centroids[c * D + d] = km.getCentroids()[c][d];
}

}  // extern "C"
