// Autotuner of the batched engine's launch parameters.
//
// Every parameter it touches changes only how work is spread over the GPU,
// never the results; the tuner checks that by hashing each candidate's
// results on the device and comparing with the starting configuration.
// Candidates are timed with HIP events on the engine's stream (median of
// `reps` runs), one parameter at a time (coordinate descent in a fixed order).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../batch_internal.hpp"

namespace tblaeq {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Order-sensitive hash of n words: sum of mix(word + position).
__global__ void k_hash_words(const uint64_t* __restrict__ a, const uint64_t* dn, uint64_t n0, uint64_t* out) {
    const uint64_t n = dn ? *dn : n0;
    uint64_t h = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        h += mix64(a[i] + 0x9e3779b97f4a7c15ull * (i + 1));
    h = wave_sum_u64(h);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(reinterpret_cast<unsigned long long*>(out), h);
}

// Hash of KNN rows: the first n[q] (id, d2) pairs of every row.
__global__ void k_hash_rows(const uint64_t* __restrict__ ids, const double* __restrict__ d2,
                            const uint64_t* __restrict__ n, uint64_t nq, uint64_t Kmax, uint64_t* out) {
    uint64_t h = 0;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; q < nq;
         q += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t k = n[q];
        h += mix64(k + q * 0x9e3779b97f4a7c15ull);
        for (uint64_t e = 0; e < k; ++e)
            h += mix64(ids[q * Kmax + e] ^ mix64(__double_as_longlong(d2[q * Kmax + e]) + e));
    }
    h = wave_sum_u64(h);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(reinterpret_cast<unsigned long long*>(out), h);
}

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
            set_error("tblaeq_batch_autotune: device allocation failed");
            return TBLAEQ_ERR_NOMEM;
        }
        return TBLAEQ_OK;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct Candidate {
    int param;
    std::vector<uint64_t> values;
};

const char* param_name(int id) {
    switch (id) {
    case TBLAEQ_PARAM_KNN_SEED: return "knn_seed";
    case TBLAEQ_PARAM_KNN_WG_PER_CU: return "knn_wg_per_cu";
    case TBLAEQ_PARAM_KNN_MODE: return "knn_mode";
    case TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU: return "range_blocks_per_cu";
    case TBLAEQ_PARAM_STEP_BLOCKS_PER_CU: return "step_blocks_per_cu";
    case TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU: return "radix_blocks_per_cu";
    case TBLAEQ_PARAM_HIT_STAGE: return "hit_stage";
    case TBLAEQ_PARAM_RANGE_ROW_QUERIES: return "range_row_queries";
    case TBLAEQ_PARAM_KNN_ROW_QUERIES: return "knn_row_queries";
    case TBLAEQ_PARAM_KNN_LEAF_REFINE: return "knn_leaf_refine";
    default: return nullptr;
    }
}
const int kAllParams[] = {TBLAEQ_PARAM_KNN_SEED,           TBLAEQ_PARAM_KNN_WG_PER_CU,
                          TBLAEQ_PARAM_KNN_MODE,           TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU,
                          TBLAEQ_PARAM_STEP_BLOCKS_PER_CU, TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU,
                          TBLAEQ_PARAM_HIT_STAGE,          TBLAEQ_PARAM_RANGE_ROW_QUERIES,
                          TBLAEQ_PARAM_KNN_ROW_QUERIES,    TBLAEQ_PARAM_KNN_LEAF_REFINE};

}  // namespace
}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_batch_autotune(tblaeq_batch* b, const double* d_boxes, uint64_t n_boxes, const double* d_points,
                                     const uint64_t* d_K, uint64_t n_points, uint64_t Kmax, int reps,
                                     double* out_ms) {
    if (!b || (n_boxes && !d_boxes) || (n_points && (!d_points || !d_K)) || (n_boxes == 0 && n_points == 0) ||
        reps < 1) {
        set_error("tblaeq_batch_autotune: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    tblaeq_ctx* ctx = b->ctx;
    const bool timing = ctx->timing;
    ctx->timing = false;
    int rc = TBLAEQ_OK;
    DevBuf off, ids, status, hash, kids, kd2, kn;
    const uint64_t outCap = std::max<uint64_t>(b->hitCap, 1);
    if (n_boxes) {
        if ((rc = off.alloc((n_boxes + 1) * 8)) || (rc = ids.alloc(outCap * 8))) return rc;
    }
    if (n_points) {
        if ((rc = kids.alloc(n_points * Kmax * 8)) || (rc = kd2.alloc(n_points * Kmax * 8)) ||
            (rc = kn.alloc(n_points * 8)))
            return rc;
    }
    if ((rc = status.alloc(4 * 8)) || (rc = hash.alloc(8))) return rc;
    hipEvent_t e0, e1;
    TB_HIP(hipEventCreate(&e0));
    TB_HIP(hipEventCreate(&e1));
    auto done = [&](int code) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->timing = timing;
        return code;
    };
    // runs one batch kind `reps` times, returns the median ms and the result hash
    auto run = [&](bool range, double* ms, uint64_t* h) -> int {
        std::vector<float> t;
        for (int r = 0; r < reps + 1; ++r) {
            TB_HIP(hipEventRecord(e0, ctx->stream));
            int c = range ? tblaeq_batch_range(b, d_boxes, n_boxes, off.as<uint64_t>(), ids.as<uint64_t>(), outCap,
                                               status.as<uint64_t>(), nullptr, nullptr)
                          : tblaeq_batch_knn(b, d_points, d_K, n_points, Kmax, kids.as<uint64_t>(), kd2.as<double>(),
                                             kn.as<uint64_t>(), status.as<uint64_t>(), nullptr, nullptr);
            if (c) return c;
            TB_HIP(hipEventRecord(e1, ctx->stream));
            TB_HIP(hipEventSynchronize(e1));
            float f = 0;
            TB_HIP(hipEventElapsedTime(&f, e0, e1));
            if (r > 0) t.push_back(f);  // the first run warms caches and workspaces
        }
        uint64_t st[4];
        TB_HIP(hipMemcpyAsync(st, status.p, 32, hipMemcpyDeviceToHost, ctx->stream));
        TB_HIP(hipStreamSynchronize(ctx->stream));
        if (st[1] != 0) {
            set_error("tblaeq_batch_autotune: the sample batch exceeds the engine capacities");
            return TBLAEQ_ERR_CAPACITY;
        }
        TB_HIP(hipMemsetAsync(hash.p, 0, 8, ctx->stream));
        const unsigned g = static_cast<unsigned>(ctx->cuCount) * 4;
        if (range) {
            hipLaunchKernelGGL(k_hash_words, dim3(g), dim3(256), 0, ctx->stream, off.as<uint64_t>(), nullptr,
                               n_boxes + 1, hash.as<uint64_t>());
            hipLaunchKernelGGL(k_hash_words, dim3(g), dim3(256), 0, ctx->stream, ids.as<uint64_t>(), nullptr, st[0],
                               hash.as<uint64_t>());
        } else {
            hipLaunchKernelGGL(k_hash_rows, dim3(g), dim3(256), 0, ctx->stream, kids.as<uint64_t>(), kd2.as<double>(),
                               kn.as<uint64_t>(), n_points, Kmax, hash.as<uint64_t>());
        }
        TB_HIP(hipMemcpyAsync(h, hash.p, 8, hipMemcpyDeviceToHost, ctx->stream));
        TB_HIP(hipStreamSynchronize(ctx->stream));
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return TBLAEQ_OK;
    };
    const std::vector<Candidate> rangeC = {{TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU, {4, 8, 16, 32}},
                                           {TBLAEQ_PARAM_HIT_STAGE, {256, 512, 1024}},
                                           {TBLAEQ_PARAM_RANGE_ROW_QUERIES, {0, 4, 8, 16, 32, 64}},
                                           {TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU, {2, 4, 8, 16}}};
    const std::vector<Candidate> knnC = {{TBLAEQ_PARAM_KNN_SEED, {16, 32, 64, 128}},
                                         {TBLAEQ_PARAM_KNN_WG_PER_CU, {2, 4, 8}},
                                         {TBLAEQ_PARAM_STEP_BLOCKS_PER_CU, {4, 8, 16, 32}},
                                         {TBLAEQ_PARAM_KNN_ROW_QUERIES, {0, 1, 2, 4}},
                                         {TBLAEQ_PARAM_KNN_LEAF_REFINE, {0, 1}}};
    for (int kind = 0; kind < 2; ++kind) {
        const bool range = kind == 0;
        if ((range && !n_boxes) || (!range && !n_points)) {
            if (out_ms) out_ms[2 * kind] = out_ms[2 * kind + 1] = 0.0;
            continue;
        }
        double base = 0;
        uint64_t ref = 0;
        if ((rc = run(range, &base, &ref))) return done(rc);
        if (out_ms) out_ms[2 * kind] = base;
        double best = base;
        for (const Candidate& c : range ? rangeC : knnC) {
            uint64_t keep = 0;
            tblaeq_batch_get_param(b, c.param, &keep);
            for (uint64_t v : c.values) {
                if (v == keep) continue;
                if ((rc = tblaeq_batch_set_param(b, c.param, v))) return done(rc);
                double ms = 0;
                uint64_t hsh = 0;
                if ((rc = run(range, &ms, &hsh))) return done(rc);
                if (hsh != ref) {
                    tblaeq_batch_set_param(b, c.param, keep);
                    set_error(std::string("tblaeq_batch_autotune: results changed with ") + param_name(c.param));
                    return done(TBLAEQ_ERR_INVALID);
                }
                if (ms < best * 0.98) {
                    best = ms;
                    keep = v;
                }
            }
            tblaeq_batch_set_param(b, c.param, keep);
        }
        if (out_ms) out_ms[2 * kind + 1] = best;
    }
    return done(TBLAEQ_OK);
}

extern "C" int tblaeq_batch_save_params(const tblaeq_batch* b, const char* path) {
    if (!b || !path) {
        set_error("tblaeq_batch_save_params: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    std::ofstream f(path);
    if (!f) {
        set_error(std::string("tblaeq_batch_save_params: cannot write ") + path);
        return TBLAEQ_ERR_IO;
    }
    f << "# tblaeq batched-engine launch parameters\n";
    f << "# shape D=" << b->D << " levels=" << b->L;
    for (uint64_t c : b->count) f << ' ' << c;
    f << '\n';
    for (int id : kAllParams) {
        uint64_t v = 0;
        tblaeq_batch_get_param(b, id, &v);
        f << param_name(id) << ' ' << v << '\n';
    }
    return f.good() ? TBLAEQ_OK : TBLAEQ_ERR_IO;
}

extern "C" int tblaeq_batch_load_params(tblaeq_batch* b, const char* path) {
    if (!b || !path) {
        set_error("tblaeq_batch_load_params: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    std::ifstream f(path);
    if (!f) {
        set_error(std::string("tblaeq_batch_load_params: cannot read ") + path);
        return TBLAEQ_ERR_IO;
    }
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        std::string name;
        uint64_t v = 0;
        if (!(ls >> name >> v)) {
            set_error("tblaeq_batch_load_params: malformed line: " + line);
            return TBLAEQ_ERR_IO;
        }
        int id = -1;
        for (int p : kAllParams)
            if (name == param_name(p)) id = p;
        if (id < 0) {
            set_error("tblaeq_batch_load_params: unknown parameter " + name);
            return TBLAEQ_ERR_IO;
        }
        const int rc = tblaeq_batch_set_param(b, id, v);
        if (rc) return rc;
    }
    return TBLAEQ_OK;
}
