// Learned KNN radius proposal: a 48 -> 64 -> 64 -> 4 ReLU MLP with BF16
// weights and activations on the MFMA cores (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), FP32 master weights and Adam on the device.
//
// Inputs (48 features per query): the query point scaled to the domain (16
// slots, zero-padded), the 31 smallest lower bounds lb = |q - c| - r over the
// coarsest clusters (scaled, ascending) and log(K) / log(256).  Outputs: [0]
// log(d_K / scale) of the K-th nearest candidate, [1..3] log(1 + survivors)
// of the first three levels / 10 (survivor budgets).  The engine turns output
// 0 into a search radius; exact certification decides whether the answer
// stands or the query falls back to the seed method.
//
// Every GEMM is C[M x N] (+)= op(A)[M x K] op(B)[K x N] with M, N, K multiples
// of 16, one 16 x 16 output tile per wave and an optional split over K
// (partial tiles added with FP32 atomics).  Per lane, the MFMA operands are
// A[row = lane % 16][k = 4 (lane / 16) + v] and B[k = 4 (lane / 16) + v][col =
// lane % 16], v = 0..3; the result is D[row = 4 (lane / 16) + v][col =
// lane % 16].
#include <cmath>
#include <cstring>
#include <vector>

#include "../../../include/tblaeq_learn.h"
#include "../batch_internal.hpp"

namespace tblaeq {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kIn = 48, kH = 64, kOutPad = 16, kOut = 4;
// packed FP32 parameters: W1 [48 x 64] | b1 [64] | W2 [64 x 64] | b2 [64] | W3 [64 x 16] | b3 [16]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOutPad, kParams = kOffB3 + kOutPad;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

struct GemmArgs {
    int M, N, K;
    const uint16_t* A;
    int lda, transA;  // op(A)[i][k] = transA ? A[k lda + i] : A[i lda + k]
    const uint16_t* B;
    int ldb, transB;  // op(B)[k][j] = transB ? B[j ldb + k] : B[k ldb + j]
    float* C;
    int ldc;
    int atomic;  // add into C with atomics (split K or accumulation)
    int kSplit;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = blockIdx.x * 4 + threadIdx.x / kWave;
    const int tilesN = g.N / 16, tiles = (g.M / 16) * tilesN;
    if (wave >= tiles * g.kSplit) return;
    const int t = wave % tiles, sidx = wave / tiles;
    const int tm = t / tilesN, tn = t % tilesN;
    const int kChunk = ((g.K / 16 + g.kSplit - 1) / g.kSplit) * 16;
    const int k0 = sidx * kChunk, k1 = min(g.K, k0 + kChunk);
    const int i = tm * 16 + (lane & 15), j = tn * 16 + (lane & 15), kg = 4 * (lane >> 4);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k = k0; k < k1; k += 16) {
        v4s a, b;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int kk = k + kg + v;
            a[v] = static_cast<short>(g.transA ? g.A[static_cast<size_t>(kk) * g.lda + i]
                                               : g.A[static_cast<size_t>(i) * g.lda + kk]);
            b[v] = static_cast<short>(g.transB ? g.B[static_cast<size_t>(j) * g.ldb + kk]
                                               : g.B[static_cast<size_t>(kk) * g.ldb + j]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        float* c = g.C + static_cast<size_t>(tm * 16 + 4 * (lane >> 4) + v) * g.ldc + tn * 16 + (lane & 15);
        if (g.atomic) atomicAdd(c, acc[v]);
        else *c = acc[v];
    }
}

__global__ void k_f32_to_bf16(uint64_t n, const float* __restrict__ in, uint16_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = f2bf(in[i]);
}

__global__ void k_zero_f32(uint64_t n, float* p) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = 0.f;
}

// out = act(C + bias) as BF16 (and FP32 into C); rows >= valid are zeroed.
__global__ void k_bias_act(uint64_t rows, uint64_t valid, int cols, float* __restrict__ C, const float* __restrict__ bias,
                           int relu, uint16_t* __restrict__ out) {
    const uint64_t n = rows * cols;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t r = e / cols;
        float v = C[e] + bias[e % cols];
        if (relu) v = fmaxf(v, 0.f);
        if (r >= valid) v = 0.f;
        C[e] = v;
        if (out) out[e] = f2bf(v);
    }
}

// Weighted MSE over the first kOut columns: loss += sum w_c (y - t)^2 / n,
// dY = 2 w_c (y - t) / n as BF16 (0 in padding rows and columns).
__global__ void k_mse_grad(uint64_t rows, uint64_t valid, const float* __restrict__ Y, const float* __restrict__ T,
                           const float* __restrict__ w, uint16_t* __restrict__ dY, float* loss) {
    float acc = 0.f;
    const uint64_t n = rows * kOutPad;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t r = e / kOutPad;
        const int c = static_cast<int>(e % kOutPad);
        float g = 0.f;
        if (r < valid && c < kOut) {
            const float d = Y[e] - T[r * kOut + c];
            acc += w[c] * d * d / static_cast<float>(valid);
            g = 2.f * w[c] * d / static_cast<float>(valid);
        }
        dY[e] = f2bf(g);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && loss) atomicAdd(loss, acc);
}

// dH = G * [H > 0] as BF16.
__global__ void k_relu_grad(uint64_t n, const float* __restrict__ G, const uint16_t* __restrict__ H,
                            uint16_t* __restrict__ dH) {
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        dH[e] = bf2f(H[e]) > 0.f ? f2bf(G[e]) : static_cast<uint16_t>(0);
}

// db[c] += sum over rows of dH[r][c] (BF16 input).
__global__ void k_col_sum(uint64_t rows, int cols, const uint16_t* __restrict__ dH, float* db) {
    const int c = threadIdx.x % cols;
    const int lanesPerCol = blockDim.x / cols;
    const int sub = threadIdx.x / cols;
    float acc = 0.f;
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(lanesPerCol) + sub; r < rows;
         r += static_cast<uint64_t>(gridDim.x) * lanesPerCol)
        acc += bf2f(dH[r * cols + c]);
    if (sub < lanesPerCol) atomicAdd(&db[c], acc);
}

__global__ void k_adam(uint64_t n, float* __restrict__ W, const float* __restrict__ G, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const float g = G[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        W[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Features of KNN queries (see the file comment); X is [nq x 48] FP32.
__global__ void k_knn_features(uint64_t nq, int D, const double* __restrict__ points, const uint64_t* __restrict__ Ks,
                               uint64_t n0, const double* __restrict__ coords0, const double* __restrict__ radius0,
                               double vmin, double scale, float* __restrict__ X) {
    constexpr int kTop = 31;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; q < nq;
         q += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const double* p = points + q * D;
        float* x = X + q * kIn;
        for (int d = 0; d < 16; ++d) x[d] = d < D ? static_cast<float>((p[d] - vmin) / scale) : 0.f;
        double best[kTop];
        int nb = 0;
        for (uint64_t c = 0; c < n0; ++c) {
            double acc = 0.0;
            for (int d = 0; d < D; ++d) {
                const double diff = p[d] - coords0[c * D + d];
                acc = fma(diff, diff, acc);
            }
            double lb = sqrt(acc) - radius0[c];
            if (nb < kTop) {
                int k = nb++;
                while (k > 0 && best[k - 1] > lb) {
                    best[k] = best[k - 1];
                    --k;
                }
                best[k] = lb;
            } else if (lb < best[kTop - 1]) {
                int k = kTop - 1;
                while (k > 0 && best[k - 1] > lb) {
                    best[k] = best[k - 1];
                    --k;
                }
                best[k] = lb;
            }
        }
        for (int k = 0; k < kTop; ++k) x[16 + k] = k < nb ? static_cast<float>(best[k] / scale) : (nb ? x[16 + nb - 1] : 0.f);
        x[47] = static_cast<float>(log(static_cast<double>(Ks[q] > 0 ? Ks[q] : 1)) / log(256.0));
    }
}

// Proposed radius from the network output: scale * exp(y0) * factor.
__global__ void k_radius_from_output(uint64_t nq, const float* __restrict__ Y, double scale, double factor,
                                     double* __restrict__ radius) {
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; q < nq;
         q += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        radius[q] = scale * exp(static_cast<double>(Y[q * kOutPad])) * factor;
}

}  // namespace tblaeq

using namespace tblaeq;

struct tblaeq_mlp {
    tblaeq_ctx* ctx = nullptr;  // private context (own stream) on the creating context's device
    double vmin = 0.0, scale = 1.0;  // feature normalisation (domain origin and extent)
    float* W = nullptr;   // FP32 master parameters [kParams]
    float* G = nullptr;   // gradients
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* Wb = nullptr;  // BF16 W1 | W2 | W3
    uint64_t step = 0;
    // activations for up to cap rows
    uint64_t cap = 0;
    uint16_t *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dY = nullptr, *dH = nullptr;
    float *C1 = nullptr, *C2 = nullptr, *Y = nullptr, *Gt = nullptr, *loss = nullptr, *lw = nullptr;
    std::vector<void*> owned;
};

namespace {

int mlp_alloc(tblaeq_mlp* m, void** p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
        set_error("tblaeq_mlp: device allocation failed");
        return TBLAEQ_ERR_NOMEM;
    }
    m->owned.push_back(*p);
    return TBLAEQ_OK;
}

unsigned grid_for(const tblaeq_mlp* m, uint64_t n) {
    const uint64_t g = (n + 255) / 256, cap = static_cast<uint64_t>(m->ctx->cuCount) * 8;
    return static_cast<unsigned>(g < 1 ? 1 : (g < cap ? g : cap));
}

int ensure_rows(tblaeq_mlp* m, uint64_t rows) {
    if (rows <= m->cap) return TBLAEQ_OK;
    TB_HIP(hipStreamSynchronize(m->ctx->stream));
    // grow: old activation buffers stay owned until destroy (small)
    int rc;
    if ((rc = mlp_alloc(m, reinterpret_cast<void**>(&m->X), rows * kIn * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->H1), rows * kH * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->H2), rows * kH * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->dY), rows * kOutPad * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->dH), rows * kH * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->C1), rows * kH * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->C2), rows * kH * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->Y), rows * kOutPad * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->Gt), rows * kH * 4)))
        return rc;
    m->cap = rows;
    return TBLAEQ_OK;
}

void gemm(tblaeq_mlp* m, int M, int N, int K, const uint16_t* A, int lda, int tA, const uint16_t* B, int ldb, int tB,
          float* C, int ldc, int atomic) {
    GemmArgs g{M, N, K, A, lda, tA, B, ldb, tB, C, ldc, atomic, 1};
    const int tiles = (M / 16) * (N / 16);
    if (atomic) {
        // split the (long) K dimension so that about 2048 waves share it
        int split = 2048 / (tiles > 0 ? tiles : 1);
        const int kt = K / 16;
        if (split > kt) split = kt;
        if (split < 1) split = 1;
        g.kSplit = split;
    }
    const int waves = tiles * g.kSplit;
    TimedScope ts(m->ctx, kKBatch0 + 11);
    hipLaunchKernelGGL(k_gemm_bf16, dim3((waves + 3) / 4), dim3(256), 0, m->ctx->stream, g);
}

void refresh_bf16(tblaeq_mlp* m) {
    hipStream_t s = m->ctx->stream;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(m, kIn * kH)), dim3(256), 0, s, static_cast<uint64_t>(kIn * kH),
                       m->W + kOffW1, m->Wb);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(m, kH * kH)), dim3(256), 0, s, static_cast<uint64_t>(kH * kH),
                       m->W + kOffW2, m->Wb + kIn * kH);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(m, kH * kOutPad)), dim3(256), 0, s,
                       static_cast<uint64_t>(kH * kOutPad), m->W + kOffW3, m->Wb + kIn * kH + kH * kH);
}

// Forward over rows (multiple of 16) of the FP32 features Xf; valid rows < rows are real.
void forward(tblaeq_mlp* m, const float* Xf, uint64_t rows, uint64_t valid) {
    hipStream_t s = m->ctx->stream;
    const uint16_t* W1 = m->Wb;
    const uint16_t* W2 = m->Wb + kIn * kH;
    const uint16_t* W3 = m->Wb + kIn * kH + kH * kH;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(m, rows * kIn)), dim3(256), 0, s, rows * kIn, Xf, m->X);
    gemm(m, static_cast<int>(rows), kH, kIn, m->X, kIn, 0, W1, kH, 0, m->C1, kH, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(m, rows * kH)), dim3(256), 0, s, rows, valid, kH, m->C1,
                       m->W + kOffB1, 1, m->H1);
    gemm(m, static_cast<int>(rows), kH, kH, m->H1, kH, 0, W2, kH, 0, m->C2, kH, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(m, rows * kH)), dim3(256), 0, s, rows, valid, kH, m->C2,
                       m->W + kOffB2, 1, m->H2);
    gemm(m, static_cast<int>(rows), kOutPad, kH, m->H2, kH, 0, W3, kOutPad, 0, m->Y, kOutPad, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(m, rows * kOutPad)), dim3(256), 0, s, rows, valid, kOutPad, m->Y,
                       m->W + kOffB3, 0, static_cast<uint16_t*>(nullptr));
}

// Gradients of the weighted MSE for rows [0, rows) into m->G (zeroed first).
void backward(tblaeq_mlp* m, const float* Tf, uint64_t rows, uint64_t valid) {
    hipStream_t s = m->ctx->stream;
    const uint16_t* W2 = m->Wb + kIn * kH;
    const uint16_t* W3 = m->Wb + kIn * kH + kH * kH;
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(m, kParams)), dim3(256), 0, s, static_cast<uint64_t>(kParams), m->G);
    hipLaunchKernelGGL(k_mse_grad, dim3(grid_for(m, rows * kOutPad)), dim3(256), 0, s, rows, valid, m->Y, Tf, m->lw,
                       m->dY, m->loss);
    const int R = static_cast<int>(rows);
    // layer 3
    gemm(m, kH, kOutPad, R, m->H2, kH, 1, m->dY, kOutPad, 0, m->G + kOffW3, kOutPad, 1);
    hipLaunchKernelGGL(k_col_sum, dim3(grid_for(m, rows)), dim3(256), 0, s, rows, kOutPad, m->dY, m->G + kOffB3);
    gemm(m, R, kH, kOutPad, m->dY, kOutPad, 0, W3, kOutPad, 1, m->Gt, kH, 0);
    hipLaunchKernelGGL(k_relu_grad, dim3(grid_for(m, rows * kH)), dim3(256), 0, s, rows * kH, m->Gt, m->H2, m->dH);
    // layer 2
    gemm(m, kH, kH, R, m->H1, kH, 1, m->dH, kH, 0, m->G + kOffW2, kH, 1);
    hipLaunchKernelGGL(k_col_sum, dim3(grid_for(m, rows)), dim3(256), 0, s, rows, kH, m->dH, m->G + kOffB2);
    gemm(m, R, kH, kH, m->dH, kH, 0, W2, kH, 1, m->Gt, kH, 0);
    hipLaunchKernelGGL(k_relu_grad, dim3(grid_for(m, rows * kH)), dim3(256), 0, s, rows * kH, m->Gt, m->H1, m->dH);
    // layer 1
    gemm(m, kIn, kH, R, m->X, kIn, 1, m->dH, kH, 0, m->G + kOffW1, kH, 1);
    hipLaunchKernelGGL(k_col_sum, dim3(grid_for(m, rows)), dim3(256), 0, s, rows, kH, m->dH, m->G + kOffB1);
}

uint64_t round16(uint64_t n) { return (n + 15) / 16 * 16; }

}  // namespace

extern "C" {

uint64_t tblaeq_mlp_param_count(void) { return kParams; }

int tblaeq_mlp_create(tblaeq_ctx* ctx, uint64_t seed, tblaeq_mlp** out) {
    if (!ctx || !out) {
        set_error("tblaeq_mlp_create: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    auto* m = new tblaeq_mlp();
    int rc = tblaeq_ctx_create(ctx->device, nullptr, &m->ctx);
    if (rc) {
        delete m;
        return rc;
    }
    if ((rc = mlp_alloc(m, reinterpret_cast<void**>(&m->W), kParams * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->G), kParams * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->m), kParams * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->v), kParams * 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->Wb), (kIn * kH + kH * kH + kH * kOutPad) * 2)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->loss), 4)) ||
        (rc = mlp_alloc(m, reinterpret_cast<void**>(&m->lw), kOut * 4))) {
        tblaeq_mlp_destroy(m);
        return rc;
    }
    // He-uniform initialisation from a 64-bit LCG (weights), zero biases
    std::vector<float> w(kParams, 0.f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<float>((st >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
    };
    for (int i = 0; i < kIn * kH; ++i) w[kOffW1 + i] = uni() * std::sqrt(6.f / kIn);
    for (int i = 0; i < kH * kH; ++i) w[kOffW2 + i] = uni() * std::sqrt(6.f / kH);
    for (int r = 0; r < kH; ++r)
        for (int c = 0; c < kOutPad; ++c) w[kOffW3 + r * kOutPad + c] = c < kOut ? uni() * std::sqrt(6.f / kH) : 0.f;
    const float lw[kOut] = {1.f, 1.f, 1.f, 1.f};
    if (hipMemcpy(m->W, w.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->lw, lw, sizeof(lw), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m, 0, kParams * 4) != hipSuccess || hipMemset(m->v, 0, kParams * 4) != hipSuccess) {
        tblaeq_mlp_destroy(m);
        set_error("tblaeq_mlp_create: upload failed");
        return TBLAEQ_ERR_HIP;
    }
    refresh_bf16(m);
    *out = m;
    return TBLAEQ_OK;
}

int tblaeq_mlp_destroy(tblaeq_mlp* m) {
    if (!m) return TBLAEQ_OK;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    for (void* p : m->owned) (void)hipFree(p);
    if (m->ctx) tblaeq_ctx_destroy(m->ctx);
    delete m;
    return TBLAEQ_OK;
}

int tblaeq_mlp_set_params(tblaeq_mlp* m, const float* params) {
    if (!m || !params) return TBLAEQ_ERR_INVALID;
    TB_HIP(hipMemcpyAsync(m->W, params, kParams * 4, hipMemcpyHostToDevice, m->ctx->stream));
    TB_HIP(hipMemsetAsync(m->m, 0, kParams * 4, m->ctx->stream));
    TB_HIP(hipMemsetAsync(m->v, 0, kParams * 4, m->ctx->stream));
    m->step = 0;
    refresh_bf16(m);
    TB_HIP(hipStreamSynchronize(m->ctx->stream));
    return TBLAEQ_OK;
}

int tblaeq_mlp_get_params(tblaeq_mlp* m, float* params) {
    if (!m || !params) return TBLAEQ_ERR_INVALID;
    TB_HIP(hipMemcpyAsync(params, m->W, kParams * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    TB_HIP(hipStreamSynchronize(m->ctx->stream));
    return TBLAEQ_OK;
}

int tblaeq_mlp_set_norm(tblaeq_mlp* m, double vmin, double scale) {
    if (!m || !(scale > 0.0)) return TBLAEQ_ERR_INVALID;
    m->vmin = vmin;
    m->scale = scale;
    return TBLAEQ_OK;
}

int tblaeq_mlp_get_norm(const tblaeq_mlp* m, double* vmin, double* scale) {
    if (!m || !vmin || !scale) return TBLAEQ_ERR_INVALID;
    *vmin = m->vmin;
    *scale = m->scale;
    return TBLAEQ_OK;
}

int tblaeq_mlp_set_loss_weights(tblaeq_mlp* m, const float* w4) {
    if (!m || !w4) return TBLAEQ_ERR_INVALID;
    TB_HIP(hipMemcpyAsync(m->lw, w4, kOut * 4, hipMemcpyHostToDevice, m->ctx->stream));
    TB_HIP(hipStreamSynchronize(m->ctx->stream));
    return TBLAEQ_OK;
}

int tblaeq_mlp_forward_device(tblaeq_mlp* m, const float* d_X, uint64_t n, float* d_Y16) {
    if (!m || !d_X || n == 0) return TBLAEQ_ERR_INVALID;
    const uint64_t rows = round16(n);
    int rc = ensure_rows(m, rows);
    if (rc) return rc;
    // rows beyond n read zero-padded features
    // stage the features zero-padded to whole 16-row tiles in C1 ([rows x 64]
    // holds [rows x 48]); forward() converts them to BF16 before C1 is reused
    float* Xp = m->C1;
    TB_HIP(hipMemsetAsync(Xp, 0, rows * kIn * 4, m->ctx->stream));
    TB_HIP(hipMemcpyAsync(Xp, d_X, n * kIn * 4, hipMemcpyDeviceToDevice, m->ctx->stream));
    forward(m, Xp, rows, n);
    if (d_Y16) TB_HIP(hipMemcpyAsync(d_Y16, m->Y, n * kOutPad * 4, hipMemcpyDeviceToDevice, m->ctx->stream));
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

int tblaeq_mlp_forward(tblaeq_mlp* m, const float* X, uint64_t n, float* Y) {
    if (!m || !X || !Y || n == 0) return TBLAEQ_ERR_INVALID;
    float *dX = nullptr, *dY = nullptr;
    if (hipMalloc(&dX, n * kIn * 4) != hipSuccess || hipMalloc(&dY, n * kOutPad * 4) != hipSuccess) {
        if (dX) (void)hipFree(dX);
        set_error("tblaeq_mlp_forward: device allocation failed");
        return TBLAEQ_ERR_NOMEM;
    }
    int rc = TBLAEQ_OK;
    std::vector<float> y16(n * kOutPad);
    if (hipMemcpyAsync(dX, X, n * kIn * 4, hipMemcpyHostToDevice, m->ctx->stream) != hipSuccess) rc = TBLAEQ_ERR_HIP;
    if (!rc) rc = tblaeq_mlp_forward_device(m, dX, n, dY);
    if (!rc && (hipMemcpyAsync(y16.data(), dY, n * kOutPad * 4, hipMemcpyDeviceToHost, m->ctx->stream) != hipSuccess ||
                hipStreamSynchronize(m->ctx->stream) != hipSuccess))
        rc = TBLAEQ_ERR_HIP;
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (rc) return rc;
    for (uint64_t i = 0; i < n; ++i) std::memcpy(Y + i * kOut, y16.data() + i * kOutPad, kOut * 4);
    return TBLAEQ_OK;
}

// Gradients of the weighted MSE over n rows (host X [n x 48], T [n x 4]) into
// grads [kParams] (host); *loss receives the loss.  Weights unchanged.
int tblaeq_mlp_gradients(tblaeq_mlp* m, const float* X, const float* T, uint64_t n, float* grads, float* loss) {
    if (!m || !X || !T || n == 0) return TBLAEQ_ERR_INVALID;
    const uint64_t rows = round16(n);
    int rc = ensure_rows(m, rows);
    if (rc) return rc;
    float *dX = nullptr, *dT = nullptr;
    if (hipMalloc(&dX, rows * kIn * 4) != hipSuccess || hipMalloc(&dT, rows * kOut * 4) != hipSuccess) {
        if (dX) (void)hipFree(dX);
        set_error("tblaeq_mlp_gradients: device allocation failed");
        return TBLAEQ_ERR_NOMEM;
    }
    hipStream_t s = m->ctx->stream;
    bool ok = hipMemsetAsync(dX, 0, rows * kIn * 4, s) == hipSuccess &&
              hipMemcpyAsync(dX, X, n * kIn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemsetAsync(dT, 0, rows * kOut * 4, s) == hipSuccess &&
              hipMemcpyAsync(dT, T, n * kOut * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemsetAsync(m->loss, 0, 4, s) == hipSuccess;
    if (ok) {
        forward(m, dX, rows, n);
        backward(m, dT, rows, n);
        ok = hipGetLastError() == hipSuccess;
        if (ok && grads) ok = hipMemcpyAsync(grads, m->G, kParams * 4, hipMemcpyDeviceToHost, s) == hipSuccess;
        if (ok && loss) ok = hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, s) == hipSuccess;
        ok = ok && hipStreamSynchronize(s) == hipSuccess;
    }
    (void)hipFree(dX);
    (void)hipFree(dT);
    if (!ok) {
        set_error("tblaeq_mlp_gradients: HIP error");
        return TBLAEQ_ERR_HIP;
    }
    return TBLAEQ_OK;
}

// Adam training on n rows (host X [n x 48], T [n x 4]) in mini-batches of
// `batch` rows (a multiple of 16), in the given row order, for `epochs`
// passes; loss_per_epoch[e] (host, optional) = mean mini-batch loss.
int tblaeq_mlp_train(tblaeq_mlp* m, const float* X, const float* T, uint64_t n, uint64_t epochs, uint64_t batch,
                     float lr, float* loss_per_epoch) {
    if (!m || !X || !T || n == 0 || batch == 0 || batch % 16) {
        set_error("tblaeq_mlp_train: invalid argument (batch must be a positive multiple of 16)");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t rows = round16(n);
    int rc = ensure_rows(m, batch);
    if (rc) return rc;
    float *dX = nullptr, *dT = nullptr, *dLoss = nullptr;
    if (hipMalloc(&dX, rows * kIn * 4) != hipSuccess || hipMalloc(&dT, rows * kOut * 4) != hipSuccess ||
        hipMalloc(&dLoss, (epochs ? epochs : 1) * 4) != hipSuccess) {
        if (dX) (void)hipFree(dX);
        if (dT) (void)hipFree(dT);
        set_error("tblaeq_mlp_train: device allocation failed");
        return TBLAEQ_ERR_NOMEM;
    }
    hipStream_t s = m->ctx->stream;
    bool ok = hipMemsetAsync(dX, 0, rows * kIn * 4, s) == hipSuccess &&
              hipMemcpyAsync(dX, X, n * kIn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemsetAsync(dT, 0, rows * kOut * 4, s) == hipSuccess &&
              hipMemcpyAsync(dT, T, n * kOut * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<float> losses(epochs, 0.f);
    for (uint64_t e = 0; ok && e < epochs; ++e) {
        ok = hipMemsetAsync(m->loss, 0, 4, s) == hipSuccess;
        uint64_t nb = 0;
        for (uint64_t b0 = 0; ok && b0 < n; b0 += batch, ++nb) {
            const uint64_t valid = n - b0 < batch ? n - b0 : batch;
            const uint64_t r = round16(valid);
            forward(m, dX + b0 * kIn, r, valid);
            backward(m, dT + b0 * kOut, r, valid);
            ++m->step;
            const float c1 = 1.f - std::pow(b1, static_cast<float>(m->step));
            const float c2 = 1.f - std::pow(b2, static_cast<float>(m->step));
            hipLaunchKernelGGL(k_adam, dim3(grid_for(m, kParams)), dim3(256), 0, s, static_cast<uint64_t>(kParams),
                               m->W, m->G, m->m, m->v, lr, b1, b2, eps, c1, c2);
            refresh_bf16(m);
        }
        if (ok) ok = hipMemcpyAsync(dLoss + e, m->loss, 4, hipMemcpyDeviceToDevice, s) == hipSuccess;
        if (ok && loss_per_epoch) losses[e] = static_cast<float>(nb);
    }
    if (ok && loss_per_epoch && epochs) {
        std::vector<float> l(epochs);
        ok = hipMemcpyAsync(l.data(), dLoss, epochs * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        for (uint64_t e = 0; ok && e < epochs; ++e) loss_per_epoch[e] = l[e] / losses[e];
    }
    ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(dX);
    (void)hipFree(dT);
    (void)hipFree(dLoss);
    if (!ok) {
        set_error("tblaeq_mlp_train: HIP error");
        return TBLAEQ_ERR_HIP;
    }
    return TBLAEQ_OK;
}

int tblaeq_knn_features(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, double vmin,
                        double scale, float* d_X) {
    if (!b || !d_points || !d_K || !d_X || nq == 0 || !(scale > 0.0)) {
        set_error("tblaeq_knn_features: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t g = std::min<uint64_t>((nq + 63) / 64, static_cast<uint64_t>(b->ctx->cuCount) * 8);
    hipLaunchKernelGGL(k_knn_features, dim3(static_cast<unsigned>(g ? g : 1)), dim3(64), 0, b->ctx->stream, nq, b->D,
                       d_points, d_K, b->count[0], b->coords[0], b->radiusPos[0], vmin, scale, d_X);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

int tblaeq_mlp_propose_radius(tblaeq_mlp* m, const float* d_X, uint64_t n, double scale, double factor,
                              double* d_radius) {
    if (!m || !d_X || !d_radius || n == 0) return TBLAEQ_ERR_INVALID;
    const uint64_t rows = round16(n);
    int rc = ensure_rows(m, rows);
    if (rc) return rc;
    float* Xp = m->C1;
    TB_HIP(hipMemsetAsync(Xp, 0, rows * kIn * 4, m->ctx->stream));
    TB_HIP(hipMemcpyAsync(Xp, d_X, n * kIn * 4, hipMemcpyDeviceToDevice, m->ctx->stream));
    forward(m, Xp, rows, n);
    hipLaunchKernelGGL(k_radius_from_output, dim3(grid_for(m, n)), dim3(256), 0, m->ctx->stream, n, m->Y, scale,
                       factor, d_radius);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

}  // extern "C"
