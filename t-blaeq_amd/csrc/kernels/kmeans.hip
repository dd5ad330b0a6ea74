// GPU index build: k-means levels (k-means++ seeding, Lloyd iterations) and the
// CSC P-tensor of each level.
//
// Assignment.  The squared distance ranking of the centers for a point x is
// the ranking of s(c) = ||c||^2 - 2 x.c (||x||^2 is common to all centers).
// A workgroup takes 256 points (64 per wave, four 16-row tiles) and streams the
// centers through LDS 64 at a time; each 16 x 16 (points x centers) tile is
// D/4 v_mfma_f32_16x16x4f32 instructions with the accumulator seeded with
// ||c||^2 and the A operand holding -2x, so the MFMA result is s(c) itself.
// Every lane keeps, for its four rows of each tile, the best and second-best
// score and the best index; the 16 lanes of a row group are merged at the end.
// Coordinates are centred on the data's mid-range and rounded to FP32; the
// FP32 score of any center is within E(x) = g (|x| + max|c|)^2 of the exact
// one (g = 2 (D + 8) 2^-24: input rounding, norms, the dot product's
// accumulation), so when the second-best score exceeds the best by more than
// 2 E(x) the FP32 winner is the exact winner.  Other points (near ties) are
// re-scored in FP64 with the CPU build's arithmetic over every center, strict
// < with the lowest index winning, which makes the labels the CPU build's.
//
// Update.  (label, point) keys are sorted stably by label, and one thread per
// (cluster, dimension) sums the cluster's coordinates in point order, the
// order in which the CPU build accumulates them (CPUKmeans.cuh:160-205), so
// the new centroids are bit-identical for identical labels.
//
// Seeding.  k-means++ as in the CPU build (CPUKmeans.cuh:67-118): the random
// draws come from the same mt19937_64 stream on the host; the min distances
// and their cumulative sums live on the device (per 1024-point tile sums, a
// block scan over tiles, a scan inside the chosen tile).
#include <algorithm>
#include <cstring>
#include <cmath>
#include <random>
#include <vector>

#include "../../../include/tblaeq_build.h"
#include "../batch_internal.hpp"
#include "../scan_common.hpp"

namespace tblaeq {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kKmThreads = 256;
constexpr int kKmPoints = 256;   // points per assignment workgroup
constexpr int kKmChunk = 128;    // centers per LDS chunk
constexpr int kPpTile = 1024;    // points per seeding tile sum

// Squared distance with the CPU build's arithmetic (no contraction: the
// library is compiled with -ffp-contract=off).
__device__ __forceinline__ double sq64(const double* a, const double* b, int D) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        const double df = a[d] - b[d];
        const double p = df * df;
        s = s + p;
    }
    return s;
}

// x32[i][d] = float(x[i][d] - mid[d]) (zero padding up to DP), xn[i] = |x32[i]|.
__global__ void k_km_points(const double* __restrict__ x, uint64_t n, int D, int DP, const double* __restrict__ mid,
                            float* __restrict__ x32, float* __restrict__ xn) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        float s = 0.f;
        for (int d = 0; d < DP; ++d) {
            const float v = d < D ? static_cast<float>(x[i * D + d] - mid[d]) : 0.f;
            x32[i * DP + d] = v;
            s += v * v;
        }
        xn[i] = __builtin_sqrtf(s);
    }
}

// Centers in the same form; cn[j] = |c32[j]|^2, *cmax = max |c32[j]| (as float bits).
__global__ void k_km_centers(const double* __restrict__ c, uint64_t k, int D, int DP, const double* __restrict__ mid,
                             float* __restrict__ c32, float* __restrict__ cn, uint32_t* cmax) {
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < k;
         j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        float s = 0.f;
        for (int d = 0; d < DP; ++d) {
            const float v = d < D ? static_cast<float>(c[j * D + d] - mid[d]) : 0.f;
            c32[j * DP + d] = v;
            s += v * v;
        }
        cn[j] = s;
        atomicMax(cmax, __float_as_uint(__builtin_sqrtf(s)));
    }
}

struct AssignArgs {
    uint64_t n, k;
    const float* x32;
    const float* xn;
    const float* c32;
    const float* cn;
    const uint32_t* cmax;
    float g;              // error-bound factor
    uint64_t* labels;     // in: previous labels; out: new labels of decided points
    uint64_t* amb;        // points left for the FP64 re-score
    uint64_t* nAmb;
    uint64_t* changed;
};

// Best and second-best score with the best's index, centres arriving in
// increasing index order (strict <: the lowest index keeps a tie).  With
// b1 <= b2: b2' = min(b2, max(b1, v)) and b1' = min(b1, v) -- one compare, one
// select and three min / max, the same values as the branching form.
__device__ __forceinline__ void keep_best(float v, uint32_t j, float& b1, float& b2, uint32_t& i1) {
    i1 = v < b1 ? j : i1;
    b2 = fminf(b2, fmaxf(b1, v));
    b1 = fminf(b1, v);
}

template <int DP>
__global__ __launch_bounds__(kKmThreads) void k_km_assign(AssignArgs A) {
    constexpr int KS = DP / 4;  // MFMA k-steps
    __shared__ float sc[kKmChunk][DP + 1];
    __shared__ float sn[kKmChunk];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int li = lane & 15, lk = lane >> 4;
    const uint64_t p0 = blockIdx.x * static_cast<uint64_t>(kKmPoints) + static_cast<uint64_t>(wid) * 64;
    // A operands: -2 x of row li of each 16-row tile, dimension 4 s + lk
    float a[4][KS];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint64_t p = p0 + 16 * t + li;
#pragma unroll
        for (int s = 0; s < KS; ++s) a[t][s] = p < A.n ? -2.f * A.x32[p * DP + 4 * s + lk] : 0.f;
    }
    float b1[4][4], b2[4][4];
    uint32_t i1[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            b1[t][v] = __builtin_inff();
            b2[t][v] = __builtin_inff();
            i1[t][v] = 0;
        }
    for (uint64_t c0 = 0; c0 < A.k; c0 += kKmChunk) {
        __syncthreads();
        for (int e = tid; e < kKmChunk * DP; e += kKmThreads) {
            const int j = e / DP, d = e % DP;
            sc[j][d] = c0 + j < A.k ? A.c32[(c0 + j) * DP + d] : 0.f;
        }
        if (tid < kKmChunk) sn[tid] = c0 + tid < A.k ? A.cn[c0 + tid] : __builtin_inff();
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kKmChunk / 16; ++u) {
            const float nrm = sn[16 * u + li];
            float bop[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) bop[s] = sc[16 * u + li][4 * s + lk];
            const uint32_t j = static_cast<uint32_t>(c0) + 16 * u + li;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4f acc = {nrm, nrm, nrm, nrm};
#pragma unroll
                for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][s], bop[s], acc, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; ++v) keep_best(acc[v], j, b1[t][v], b2[t][v], i1[t][v]);
            }
        }
    }
    // merge the 16 lanes that hold the same rows (same lk)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float x1 = b1[t][v], x2 = b2[t][v];
            uint32_t xi = i1[t][v];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const float y1 = __shfl_xor(x1, o, kWave), y2 = __shfl_xor(x2, o, kWave);
                const uint32_t yi = __shfl_xor(xi, o, kWave);
                const bool takeY = y1 < x1 || (y1 == x1 && yi < xi);
                const float other = takeY ? x1 : y1;
                if (takeY) {
                    x1 = y1;
                    xi = yi;
                }
                x2 = fminf(other, fminf(x2, y2));
            }
            b1[t][v] = x1;
            b2[t][v] = x2;
            i1[t][v] = xi;
        }
    const float cm = __uint_as_float(*A.cmax);
    uint32_t changed = 0;
    if (li == 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint64_t p = p0 + 16 * t + 4 * lk + v;
                if (p >= A.n) continue;
                const float r = A.xn[p] + cm;
                const float e = A.g * r * r;
                if (b2[t][v] - b1[t][v] > 2.f * e) {
                    const uint64_t lab = i1[t][v];
                    if (A.labels[p] != lab) {
                        A.labels[p] = lab;
                        ++changed;
                    }
                } else {
                    const uint64_t slot = atomicAdd(reinterpret_cast<unsigned long long*>(A.nAmb), 1ull);
                    A.amb[slot] = p;
                }
            }
    }
    for (int o = 32; o > 0; o >>= 1) changed += __shfl_xor(changed, o, kWave);
    if (lane == 0 && changed) atomicAdd(reinterpret_cast<unsigned long long*>(A.changed), changed);
}

// FP64 re-score of the undecided points: a wave per point over every center.
__global__ void k_km_rescore(const double* __restrict__ x, int D, const double* __restrict__ c, uint64_t k,
                             const uint64_t* __restrict__ amb, const uint64_t* nAmb, uint64_t* __restrict__ labels,
                             uint64_t* changed) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t w0 = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
    const uint64_t na = *nAmb;
    for (uint64_t a = w0; a < na; a += nw) {
        const uint64_t p = amb[a];
        const double* xv = x + p * D;
        double best = __DBL_MAX__;
        uint64_t bi = 0;
        for (uint64_t j = lane; j < k; j += kWave) {
            const double s = sq64(xv, c + j * D, D);
            if (s < best) {
                best = s;
                bi = j;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, kWave);
            const uint64_t oi = __shfl_xor(bi, o, kWave);
            if (ob < best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0 && labels[p] != bi) {
            labels[p] = bi;
            atomicAdd(reinterpret_cast<unsigned long long*>(changed), 1ull);
        }
    }
}

__global__ void k_km_keys(const uint64_t* __restrict__ labels, uint64_t n, int idxBits, uint64_t* __restrict__ keys,
                          uint64_t* __restrict__ cnt) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        keys[i] = (labels[i] << idxBits) | i;
        atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[labels[i]]), 1ull);
    }
}

// New centroid coordinate (cluster c, dimension d): the cluster's points summed
// in point order, divided by their count; an empty cluster keeps its centroid.
__global__ void k_km_update(const double* __restrict__ x, int D, const uint64_t* __restrict__ sorted, uint64_t idxMask,
                            const uint64_t* __restrict__ off, uint64_t k, double* __restrict__ c) {
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < k * D;
         t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t cl = t / D;
        const int d = static_cast<int>(t % D);
        const uint64_t b = off[cl], e = off[cl + 1];
        if (b == e) continue;
        double s = 0.0;
        for (uint64_t p = b; p < e; ++p) s = s + x[(sorted[p] & idxMask) * D + d];
        c[t] = s / static_cast<double>(e - b);
    }
}

// Seeding: minD[i] = min(minD[i], |x_i - last|^2); tileSum[t] = sum of minD over
// tile t (256 threads x 4 consecutive points, then a fixed-order tree), each
// term times w[i] when weights are given (weighted k-means++ on k-means||
// candidates).
__global__ __launch_bounds__(256) void k_kmpp_update(const double* __restrict__ x, uint64_t n, int D,
                                                     const double* __restrict__ last, double* __restrict__ minD,
                                                     double* __restrict__ tileSum, const double* __restrict__ w) {
    __shared__ double s_w[4];
    __shared__ double s_last[64];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    if (tid < D) s_last[tid] = last[tid];
    __syncthreads();
    const uint64_t ntiles = (n + kPpTile - 1) / kPpTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        double s = 0.0;
        for (int r = 0; r < 4; ++r) {
            const uint64_t i = t * kPpTile + static_cast<uint64_t>(tid) * 4 + r;
            if (i < n) {
                const double dd = sq64(x + i * D, s_last, D);
                double m = minD[i];
                if (dd < m) {
                    m = dd;
                    minD[i] = m;
                }
                s = s + (w ? w[i] * m : m);
            }
        }
        for (int o = 1; o < kWave; o <<= 1) s = s + __shfl_xor(s, o, kWave);
        if (lane == 0) s_w[wid] = s;
        __syncthreads();
        if (tid == 0) tileSum[t] = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
        __syncthreads();
    }
}

// Seeding pick (one workgroup of 1024 threads): total = sum of the tile sums,
// thr = u * total, chosen = first point whose running sum of minD reaches thr
// (n - 1 if none), centers[cc] = x[chosen].
__global__ __launch_bounds__(1024) void k_kmpp_pick(const double* __restrict__ x, uint64_t n, int D,
                                                    const double* __restrict__ minD,
                                                    const double* __restrict__ tileSum, const double* __restrict__ u,
                                                    uint64_t cc, double* __restrict__ centers,
                                                    const double* __restrict__ w) {
    __shared__ double s_part[1024];
    __shared__ uint64_t s_tile, s_pick;
    __shared__ double s_thr, s_base;
    __shared__ uint32_t s_found;
    const int tid = threadIdx.x;
    const uint64_t ntiles = (n + kPpTile - 1) / kPpTile;
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t tb = static_cast<uint64_t>(tid) * per, te = std::min<uint64_t>(tb + per, ntiles);
    double local = 0.0;
    for (uint64_t t = tb; t < te; ++t) local = local + tileSum[t];
    s_part[tid] = local;
    if (tid == 0) {
        s_tile = ~0ull;
        s_pick = n - 1;
        s_found = 0;
    }
    __syncthreads();
    // inclusive scan of the thread sums (Hillis-Steele, fixed order)
    for (int o = 1; o < 1024; o <<= 1) {
        const double v = tid >= o ? s_part[tid - o] : 0.0;
        __syncthreads();
        s_part[tid] = s_part[tid] + v;
        __syncthreads();
    }
    if (tid == 0) s_thr = u[cc] * (s_part[1023] - 0.0) + 0.0;
    __syncthreads();
    const double thr = s_thr;
    // the first tile whose running sum reaches thr
    double run = tid ? s_part[tid - 1] : 0.0;
    for (uint64_t t = tb; t < te; ++t) {
        const double nx = run + tileSum[t];
        if (nx >= thr) {
            atomicMin(reinterpret_cast<unsigned long long*>(&s_tile), t);
            break;
        }
        run = nx;
    }
    __syncthreads();
    const uint64_t tile = s_tile;
    if (tile != ~0ull) {
        // running sum before the tile: recompute from the owning thread's prefix
        if (tid == 0) {
            const uint64_t owner = tile / per;
            double r = owner ? s_part[owner - 1] : 0.0;
            for (uint64_t t = owner * per; t < tile; ++t) r = r + tileSum[t];
            s_base = r;
        }
        __syncthreads();
        // scan of the tile's minD (one point per thread)
        const uint64_t i = tile * kPpTile + tid;
        s_part[tid] = i < n ? (w ? w[i] * minD[i] : minD[i]) : 0.0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const double v = tid >= o ? s_part[tid - o] : 0.0;
            __syncthreads();
            s_part[tid] = s_part[tid] + v;
            __syncthreads();
        }
        if (i < n && s_base + s_part[tid] >= thr) {
            atomicMin(reinterpret_cast<unsigned long long*>(&s_pick), i);
            s_found = 1;
        }
        __syncthreads();
        // the tile reached thr but its element scan (another summation order)
        // did not: its last point
        if (tid == 0 && !s_found) s_pick = std::min<uint64_t>(n, (tile + 1) * kPpTile) - 1;
        __syncthreads();
    }
    const uint64_t pick = s_pick;
    if (tid < D) centers[cc * D + tid] = x[pick * D + tid];
}

// ---- k-means|| (Bahmani et al., 2012) --------------------------------------

// Uniform draw in [0, 1) of point i in round r: SplitMix64 of (seed, r, i).
__device__ __forceinline__ double unit_draw(uint64_t seed, uint64_t r, uint64_t i) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (r * 0x100000001B3ull + i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1.0p-53;
}

// minD[i] = |x_i - c0|^2, nearest[i] = 0.
__global__ void k_kmpar_init(const double* __restrict__ x, uint64_t n, int D, const double* __restrict__ c0,
                             double* __restrict__ minD, uint64_t* __restrict__ nearest) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        minD[i] = sq64(x + i * D, c0, D);
        nearest[i] = 0;
    }
}

// Per 1024-point tile, the sum of v[i] (256 threads x 4 consecutive points,
// then a fixed-order tree), and one workgroup's fixed-order sum of the tiles.
__global__ __launch_bounds__(256) void k_tile_sum(const double* __restrict__ v, uint64_t n,
                                                  double* __restrict__ tileSum) {
    __shared__ double s_w[4];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const uint64_t ntiles = (n + kPpTile - 1) / kPpTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        double s = 0.0;
        for (int r = 0; r < 4; ++r) {
            const uint64_t i = t * kPpTile + static_cast<uint64_t>(tid) * 4 + r;
            if (i < n) s = s + v[i];
        }
        for (int o = 1; o < kWave; o <<= 1) s = s + __shfl_xor(s, o, kWave);
        if (lane == 0) s_w[wid] = s;
        __syncthreads();
        if (tid == 0) tileSum[t] = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_sum_tiles(const double* __restrict__ tileSum, uint64_t ntiles,
                                                    double* __restrict__ out) {
    __shared__ double s_part[1024];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (uint64_t t = tid; t < ntiles; t += 1024) s = s + tileSum[t];
    s_part[tid] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (tid < o) s_part[tid] = s_part[tid] + s_part[tid + o];
        __syncthreads();
    }
    if (tid == 0) *out = s_part[0];
}

// Round r's sample: point i is drawn when unit_draw < ell * minD[i] / phi.
__global__ void k_kmpar_mark(const double* __restrict__ minD, uint64_t n, const double* __restrict__ phi, double ell,
                             uint64_t seed, uint64_t round, uint64_t* __restrict__ flag) {
    const double p = *phi;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        flag[i] = (p > 0.0 && unit_draw(seed, round, i) * p < ell * minD[i]) ? 1 : 0;
}

// Drawn points, in point order, appended to the candidates (off = exclusive
// scan of the flags).
__global__ void k_kmpar_gather(const double* __restrict__ x, uint64_t n, int D, const uint64_t* __restrict__ flag,
                               const uint64_t* __restrict__ off, double* __restrict__ cand) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        if (flag[i])
            for (int d = 0; d < D; ++d) cand[off[i] * D + d] = x[i * D + d];
}

// After the nearest of a round's new candidates is found (lab, exact): keep
// it when strictly closer than the point's current candidate.
__global__ void k_kmpar_improve(const double* __restrict__ x, uint64_t n, int D, const double* __restrict__ cand,
                                const uint64_t* __restrict__ lab, uint64_t base, double* __restrict__ minD,
                                uint64_t* __restrict__ nearest) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const double d = sq64(x + i * D, cand + lab[i] * D, D);
        if (d < minD[i]) {
            minD[i] = d;
            nearest[i] = base + lab[i];
        }
    }
}

// Candidate weights: the points whose nearest candidate each one is.
__global__ void k_kmpar_weights(const uint64_t* __restrict__ nearest, uint64_t n, uint64_t* __restrict__ wcnt) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        atomicAdd(reinterpret_cast<unsigned long long*>(&wcnt[nearest[i]]), 1ull);
}

__global__ void k_u64_to_f64(const uint64_t* __restrict__ a, uint64_t n, double* __restrict__ b) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        b[i] = static_cast<double>(a[i]);
}

// v[i] = |x_i - c[lab[i]]|^2 (the inertia terms).
__global__ void k_km_sqdist(const double* __restrict__ x, uint64_t n, int D, const double* __restrict__ c,
                            const uint64_t* __restrict__ lab, double* __restrict__ v) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        v[i] = sq64(x + i * D, c + lab[i] * D, D);
}

__global__ void k_f32_to_f64(const float* __restrict__ a, uint64_t n, double* __restrict__ b) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        b[i] = static_cast<double>(a[i]);
}

// CSC build: map[pos], vals[pos], per-cluster max squared radius (as u64 bits).
__global__ void k_csc_rows(const double* __restrict__ fine, uint64_t n, int D, const double* __restrict__ centers,
                           const uint64_t* __restrict__ sorted, int idxBits, uint64_t* __restrict__ map,
                           double* __restrict__ vals, uint64_t* __restrict__ r2bits) {
    const uint64_t mask = (1ull << idxBits) - 1;
    for (uint64_t pos = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; pos < n;
         pos += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t key = sorted[pos];
        const uint64_t i = key & mask, col = key >> idxBits;
        map[pos] = i;
        const double* f = fine + i * D;
        const double* cv = centers + col * D;
        double s = 0.0;
        for (int d = 0; d < D; ++d) {
            vals[pos * D + d] = f[d] / cv[d];
            const double df = cv[d] - f[d];
            const double sq = df * df;
            s = s + sq;
        }
        atomicMax(reinterpret_cast<unsigned long long*>(&r2bits[col]),
                  static_cast<unsigned long long>(__double_as_longlong(s)));
    }
}

__global__ void k_csc_radius(const uint64_t* __restrict__ r2bits, uint64_t k, double* __restrict__ radius) {
    for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < k;
         c += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        radius[c] = __builtin_sqrt(__longlong_as_double(static_cast<long long>(r2bits[c])));
}

__global__ void k_km_zero(uint64_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = 0;
}

__global__ void k_km_fill_f64(double* p, uint64_t n, double v) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Device scratch of one call, freed on every exit.
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t n, int* rc) {
        void* p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("GPU build: out of device memory");
            *rc = TBLAEQ_ERR_HIP;
            return nullptr;
        }
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

unsigned grid_for(tblaeq_ctx* ctx, uint64_t n, int threads) {
    const uint64_t g = std::min<uint64_t>((n + threads - 1) / threads, static_cast<uint64_t>(ctx->cuCount) * 16);
    return static_cast<unsigned>(g ? g : 1);
}

void launch_assign(int DP, dim3 g, tblaeq_ctx* ctx, const AssignArgs& A) {
    const dim3 b(kKmThreads);
    switch (DP) {
    case 4: hipLaunchKernelGGL(k_km_assign<4>, g, b, 0, ctx->stream, A); break;
    case 8: hipLaunchKernelGGL(k_km_assign<8>, g, b, 0, ctx->stream, A); break;
    case 16: hipLaunchKernelGGL(k_km_assign<16>, g, b, 0, ctx->stream, A); break;
    case 32: hipLaunchKernelGGL(k_km_assign<32>, g, b, 0, ctx->stream, A); break;
    default: hipLaunchKernelGGL(k_km_assign<64>, g, b, 0, ctx->stream, A); break;
    }
}

int padded_dim(uint64_t D) {
    for (int p : {4, 8, 16, 32, 64})
        if (D <= static_cast<uint64_t>(p)) return p;
    return 0;
}

}  // namespace
}  // namespace tblaeq

using namespace tblaeq;

extern "C" void tblaeq_kmeans_options_init(tblaeq_kmeans_options* o) {
    if (!o) return;
    o->max_iter = 7;
    o->seed = 42;
    o->seeding = 0;
    o->oversampling = 0.5;
    o->rounds = 4;
}

extern "C" int tblaeq_kmeans(tblaeq_ctx* ctx, const double* d_x, uint64_t n, uint64_t D, uint64_t k,
                             uint64_t max_iter, uint64_t seed, double* d_centers, uint64_t* d_labels,
                             uint64_t* stats) {
    tblaeq_kmeans_options o;
    tblaeq_kmeans_options_init(&o);
    o.max_iter = max_iter;
    o.seed = seed;
    uint64_t st6[6] = {0, 0, 0, 0, 0, 0};
    const int rc = tblaeq_kmeans_ex(ctx, d_x, n, D, k, &o, d_centers, d_labels, st6);
    if (rc == TBLAEQ_OK && stats) std::copy(st6, st6 + 4, stats);
    return rc;
}

extern "C" int tblaeq_kmeans_ex(tblaeq_ctx* ctx, const double* d_x, uint64_t n, uint64_t D, uint64_t k,
                                const tblaeq_kmeans_options* opt, double* d_centers, uint64_t* d_labels,
                                uint64_t* stats) {
    if (!ctx || !d_x || !d_centers || !d_labels || !opt || n == 0 || D == 0 || k == 0 || k > n ||
        (opt->seeding != 0 && opt->seeding != 1) || (opt->seeding == 1 && (!(opt->oversampling > 0.0) || opt->rounds == 0))) {
        set_error("tblaeq_kmeans: invalid argument (1 <= k <= n; seeding 0 or 1; k-means|| needs oversampling > 0 "
                  "and rounds >= 1)");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t max_iter = opt->max_iter, seed = opt->seed;
    const int DP = padded_dim(D);
    if (DP == 0 || k >= (1ull << 32)) {
        set_error("tblaeq_kmeans: D above 64 or k above 2^32 is not supported");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    const int idxBits = bits_for(n - 1), labBits = bits_for(k - 1);
    if (idxBits + labBits > 64) {
        set_error("tblaeq_kmeans: point and cluster index bits exceed 64");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    hipStream_t st = ctx->stream;
    int rc = TBLAEQ_OK;
    Scratch S;
    // mid-range centring (only the FP32 ranking uses it)
    double* mid = S.get<double>(64, &rc);
    double* minD = S.get<double>(n, &rc);
    const uint64_t ntiles = (n + kPpTile - 1) / kPpTile;
    double* tileSum = S.get<double>(ntiles, &rc);
    double* u = S.get<double>(k, &rc);
    float* x32 = S.get<float>(n * DP, &rc);
    float* xn = S.get<float>(n, &rc);
    uint64_t* ctl = S.get<uint64_t>(8, &rc);  // [0] changed [1] nAmb [2] cmax bits [3..4] scan scratch
    uint64_t* amb = S.get<uint64_t>(n, &rc);
    uint64_t* keysA = S.get<uint64_t>(n, &rc);
    uint64_t* keysB = S.get<uint64_t>(n, &rc);
    uint64_t* cnt = S.get<uint64_t>(k, &rc);
    uint64_t* off = S.get<uint64_t>(k + 1, &rc);
    uint64_t* hist = S.get<uint64_t>(radix_hist_words(n, 0), &rc);
    uint64_t* dn = S.get<uint64_t>(1, &rc);
    if (rc) return rc;
    // data bounds for the centring (per dimension min / max on the host side of a small copy)
    {
        std::vector<double> lo(D, __DBL_MAX__), hi(D, -__DBL_MAX__);
        const uint64_t chunk = 1 << 20;
        std::vector<double> buf(std::min<uint64_t>(n, chunk) * D);
        for (uint64_t b = 0; b < n; b += chunk) {
            const uint64_t m = std::min<uint64_t>(chunk, n - b);
            TB_HIP(hipMemcpyAsync(buf.data(), d_x + b * D, m * D * sizeof(double), hipMemcpyDeviceToHost, st));
            TB_HIP(hipStreamSynchronize(st));
            for (uint64_t i = 0; i < m; ++i)
                for (uint64_t d = 0; d < D; ++d) {
                    lo[d] = std::min(lo[d], buf[i * D + d]);
                    hi[d] = std::max(hi[d], buf[i * D + d]);
                }
        }
        std::vector<double> m64(64, 0.0);
        for (uint64_t d = 0; d < D; ++d) m64[d] = 0.5 * (lo[d] + hi[d]);
        TB_HIP(hipMemcpyAsync(mid, m64.data(), 64 * sizeof(double), hipMemcpyHostToDevice, st));
        TB_HIP(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(k_km_points, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_x, n, static_cast<int>(D), DP, mid,
                       x32, xn);
    const float g = 2.f * static_cast<float>(DP + 8) * 5.9604645e-8f;  // 2 (D + 8) 2^-24
    // Nearest of `m` centres (FP64, cent) for every point: the FP32 matrix-core
    // ranking with the FP64 re-score of near ties, labels into lab (the
    // assignment of the Lloyd iterations); c32 / cn hold at least m rows.
    auto assign = [&](const double* cent, uint64_t m, float* c32, float* cn, uint64_t* lab) {
        hipLaunchKernelGGL(k_km_zero, dim3(1), dim3(64), 0, st, ctl, 8);
        hipLaunchKernelGGL(k_km_centers, dim3(grid_for(ctx, m, 256)), dim3(256), 0, st, cent, m, static_cast<int>(D),
                           DP, mid, c32, cn, reinterpret_cast<uint32_t*>(ctl + 2));
        AssignArgs A{};
        A.n = n;
        A.k = m;
        A.x32 = x32;
        A.xn = xn;
        A.c32 = c32;
        A.cn = cn;
        A.cmax = reinterpret_cast<const uint32_t*>(ctl + 2);
        A.g = g;
        A.labels = lab;
        A.amb = amb;
        A.nAmb = ctl + 1;
        A.changed = ctl;
        launch_assign(DP, dim3(static_cast<unsigned>((n + kKmPoints - 1) / kKmPoints)), ctx, A);
        hipLaunchKernelGGL(k_km_rescore, dim3(static_cast<unsigned>(ctx->cuCount) * 4), dim3(256), 0, st, d_x,
                           static_cast<int>(D), cent, m, amb, ctl + 1, lab, ctl);
    };
    // random draws of the CPU build's k-means++ (CPUKmeans.cuh:67-118): the
    // first centre, then one uniform draw per further pick
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<size_t> uni(0, n - 1);
    const uint64_t first = uni(rng);
    std::vector<double> hu(k, 0.0);
    {
        std::uniform_real_distribution<double> unit(0.0, 1.0);
        for (uint64_t cc = 1; cc < k; ++cc) hu[cc] = unit(rng);
    }
    const uint64_t hn = n;
    TB_HIP(hipMemcpyAsync(u, hu.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
    TB_HIP(hipMemcpyAsync(dn, &hn, sizeof(uint64_t), hipMemcpyHostToDevice, st));
    TB_HIP(hipStreamSynchronize(st));
    // Seeding k-means++ over `m` points `pts` (optionally weighted) into
    // d_centers, starting from the point `start`.
    auto kmeanspp = [&](const double* pts, uint64_t m, const double* w, uint64_t start, double* mD, double* tS) {
        TB_HIP(hipMemcpyAsync(d_centers, pts + start * D, D * sizeof(double), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_km_fill_f64, dim3(grid_for(ctx, m, 256)), dim3(256), 0, st, mD, m, __DBL_MAX__);
        const uint64_t mt = (m + kPpTile - 1) / kPpTile;
        const unsigned ppGrid = static_cast<unsigned>(std::min<uint64_t>(mt, static_cast<uint64_t>(ctx->cuCount) * 8));
        const uint64_t report = std::max<uint64_t>(1, k / 8);
        for (uint64_t cc = 1; cc < k; ++cc) {
            hipLaunchKernelGGL(k_kmpp_update, dim3(ppGrid), dim3(256), 0, st, pts, m, static_cast<int>(D),
                               d_centers + (cc - 1) * D, mD, tS, w);
            hipLaunchKernelGGL(k_kmpp_pick, dim3(1), dim3(1024), 0, st, pts, m, static_cast<int>(D), mD, tS, u, cc,
                               d_centers, w);
            if (cc % report == 0 && k >= 4096) {
                TB_HIP(hipStreamSynchronize(st));
                progress("k-means seeding: %llu of %llu centres picked", static_cast<unsigned long long>(cc),
                         static_cast<unsigned long long>(k));
            }
        }
        TB_HIP(hipGetLastError());
        return TBLAEQ_OK;
    };
    uint64_t candidates = 0;
    if (opt->seeding == 1) {
        // k-means||: rounds of independent D^2-weighted draws (expected
        // oversampling * k per round) from a counter-based stream, each
        // round's draws in point order; the nearest candidate of every point
        // is tracked, the candidates weighted by the points nearest to them,
        // then weighted k-means++ on the candidates picks the k centres
        const double ell = opt->oversampling * static_cast<double>(k);
        uint64_t capC = 1 + opt->rounds * (static_cast<uint64_t>(2.0 * ell) + 1024);
        double* cand = S.get<double>(capC * D, &rc);
        uint64_t* nearest = S.get<uint64_t>(n, &rc);
        uint64_t* flag = S.get<uint64_t>(n, &rc);
        uint64_t* foff = S.get<uint64_t>(n + 1, &rc);
        uint64_t* lab = S.get<uint64_t>(n, &rc);
        double* phi = S.get<double>(1, &rc);
        uint64_t* scr = S.get<uint64_t>(2, &rc);
        if (rc) return rc;
        TB_HIP(hipMemcpyAsync(cand, d_x + first * D, D * sizeof(double), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_kmpar_init, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_x, n, static_cast<int>(D),
                           cand, minD, nearest);
        uint64_t m = 1;
        const unsigned tGrid = static_cast<unsigned>(std::min<uint64_t>(ntiles, static_cast<uint64_t>(ctx->cuCount) * 8));
        float* c32r = nullptr;
        float* cnr = nullptr;
        uint64_t capR = 0;
        for (uint64_t r = 0; r < opt->rounds; ++r) {
            hipLaunchKernelGGL(k_tile_sum, dim3(tGrid), dim3(256), 0, st, minD, n, tileSum);
            hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(1024), 0, st, tileSum, ntiles, phi);
            hipLaunchKernelGGL(k_kmpar_mark, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, minD, n, phi, ell, seed,
                               r, flag);
            if ((rc = scan_counts(ctx, flag, n, foff, scr))) return rc;
            uint64_t drawn = 0;
            TB_HIP(hipMemcpyAsync(&drawn, foff + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
            TB_HIP(hipStreamSynchronize(st));
            if (drawn == 0) break;
            if (m + drawn > capC) {
                // more draws than the expected bound: grow the candidate array
                const uint64_t nc = std::max(capC * 2, m + drawn);
                double* bigger = S.get<double>(nc * D, &rc);
                if (rc) return rc;
                TB_HIP(hipMemcpyAsync(bigger, cand, m * D * sizeof(double), hipMemcpyDeviceToDevice, st));
                cand = bigger;
                capC = nc;
            }
            hipLaunchKernelGGL(k_kmpar_gather, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_x, n,
                               static_cast<int>(D), flag, foff, cand + m * D);
            if (drawn > capR) {
                c32r = S.get<float>(drawn * DP, &rc);
                cnr = S.get<float>(drawn, &rc);
                if (rc) return rc;
                capR = drawn;
            }
            hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, lab, n);
            assign(cand + m * D, drawn, c32r, cnr, lab);
            hipLaunchKernelGGL(k_kmpar_improve, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_x, n,
                               static_cast<int>(D), cand + m * D, lab, m, minD, nearest);
            TB_HIP(hipGetLastError());
            m += drawn;
            TB_HIP(hipStreamSynchronize(st));
            progress("k-means|| round %llu of %llu: %llu candidates drawn, %llu in all",
                     static_cast<unsigned long long>(r + 1), static_cast<unsigned long long>(opt->rounds),
                     static_cast<unsigned long long>(drawn), static_cast<unsigned long long>(m));
        }
        if (m > k) {
            uint64_t* wcnt = S.get<uint64_t>(m, &rc);
            double* w = S.get<double>(m, &rc);
            double* cMinD = S.get<double>(m, &rc);
            double* cTile = S.get<double>((m + kPpTile - 1) / kPpTile, &rc);
            if (rc) return rc;
            hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, m, 256)), dim3(256), 0, st, wcnt, m);
            hipLaunchKernelGGL(k_kmpar_weights, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, nearest, n, wcnt);
            hipLaunchKernelGGL(k_u64_to_f64, dim3(grid_for(ctx, m, 256)), dim3(256), 0, st, wcnt, m, w);
            if ((rc = kmeanspp(cand, m, w, 0, cMinD, cTile))) return rc;
            candidates = m;
        }
    }
    if (candidates == 0 && (rc = kmeanspp(d_x, n, nullptr, first, minD, tileSum))) return rc;
    // Lloyd
    float* c32 = S.get<float>(k * DP, &rc);
    float* cn = S.get<float>(k, &rc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_labels, n);
    const uint64_t idxMask = (1ull << idxBits) - 1;
    uint64_t it = 0, rescored = 0;
    double assignMs = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TB_HIP(hipEventCreate(&e0));
    TB_HIP(hipEventCreate(&e1));
    for (; it < max_iter;) {
        TB_HIP(hipEventRecord(e0, st));
        assign(d_centers, k, c32, cn, d_labels);
        TB_HIP(hipEventRecord(e1, st));
        // update: sort (label, point) keys by label, sum each cluster in point order
        hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, k, 256)), dim3(256), 0, st, cnt, k);
        hipLaunchKernelGGL(k_km_keys, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_labels, n, idxBits, keysA, cnt);
        uint64_t* sorted = nullptr;
        if ((rc = radix_sort(ctx, keysA, keysB, dn, n, idxBits, labBits, hist, &sorted, RadixOut(), 8))) break;
        if ((rc = scan_counts(ctx, cnt, k, off, ctl + 3))) break;
        hipLaunchKernelGGL(k_km_update, dim3(grid_for(ctx, k * D, 256)), dim3(256), 0, st, d_x, static_cast<int>(D),
                           sorted, idxMask, off, k, d_centers);
        TB_HIP(hipGetLastError());
        uint64_t hc[2] = {0, 0};
        TB_HIP(hipMemcpyAsync(hc, ctl, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        TB_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        TB_HIP(hipEventElapsedTime(&ms, e0, e1));
        assignMs += ms;
        rescored += hc[1];
        ++it;
        if (k >= 4096 || n >= (1ull << 24))
            progress("k-means Lloyd iteration %llu of %llu: %llu labels changed, assignment %.1f ms",
                     static_cast<unsigned long long>(it), static_cast<unsigned long long>(max_iter),
                     static_cast<unsigned long long>(hc[0]), static_cast<double>(ms));
        if (hc[0] == 0) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (stats) {
        // inertia: per point squared distance to its centroid, summed per
        // 1024-point tile and then over the tiles in a fixed order
        hipLaunchKernelGGL(k_km_sqdist, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_x, n, static_cast<int>(D),
                           d_centers, d_labels, minD);
        const unsigned tGrid = static_cast<unsigned>(std::min<uint64_t>(ntiles, static_cast<uint64_t>(ctx->cuCount) * 8));
        hipLaunchKernelGGL(k_tile_sum, dim3(tGrid), dim3(256), 0, st, minD, n, tileSum);
        hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(1024), 0, st, tileSum, ntiles, u);
        double inertia = 0.0;
        TB_HIP(hipMemcpyAsync(&inertia, u, sizeof(double), hipMemcpyDeviceToHost, st));
        TB_HIP(hipStreamSynchronize(st));
        stats[0] = it;
        stats[1] = rescored;
        stats[2] = it * n * k;
        stats[3] = static_cast<uint64_t>(assignMs);
        std::memcpy(&stats[4], &inertia, sizeof(double));
        stats[5] = candidates;
    }
    return TBLAEQ_OK;
}

extern "C" int tblaeq_convert_f32_f64(tblaeq_ctx* ctx, const float* d_in, uint64_t n, double* d_out) {
    if (!ctx || (n && (!d_in || !d_out))) {
        set_error("tblaeq_convert_f32_f64: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    if (n) hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, d_in, n, d_out);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

extern "C" int tblaeq_build_level(tblaeq_ctx* ctx, const double* d_fine, uint64_t n, uint64_t D,
                                  const double* d_centers, uint64_t k, const uint64_t* d_labels, uint64_t* d_colRes,
                                  uint64_t* d_nnzPerCol, uint64_t* d_map, double* d_vals, double* d_radius) {
    if (!ctx || !d_fine || !d_centers || !d_labels || !d_colRes || !d_nnzPerCol || !d_map || !d_vals || !d_radius ||
        n == 0 || D == 0 || k == 0) {
        set_error("tblaeq_build_level: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const int idxBits = bits_for(n - 1), labBits = bits_for(k - 1);
    if (idxBits + labBits > 64) {
        set_error("tblaeq_build_level: point and cluster index bits exceed 64");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    hipStream_t st = ctx->stream;
    int rc = TBLAEQ_OK;
    Scratch S;
    uint64_t* keysA = S.get<uint64_t>(n, &rc);
    uint64_t* keysB = S.get<uint64_t>(n, &rc);
    uint64_t* hist = S.get<uint64_t>(radix_hist_words(n, 0), &rc);
    uint64_t* dn = S.get<uint64_t>(1, &rc);
    uint64_t* r2 = S.get<uint64_t>(k, &rc);
    uint64_t* scratch = S.get<uint64_t>(2, &rc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, k, 256)), dim3(256), 0, st, d_nnzPerCol, k);
    hipLaunchKernelGGL(k_km_zero, dim3(grid_for(ctx, k, 256)), dim3(256), 0, st, r2, k);
    hipLaunchKernelGGL(k_km_keys, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_labels, n, idxBits, keysA,
                       d_nnzPerCol);
    const uint64_t hn = n;
    TB_HIP(hipMemcpyAsync(dn, &hn, sizeof(uint64_t), hipMemcpyHostToDevice, st));
    uint64_t* sorted = nullptr;
    if ((rc = radix_sort(ctx, keysA, keysB, dn, n, idxBits, labBits, hist, &sorted, RadixOut(), 8))) return rc;
    if ((rc = scan_counts(ctx, d_nnzPerCol, k, d_colRes, scratch))) return rc;
    hipLaunchKernelGGL(k_csc_rows, dim3(grid_for(ctx, n, 256)), dim3(256), 0, st, d_fine, n, static_cast<int>(D),
                       d_centers, sorted, idxBits, d_map, d_vals, r2);
    hipLaunchKernelGGL(k_csc_radius, dim3(grid_for(ctx, k, 256)), dim3(256), 0, st, r2, k, d_radius);
    TB_HIP(hipGetLastError());
    TB_HIP(hipStreamSynchronize(st));
    return TBLAEQ_OK;
}

extern "C" int tblaeq_kmeans_host(tblaeq_ctx* ctx, const double* x, uint64_t n, uint64_t D, uint64_t k,
                                  uint64_t max_iter, uint64_t seed, double* centers, uint64_t* labels,
                                  uint64_t* stats) {
    if (!ctx || !x || !centers || !labels || n == 0 || D == 0) {
        set_error("tblaeq_kmeans_host: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    int rc = TBLAEQ_OK;
    Scratch S;
    double* dx = S.get<double>(n * D, &rc);
    double* dc = S.get<double>(k * D, &rc);
    uint64_t* dl = S.get<uint64_t>(n, &rc);
    if (rc) return rc;
    TB_HIP(hipMemcpyAsync(dx, x, n * D * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = tblaeq_kmeans(ctx, dx, n, D, k, max_iter, seed, dc, dl, stats))) return rc;
    TB_HIP(hipMemcpyAsync(centers, dc, k * D * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TB_HIP(hipMemcpyAsync(labels, dl, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    TB_HIP(hipStreamSynchronize(ctx->stream));
    return TBLAEQ_OK;
}
