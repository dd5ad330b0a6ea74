// Device-side helpers shared by the T-BLAEQ gfx950 kernels.
//
// Arithmetic order is fixed so that the CPU oracle (built with contraction off
// and explicit fma) reproduces every FP64 result bit-for-bit:
//   box distance   dist = fma(diff, diff, dist) over d = 0..D-1     (RangePruning.cu:8-21)
//   KNN bound      lb = sqrt_rd(fma-chain of (q-c)^2) - r           (KNN.cu:14-21)
//   prolongation   y = P * x                                        (SpMSpV.cuh:64)
//   point distance fma(diff, diff, acc) with diff = q - v
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tblaeq {

constexpr int kWave = 64;

// Round-toward-minus-infinity square root built from the correctly rounded
// sqrt: the RN result s is within half an ulp, so RD is s when s*s <= x and the
// next double below s otherwise; fma gives the exact sign of s*s - x.
__device__ __forceinline__ double sqrt_rd(double x) {
    double s = __builtin_sqrt(x);
    if (!(x > 0.0) || !(x < __builtin_inf())) return s;
    // The result must be the largest double whose square does not exceed x; fma
    // gives the exact sign of s*s - x.  Walk at most two ulps either way so the
    // answer does not depend on the library sqrt being correctly rounded.
    // s is positive and finite here, so neighbours are the bit pattern -/+ 1.
    if (__builtin_fma(s, s, -x) > 0.0) {
        s = __longlong_as_double(__double_as_longlong(s) - 1);
        if (__builtin_fma(s, s, -x) > 0.0) s = __longlong_as_double(__double_as_longlong(s) - 1);
    } else {
        const double n = __longlong_as_double(__double_as_longlong(s) + 1);
        if (__builtin_fma(n, n, -x) <= 0.0) s = n;
    }
    return s;
}

// Order-preserving map from double to uint64 (total order on non-NaN values,
// -0.0 folded onto +0.0 so that it compares equal like the double does).
__device__ __forceinline__ uint64_t okey(double x) {
    uint64_t b = __double_as_longlong(x + 0.0);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_inv(uint64_t k) {
    uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double(b);
}

__device__ __forceinline__ double box_dist2(const double* __restrict__ c, const double* __restrict__ lo,
                                            const double* __restrict__ hi, int D) {
    double dist = 0.0;
    for (int d = 0; d < D; ++d) {
        const double cv = c[d];
        double diff = 0.0;
        if (cv < lo[d]) diff = lo[d] - cv;
        else if (cv > hi[d]) diff = cv - hi[d];
        dist = __builtin_fma(diff, diff, dist);
    }
    return dist;
}

__device__ __forceinline__ double point_dist2(const double* __restrict__ v, const double* __restrict__ q, int D) {
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
        const double diff = q[d] - v[d];
        acc = __builtin_fma(diff, diff, acc);
    }
    return acc;
}

__device__ __forceinline__ bool in_box(const double* __restrict__ v, const double* __restrict__ lo,
                                       const double* __restrict__ hi, int D) {
    bool ok = true;
    for (int d = 0; d < D; ++d) ok = ok && (v[d] >= lo[d]) && (v[d] <= hi[d]);
    return ok;
}

__device__ __forceinline__ double knn_lb(const double* __restrict__ c, const double* __restrict__ q, double r,
                                         int D) {
    double dist = 0.0;
    for (int d = 0; d < D; ++d) {
        const double diff = q[d] - c[d];
        dist = __builtin_fma(diff, diff, dist);
    }
    return sqrt_rd(dist) - r;
}

__device__ __forceinline__ uint64_t ld_acquire(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Orders LDS writes of one wave before its later LDS reads by other lanes
// (a wave's LDS operations execute in order; this keeps the compiler from
// moving accesses across the point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Read-only device arrays addressed through the constant address space: a
// wave-uniform load from such a pointer is a scalar load into SGPRs, so a child
// row is fetched once per wave (s_load of the whole row) and enters the vector
// compares as a scalar operand.  Only valid for arrays no kernel writes while
// the reading kernel runs.
typedef const __attribute__((address_space(4))) double* ConstF64;
typedef const __attribute__((address_space(4))) uint64_t* ConstU64;
typedef const __attribute__((address_space(4))) float* ConstF32;
// LDS pointer (its integer value is the LDS byte address).
typedef __attribute__((address_space(3))) uint64_t* LdsU64;

// Uniform copy of a value every lane holds (keeps it in SGPRs).
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Lane k's value (k wave-uniform) as a scalar.
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int k) {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), k);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), k);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Next chunk [*begin, *end) of the nItems work items for this wave, taken from
// the queue counter in chunks of `chunk`; false when the queue is dry.
__device__ __forceinline__ bool next_items(uint64_t* queue, uint64_t chunk, uint64_t nItems, int lane,
                                           uint64_t* begin, uint64_t* end) {
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(queue), chunk);
    base = uniform_u64(__shfl(base, 0, kWave));
    if (base >= nItems) return false;
    *begin = base;
    *end = base + chunk < nItems ? base + chunk : nItems;
    return true;
}

// Wave-uniform child row r of the cluster at `row` (D == DM: one wide scalar
// load; padding dimensions read as 0).
template <int DM, bool EXACT>
__device__ __forceinline__ void scalar_row(double* x, ConstF64 row, uint32_t r, int D) {
#pragma unroll
    for (int d = 0; d < DM; ++d) x[d] = (EXACT || d < D) ? row[r * (EXACT ? DM : D) + d] : 0.0;
}

// Items per queue grab: about 16 grabs per wave.
__device__ __forceinline__ uint64_t queue_chunk(uint64_t nItems, uint64_t waves) {
    const uint64_t c = nItems / (waves * 16);
    return c ? c : 1;
}

// Per-launch control block for the single-pass scans (decoupled look-back).
// Status words carry their own epoch so the tile-status arrays never need to be
// cleared between launches: word = epoch(27b) | valid(1b) | value(36b).
struct ScanCtrl {
    uint64_t tileCounter;   // dynamic tile ids; reset by the last workgroup to leave
    uint64_t exitCounter;
    uint64_t outCount;      // survivors written by the last launch
    uint64_t outNnz;        // sum of the survivors' weights (next level's size)
    double ub;              // KNN STEP bound
    uint64_t aux;           // spare
    uint64_t seq;           // epoch of the next scan launch (advanced on the device)
    uint64_t pad;
};

constexpr int kValBits = 36;
constexpr uint64_t kValMask = (1ull << kValBits) - 1;

__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint64_t v) {
    return (static_cast<uint64_t>(epoch & 0x7FFFFFF) << 37) | (1ull << 36) | (v & kValMask);
}
__device__ __forceinline__ bool status_valid(uint64_t w, uint32_t epoch) {
    return (w >> 36) == ((static_cast<uint64_t>(epoch & 0x7FFFFFF) << 1) | 1ull);
}

}  // namespace tblaeq
