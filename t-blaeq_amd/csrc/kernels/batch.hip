// Batched range queries: thousands of boxes per launch, cluster-major.
//
// Per level, surviving (cluster, query) pairs are grouped by cluster with a
// counting sort, so that one 64-wide wave processes a slice of 64 queries that
// all selected the same cluster: each lane holds one query box in registers and
// the wave walks the cluster's children, whose reconstructed coordinates,
// radius and id are wave-uniform loads.  A child is read once per slice instead
// of once per query.  Every work item is a (cluster, slice) pair taken from a
// device-resident list; counts stay on the device, so a batch runs without host
// synchronisation until its hits are sorted into per-query id sets.
//
// Pruning at level l uses the same test and operation order as the per-query
// path (RangePruning.cu:5-25: fma chain of the clamped differences against
// r*r); the last step is the exact point-in-box filter of the candidates.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../batch_internal.hpp"
#include "../scan_common.hpp"

namespace tblaeq {

// ------------------------------------------------------------------ creation

__global__ void k_gather_off(const uint64_t* __restrict__ ids, uint64_t n, const uint64_t* __restrict__ colRes,
                             uint64_t* __restrict__ off) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) off[i] = colRes[ids[i]];
}

__global__ void k_gather_f64(const uint64_t* __restrict__ ids, uint64_t n, const double* __restrict__ src,
                             double* __restrict__ dst) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = src[ids[i]];
}

__global__ void k_gather_u64(const uint64_t* __restrict__ ids, uint64_t n, const uint64_t* __restrict__ src,
                             uint64_t* __restrict__ dst) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = src[ids[i]];
}

// parent[j] = c for every CSC row j of column c.
__global__ void k_fill_parent(const uint64_t* __restrict__ colRes, uint64_t cols, uint64_t* __restrict__ parent) {
    const uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (c < cols)
        for (uint64_t j = colRes[c]; j < colRes[c + 1]; ++j) parent[j] = c;
}

__global__ void k_set_u64(uint64_t* p, uint64_t v, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = v;
}

// ------------------------------------------------------------- range steps
//
// Work item = (cluster c of the previous level, slice of <= 64 queries that kept
// c), or at level 0 a slice of queries over all coarsest points.  Lane = query,
// box in registers.  The cluster's children are read one row at a time with
// wave-uniform scalar loads (coordinates, radius, bounding box into SGPRs; the
// next row's load is issued before the current row's arithmetic) and tested
// branch-free against every lane.  Padding dimensions (D < DM) hold 0 in the
// child and in the box, which passes the box test and adds fma(0,0,d) = d to
// the distance.
//
// Survivors of a prune are grouped by cluster without materialising pairs: the
// count pass adds each child's survivor mask popcount to cnt[id]; after the
// group scan the write pass recomputes the masks and writes the queries
// straight into the next level's per-cluster query lists (cursor = cnt counted
// back down to zero).  The last prune additionally drops (cluster, query) pairs
// whose box misses the bounding box of the cluster's points: such a cluster
// holds no hit.  Survivor statistics still count every ball survivor.

enum RangeMode { kCount = 0, kWrite = 1, kFinal = 2 };

struct RangeStepArgs {
    int D;
    int level0;        // candidates are all level-0 positions, items are query slices
    uint64_t nq;
    const double* boxes;  // nq x 2D (lo | hi)
    // groups of the previous level
    const uint64_t* items;
    const uint64_t* nItems;
    uint64_t itemCap;
    const uint64_t* off;
    const uint32_t* qlist;
    uint64_t qcap;        // entries of qlist
    const uint64_t* colResPrev;
    // candidates
    uint64_t count0;
    const double* coords;
    const double* radiusPos;
    const uint64_t* weightPos;   // child counts (last prune: candidate superset size)
    const float* aabb;           // last prune: per position lo[D] | hi[D], rounded outward
    const uint64_t* ids;
    // next level's groups (count / write)
    uint64_t* cnt;
    const uint64_t* offNext;
    uint32_t* qlistNext;
    // hits (final)
    uint64_t* out;
    uint64_t cap;
    uint64_t* outCount;
    uint64_t* overflow;
    int idBits;
    // optional per-query statistics (count pass)
    uint64_t* survStat;  // [nq * L]
    uint64_t* fineStat;  // [nq]
    int level, L;
    uint64_t* work;      // final: [0] row x box tests, [1] child rows, [2] leaf x box tests
    uint64_t* queue;     // final: work-queue counter (zero at launch)
    // prune: the count pass stores each item's per-row keep masks here when
    // they fit in maskCap (item i's rows from maskOff[i], each wave reserving
    // the rows of all its items on the maskNext counter; level 0: from
    // i * count0), and the write pass reads them instead of recomputing
    uint64_t* masks;
    uint64_t* maskOff;
    uint64_t* maskNext;
    uint64_t maskCap;
    // beside each stored mask row: the row's first slot in its cluster's query
    // list (returned by the count pass's atomic), so that the write pass, when
    // every row of the level was stored, places the queries without atomics
    uint64_t* maskPos;
    // final filter: first leaf of every last-level cluster (by id) and the
    // leaves' outward-rounded float boxes lo[D] | hi[D]
    const uint64_t* leafOff;
    const float* leafBox;
    // final filter: outward-rounded float copy of every box, lo | hi with
    // stride 2 * dm_of(D) (padding dimensions 0)
    const float* fboxes;
    // with a box test (aabb): evaluate the ball test of every row, also where
    // the box test already failed for every lane (the statistics count ball
    // survivors); the kept pairs are ball AND box either way
    int ballStats;
    // final filter: clusters kept by at most rowMax queries are listed in
    // sparse (*nSparse of them) for the row-major kernel (k_range_rows, items
    // from queue + 1) and get no query-major work items
    uint64_t rowMax;
    const uint64_t* sparse;     // [sparseTiles * kScanThreads] (tile t's clusters from t * kScanThreads)
    const uint64_t* sparseCnt;  // [sparseTiles] clusters in each tile's slots
    const uint64_t* sparseUnits;      // non-empty units of kRowUnit slots
    const uint64_t* sparseUnitCount;  // their number
    uint64_t sparseTiles;
};

// Appends a wave's staged hits to the global hit list (entries past the
// capacity are dropped and flagged).
__device__ __forceinline__ void flush_hits(const RangeStepArgs& A, const uint64_t* stage, uint64_t n, int lane) {
    if (n == 0) return;
    wave_lds_sync();
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(A.outCount), n);
    base = __shfl(base, 0, kWave);
    for (uint64_t o = lane; o < n; o += kWave)
        if (base + o < A.cap) A.out[base + o] = stage[o];
    if (lane == 0 && base + n > A.cap) atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 4ull);
    wave_lds_sync();
}


// flush_hits for code that keeps few scalar registers free: the output
// pointers and capacity are read from the kernel arguments at the flush (a
// volatile read, so they are not held in registers across the caller's loop).
__device__ __forceinline__ void flush_hits_lean(const RangeStepArgs& A, const uint64_t* stage, uint64_t n, int lane) {
    if (n == 0) return;
    uint64_t* const out = *reinterpret_cast<uint64_t* const volatile*>(&A.out);
    uint64_t* const outCount = *reinterpret_cast<uint64_t* const volatile*>(&A.outCount);
    uint64_t* const overflow = *reinterpret_cast<uint64_t* const volatile*>(&A.overflow);
    const uint64_t cap = *reinterpret_cast<const volatile uint64_t*>(&A.cap);
    wave_lds_sync();
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(outCount), n);
    base = __shfl(base, 0, kWave);
    for (uint64_t o = lane; o < n; o += kWave)
        if (base + o < cap) out[base + o] = stage[o];
    if (lane == 0 && base + n > cap) atomicOr(reinterpret_cast<unsigned long long*>(overflow), 4ull);
    wave_lds_sync();
}

// Write pass of a chunk of nc child rows: lane k holds row k's keep mask
// (queries of the slice that keep child myId); each kept (child, query) pair
// takes the next slot of the child's query list (cnt counted back down).
__device__ __forceinline__ void write_kept(const RangeStepArgs& A, uint64_t myId, uint64_t myMask, int nc,
                                           uint64_t q, int lane, uint64_t ltmask) {
    uint64_t myBase = 0, myOff = 0;
    if (lane < nc && myMask) {
        const uint64_t pc = __popcll(myMask);
        myBase = atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[myId]), 0ull - pc) - pc;
        myOff = A.offNext[myId];
    }
    for (int k = 0; k < nc; ++k) {
        const uint64_t m = readlane_u64(myMask, k);
        if (m == 0) continue;
        const uint64_t base = readlane_u64(myOff, k) + readlane_u64(myBase, k);
        if ((m >> lane) & 1ull) {
            const uint64_t o = base + __popcll(m & ltmask);
            if (o < A.cap) A.qlistNext[o] = static_cast<uint32_t>(q);
            else atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
        }
    }
}

// Lanes of m whose outward-rounded float box [flo, fhi] meets the
// wave-uniform bounding box [blo, bhi] in eight dimensions: the same v_cmpx
// chain as box_mask8 (EXEC narrowed from EXEC & m, read back and restored in
// the one asm block), on floats.  blo <= fhi and bhi >= flo per dimension,
// false for NaN as the C comparisons.
__device__ __forceinline__ uint64_t aabb_mask8(uint64_t m, const float* blo, const float* bhi, const float* flo,
                                               const float* fhi) {
    uint64_t out, save;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_and_b64 exec, exec, %[m]\n\t"
        "v_cmpx_le_f32 vcc, %[a0], %[f0]\n\t"
        "v_cmpx_ge_f32 vcc, %[b0], %[g0]\n\t"
        "v_cmpx_le_f32 vcc, %[a1], %[f1]\n\t"
        "v_cmpx_ge_f32 vcc, %[b1], %[g1]\n\t"
        "v_cmpx_le_f32 vcc, %[a2], %[f2]\n\t"
        "v_cmpx_ge_f32 vcc, %[b2], %[g2]\n\t"
        "v_cmpx_le_f32 vcc, %[a3], %[f3]\n\t"
        "v_cmpx_ge_f32 vcc, %[b3], %[g3]\n\t"
        "v_cmpx_le_f32 vcc, %[a4], %[f4]\n\t"
        "v_cmpx_ge_f32 vcc, %[b4], %[g4]\n\t"
        "v_cmpx_le_f32 vcc, %[a5], %[f5]\n\t"
        "v_cmpx_ge_f32 vcc, %[b5], %[g5]\n\t"
        "v_cmpx_le_f32 vcc, %[a6], %[f6]\n\t"
        "v_cmpx_ge_f32 vcc, %[b6], %[g6]\n\t"
        "v_cmpx_le_f32 vcc, %[a7], %[f7]\n\t"
        "v_cmpx_ge_f32 vcc, %[b7], %[g7]\n\t"
        "s_mov_b64 %[out], exec\n\t"
        "s_mov_b64 exec, %[sv]"
        : [out] "=&s"(out), [sv] "=&s"(save)
        : [m] "s"(m), [a0] "s"(blo[0]), [a1] "s"(blo[1]), [a2] "s"(blo[2]), [a3] "s"(blo[3]), [a4] "s"(blo[4]),
          [a5] "s"(blo[5]), [a6] "s"(blo[6]), [a7] "s"(blo[7]), [b0] "s"(bhi[0]), [b1] "s"(bhi[1]), [b2] "s"(bhi[2]),
          [b3] "s"(bhi[3]), [b4] "s"(bhi[4]), [b5] "s"(bhi[5]), [b6] "s"(bhi[6]), [b7] "s"(bhi[7]), [f0] "v"(fhi[0]),
          [f1] "v"(fhi[1]), [f2] "v"(fhi[2]), [f3] "v"(fhi[3]), [f4] "v"(fhi[4]), [f5] "v"(fhi[5]), [f6] "v"(fhi[6]),
          [f7] "v"(fhi[7]), [g0] "v"(flo[0]), [g1] "v"(flo[1]), [g2] "v"(flo[2]), [g3] "v"(flo[3]), [g4] "v"(flo[4]),
          [g5] "v"(flo[5]), [g6] "v"(flo[6]), [g7] "v"(flo[7])
        : "vcc");
    return out;
}
template <int DM>
__device__ __forceinline__ uint64_t aabb_mask(uint64_t m, const float* blo, const float* bhi, const float* flo,
                                              const float* fhi) {
    if constexpr (DM == 4) {
        float a[8] = {blo[0], blo[1], blo[2], blo[3], 0.0f, 0.0f, 0.0f, 0.0f};
        float b[8] = {bhi[0], bhi[1], bhi[2], bhi[3], 0.0f, 0.0f, 0.0f, 0.0f};
        float f[8] = {fhi[0], fhi[1], fhi[2], fhi[3], 0.0f, 0.0f, 0.0f, 0.0f};
        float g[8] = {flo[0], flo[1], flo[2], flo[3], 0.0f, 0.0f, 0.0f, 0.0f};
        return aabb_mask8(m, a, b, g, f);
    } else {
#pragma unroll
        for (int d = 0; d < DM; d += 8) m = aabb_mask8(m, blo + d, bhi + d, flo + d, fhi + d);
        return m;
    }
}

template <int DM, int MODE, bool EXACT>
__global__ __launch_bounds__(kStepWaves * kWave) void k_range_step(RangeStepArgs A) {
    static_assert(MODE == kCount || MODE == kWrite, "prune modes only");
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t gw = blockIdx.x * static_cast<uint64_t>(kStepWaves) + wave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = EXACT ? DM : A.D;
    const ConstF64 cx = (ConstF64)A.coords;
    const ConstF64 crad = (ConstF64)A.radiusPos;
    const ConstF32 cab = (ConstF32)A.aabb;
    uint64_t nItems = A.level0 ? (A.nq + kWave - 1) / kWave : ((ConstU64)A.nItems)[0];
    if (!A.level0 && nItems > A.itemCap) nItems = A.itemCap;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    // Every level keeps a (cluster, query) pair only if the reference's ball
    // test passes.  The last prune also needs the box to meet the bounding box
    // of the cluster's points (no hit otherwise): that compare chain runs
    // first, and a row whose box test fails for every lane skips the FP64 ball
    // test unless statistics (which count ball survivors) are being gathered.
    const bool useBox = A.aabb != nullptr;
    const bool ballAlways = !useBox || A.ballStats;
    // write pass: every item's mask rows were stored (level 0: item i's rows at
    // i * count0; else the count pass reserved within the capacity)
    const bool allStored = MODE == kWrite && A.masks &&
                           (A.level0 ? ((A.nq + kWave - 1) / kWave) * A.count0 <= A.maskCap
                                     : ((ConstU64)A.maskNext)[0] <= A.maskCap);
    // count pass: one reservation of keep-mask rows for all of this wave's items
    uint64_t maskBase = 0;
    if (MODE == kCount && A.masks && !A.level0) {
        uint64_t rows = 0;
        for (uint64_t item = gw; item < nItems; item += nw) {
            const uint64_t c = ((ConstU64)A.items)[item] >> kSliceBits;
            rows += ((ConstU64)A.colResPrev)[c + 1] - ((ConstU64)A.colResPrev)[c];
        }
        if (lane == 0 && rows) maskBase = atomicAdd(reinterpret_cast<unsigned long long*>(A.maskNext), rows);
        maskBase = uniform_u64(__shfl(maskBase, 0, kWave));
    }
    for (uint64_t item = gw; item < nItems; item += nw) {
        uint64_t jb, je, q = 0;
        bool valid;
        if (A.level0) {
            const uint64_t qq = item * kWave + lane;
            valid = qq < A.nq;
            q = valid ? qq : 0;
            jb = 0;
            je = A.count0;
        } else {
            const uint64_t it = ((ConstU64)A.items)[item];
            const uint64_t c = it >> kSliceBits;
            const uint64_t start = (it & ((1ull << kSliceBits) - 1)) * kWave;
            const uint64_t g0 = ((ConstU64)A.off)[c];
            const uint64_t gn = ((ConstU64)A.off)[c + 1] - g0;
            valid = start + lane < gn && g0 + start + lane < A.qcap;
            q = valid ? A.qlist[g0 + start + lane] : 0;
            jb = ((ConstU64)A.colResPrev)[c];
            je = ((ConstU64)A.colResPrev)[c + 1];
        }
        uint64_t mo = 0;
        bool stored = false;
        if (A.masks) {
            if (A.level0) {
                mo = item * A.count0;
            } else if (MODE == kCount) {
                mo = maskBase;
                maskBase += je - jb;
                if (lane == 0) A.maskOff[item] = mo;
            } else {
                mo = A.maskOff[item];
            }
            stored = mo + (je - jb) <= A.maskCap;
        }
        if (MODE == kWrite && stored) {
            // masks of the count pass: only the writes remain
            for (uint64_t j0 = jb; j0 < je; j0 += kWave) {
                const int nc = static_cast<int>(je - j0 < static_cast<uint64_t>(kWave) ? je - j0 : kWave);
                const uint64_t myMask = lane < nc ? A.masks[mo + (j0 - jb) + lane] : 0;
                if (allStored) {
                    const bool any = myMask != 0;
                    const uint64_t myBase = any ? A.offNext[A.ids[j0 + lane]] + A.maskPos[mo + (j0 - jb) + lane] : 0;
                    for (uint64_t nz = __ballot(any); nz; nz &= nz - 1) {
                        const int k = __builtin_ctzll(nz);
                        const uint64_t m = readlane_u64(myMask, k);
                        if ((m >> lane) & 1ull) {
                            const uint64_t o = readlane_u64(myBase, k) + __popcll(m & ltmask);
                            if (o < A.cap) A.qlistNext[o] = static_cast<uint32_t>(q);
                            else atomicOr(reinterpret_cast<unsigned long long*>(A.overflow), 1ull);
                        }
                    }
                } else {
                    const uint64_t myId = lane < nc ? A.ids[j0 + lane] : 0;
                    write_kept(A, myId, myMask, nc, q, lane, ltmask);
                }
            }
            continue;
        }
        double lo[DM], hi[DM];
        float flo[DM], fhi[DM];
#pragma unroll
        for (int d = 0; d < DM; ++d) {
            lo[d] = d < D ? A.boxes[q * 2 * D + d] : 0.0;
            hi[d] = d < D ? A.boxes[q * 2 * D + D + d] : 0.0;
            flo[d] = __double2float_rd(lo[d]);
            fhi[d] = __double2float_ru(hi[d]);
        }
        // with no inverted box in the wave, the distance term max(lo - x, x - hi, 0)
        // equals the conditional form (NaN bounds and coordinates included)
        bool inverted = false;
#pragma unroll
        for (int d = 0; d < DM; ++d) inverted |= lo[d] > hi[d];
        const bool fastDist = __ballot(valid && inverted) == 0;
        uint64_t mine = 0, fine = 0;
        for (uint64_t j0 = jb; j0 < je; j0 += kWave) {
            const int nc = static_cast<int>(je - j0 < static_cast<uint64_t>(kWave) ? je - j0 : kWave);
            const uint64_t myId = lane < nc ? A.ids[j0 + lane] : 0;
            uint64_t myMask = 0, myBall = 0;
            const uint64_t vm = __ballot(valid);
            double xn[DM], rn;
#pragma unroll
            for (int d = 0; d < DM; ++d) xn[d] = d < D ? cx[j0 * D + d] : 0.0;
            rn = crad[j0];
            for (int k = 0; k < nc; ++k) {
                const uint64_t j = j0 + k;
                double xv[DM];
#pragma unroll
                for (int d = 0; d < DM; ++d) xv[d] = xn[d];
                const double r = rn;
                if (k + 1 < nc) {
#pragma unroll
                    for (int d = 0; d < DM; ++d) xn[d] = d < D ? cx[(j + 1) * D + d] : 0.0;
                    rn = crad[j + 1];
                }
                uint64_t boxm = vm;
                if (useBox) {
                    // compare chain against the row's bounding box in SGPRs
                    float blo[DM], bhi[DM];
#pragma unroll
                    for (int d = 0; d < DM; ++d) {
                        blo[d] = d < D ? cab[j * 2 * D + d] : 0.0f;
                        bhi[d] = d < D ? cab[j * 2 * D + D + d] : 0.0f;
                    }
                    boxm = aabb_mask<DM>(vm, blo, bhi, flo, fhi);
                }
                uint64_t ballm = 0;
                if (ballAlways || boxm != 0) {
                    double dist = 0.0;
                    if (fastDist) {
#pragma unroll
                        for (int d = 0; d < DM; ++d) {
                            const double diff = fmax(fmax(lo[d] - xv[d], xv[d] - hi[d]), 0.0);
                            dist = __builtin_fma(diff, diff, dist);
                        }
                    } else {
#pragma unroll
                        for (int d = 0; d < DM; ++d) {
                            const double below = lo[d] - xv[d], above = xv[d] - hi[d];
                            const double diff = xv[d] < lo[d] ? below : (xv[d] > hi[d] ? above : 0.0);
                            dist = __builtin_fma(diff, diff, dist);
                        }
                    }
                    ballm = __ballot(valid && dist <= r * r);
                }
                if (lane == k) {
                    myMask = ballm & boxm;
                    myBall = ballm;
                }
            }
            if (MODE == kCount) {
                // ball survivors per lane for the statistics; kept pairs per cluster
                if (A.survStat || A.fineStat) {
                    const uint64_t myW = (A.weightPos && lane < nc) ? A.weightPos[j0 + lane] : 0;
                    for (int k = 0; k < nc; ++k) {
                        const uint64_t mb = __shfl(myBall, k, kWave);
                        const uint64_t w = __shfl(myW, k, kWave);
                        if ((mb >> lane) & 1ull) {
                            ++mine;
                            fine += w;
                        }
                    }
                }
                uint64_t slot = 0;
                if (lane < nc && myMask)
                    slot = atomicAdd(reinterpret_cast<unsigned long long*>(&A.cnt[myId]),
                                     static_cast<unsigned long long>(__popcll(myMask)));
                if (stored && lane < nc) {
                    A.masks[mo + (j0 - jb) + lane] = myMask;
                    A.maskPos[mo + (j0 - jb) + lane] = slot;
                }
                continue;
            }
            write_kept(A, myId, myMask, nc, q, lane, ltmask);
        }
        if (MODE == kCount && valid) {
            if (A.survStat && mine)
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.survStat[q * A.L + A.level]), mine);
            if (A.fineStat && A.level == A.L - 1 && fine)
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.fineStat[q]), fine);
        }
    }
}

// Lanes of m whose box [lo, hi] contains the wave-uniform point x in four
// dimensions: a chain of v_cmpx compares narrows EXEC (starting from EXEC & m),
// the survivors are read back and EXEC is restored inside the one asm block, so
// no compiler-visible code runs with the narrowed EXEC.  Same predicate as
// x >= lo && x <= hi per dimension (a NaN on either side compares false); the
// mask stays in SGPRs with no per-compare scalar AND.
__device__ __forceinline__ uint64_t box_mask4(uint64_t m, const double* x, const double* lo, const double* hi) {
    uint64_t out, save;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_and_b64 exec, exec, %[m]\n\t"
        "v_cmpx_ge_f64 vcc, %[x0], %[l0]\n\t"
        "v_cmpx_le_f64 vcc, %[x0], %[h0]\n\t"
        "v_cmpx_ge_f64 vcc, %[x1], %[l1]\n\t"
        "v_cmpx_le_f64 vcc, %[x1], %[h1]\n\t"
        "v_cmpx_ge_f64 vcc, %[x2], %[l2]\n\t"
        "v_cmpx_le_f64 vcc, %[x2], %[h2]\n\t"
        "v_cmpx_ge_f64 vcc, %[x3], %[l3]\n\t"
        "v_cmpx_le_f64 vcc, %[x3], %[h3]\n\t"
        "s_mov_b64 %[out], exec\n\t"
        "s_mov_b64 exec, %[sv]"
        : [out] "=&s"(out), [sv] "=&s"(save)
        : [m] "s"(m), [x0] "s"(x[0]), [x1] "s"(x[1]), [x2] "s"(x[2]), [x3] "s"(x[3]), [l0] "v"(lo[0]),
          [l1] "v"(lo[1]), [l2] "v"(lo[2]), [l3] "v"(lo[3]), [h0] "v"(hi[0]), [h1] "v"(hi[1]), [h2] "v"(hi[2]),
          [h3] "v"(hi[3])
        : "vcc");
    return out;
}

__device__ __forceinline__ uint64_t box_mask8(uint64_t m, const double* x, const double* lo, const double* hi) {
    uint64_t out, save;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_and_b64 exec, exec, %[m]\n\t"
        "v_cmpx_ge_f64 vcc, %[x0], %[l0]\n\t"
        "v_cmpx_le_f64 vcc, %[x0], %[h0]\n\t"
        "v_cmpx_ge_f64 vcc, %[x1], %[l1]\n\t"
        "v_cmpx_le_f64 vcc, %[x1], %[h1]\n\t"
        "v_cmpx_ge_f64 vcc, %[x2], %[l2]\n\t"
        "v_cmpx_le_f64 vcc, %[x2], %[h2]\n\t"
        "v_cmpx_ge_f64 vcc, %[x3], %[l3]\n\t"
        "v_cmpx_le_f64 vcc, %[x3], %[h3]\n\t"
        "v_cmpx_ge_f64 vcc, %[x4], %[l4]\n\t"
        "v_cmpx_le_f64 vcc, %[x4], %[h4]\n\t"
        "v_cmpx_ge_f64 vcc, %[x5], %[l5]\n\t"
        "v_cmpx_le_f64 vcc, %[x5], %[h5]\n\t"
        "v_cmpx_ge_f64 vcc, %[x6], %[l6]\n\t"
        "v_cmpx_le_f64 vcc, %[x6], %[h6]\n\t"
        "v_cmpx_ge_f64 vcc, %[x7], %[l7]\n\t"
        "v_cmpx_le_f64 vcc, %[x7], %[h7]\n\t"
        "s_mov_b64 %[out], exec\n\t"
        "s_mov_b64 exec, %[sv]"
        : [out] "=&s"(out), [sv] "=&s"(save)
        : [m] "s"(m), [x0] "s"(x[0]), [x1] "s"(x[1]), [x2] "s"(x[2]), [x3] "s"(x[3]), [x4] "s"(x[4]),
          [x5] "s"(x[5]), [x6] "s"(x[6]), [x7] "s"(x[7]), [l0] "v"(lo[0]), [l1] "v"(lo[1]), [l2] "v"(lo[2]),
          [l3] "v"(lo[3]), [l4] "v"(lo[4]), [l5] "v"(lo[5]), [l6] "v"(lo[6]), [l7] "v"(lo[7]), [h0] "v"(hi[0]),
          [h1] "v"(hi[1]), [h2] "v"(hi[2]), [h3] "v"(hi[3]), [h4] "v"(hi[4]), [h5] "v"(hi[5]), [h6] "v"(hi[6]),
          [h7] "v"(hi[7])
        : "vcc");
    return out;
}

// Writes key | id of every lane of m into LDS at byte address
// ldsAddr + 8 * (rank of the lane among m): EXEC is set to m (skipping the
// writes when it is empty), restored in the same asm block.  The key's and id's
// halves are OR-ed into the two words of a ds_write2_b32.
__device__ __forceinline__ void stage_write(uint64_t m, uint32_t ldsAddr, uint32_t klo, uint32_t khi, uint64_t id) {
    uint64_t save;
    uint32_t t, w0, w1;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_mov_b64 exec, %[m]\n\t"
        "s_cbranch_execz .Lstage_skip%=\n\t"
        "v_mbcnt_lo_u32_b32 %[t], exec_lo, 0\n\t"
        "v_mbcnt_hi_u32_b32 %[t], exec_hi, %[t]\n\t"
        "v_lshl_add_u32 %[t], %[t], 3, %[a]\n\t"
        "v_or_b32 %[w0], %[il], %[kl]\n\t"
        "v_or_b32 %[w1], %[ih], %[kh]\n\t"
        "ds_write2_b32 %[t], %[w0], %[w1] offset1:1\n\t"
        ".Lstage_skip%=:\n\t"
        "s_mov_b64 exec, %[sv]"
        : [sv] "=&s"(save), [t] "=&v"(t), [w0] "=&v"(w0), [w1] "=&v"(w1)
        : [m] "s"(m), [a] "s"(ldsAddr), [kl] "v"(klo), [kh] "v"(khi), [il] "s"(static_cast<uint32_t>(id)),
          [ih] "s"(static_cast<uint32_t>(id >> 32))
        : "memory");
}

template <int DM>
__device__ __forceinline__ uint64_t box_mask(uint64_t m, const double* x, const double* lo, const double* hi) {
    if constexpr (DM == 4) {
        return box_mask4(m, x, lo, hi);
    } else {
#pragma unroll
        for (int d = 0; d < DM; d += 8) m = box_mask8(m, x + d, lo + d, hi + d);
        return m;
    }
}

// Queries of work item (cluster, 64-query slice) `it` for this lane: *q and
// whether the lane holds one.  Level 0 (no clusters) takes query slices.
__device__ __forceinline__ bool slice_query(const RangeStepArgs& A, uint64_t c, uint64_t slice, int lane,
                                            uint64_t* q) {
    const uint64_t g0 = ((ConstU64)A.off)[c];
    const uint64_t gn = ((ConstU64)A.off)[c + 1] - g0;
    const uint64_t i = slice * kWave + lane;
    const bool valid = i < gn && g0 + i < A.qcap;
    *q = valid ? A.qlist[g0 + i] : 0;
    return valid;
}

// The exact point-in-box filter of the last level (RangePruning.cu:5-25 with
// r = 0 reduces to x in [lo, hi] per dimension), over kd leaves.  The rows of
// every last-level cluster are stored in kd order at engine creation, so each
// run of kLeafRows consecutive rows (a leaf) is spatially tight and carries an
// outward-rounded float bounding box.  A wave takes a work item (cluster, 64-
// query slice) whose slice is a multiple of S = final_sets(DM) together with
// the next S - 1 slices of the same cluster (the other items are skipped): up
// to 64 S queries, every lane holding S outward-rounded float boxes.
// Per leaf, a float v_cmpx chain per set gives the queries whose box meets the
// leaf's box; those queries (about a fifth at the headline workload) are
// compacted through LDS into full 64-lane chunks, each lane loads its query's
// FP64 box, and the leaf's rows -- scalar loads into SGPRs, the next row in
// flight while the current one is compared -- are tested with the FP64 v_cmpx
// chain of box_mask.  A query whose box misses a leaf's box has no hit among
// its rows, so the hits are exactly those of the whole cluster.  Hits go to the
// wave's LDS stage and are appended to the global list in runs (one atomic per
// run).  Items come from a work queue.
// Box sets per lane: 4 up to D = 8, 2 up to 16, 1 above (register budget).
constexpr int final_sets(int DM) { return DM <= 8 ? 4 : (DM <= 16 ? 2 : 1); }

template <int DM, bool EXACT, int STAGE>
__global__ __launch_bounds__(kStepWaves * kWave) void k_range_final(RangeStepArgs A) {
    constexpr int SETS = final_sets(DM);
    __shared__ uint64_t s_hit[kStepWaves][STAGE];
    __shared__ uint32_t s_q[kStepWaves][SETS * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = EXACT ? DM : A.D;
    uint64_t nItems = ((ConstU64)A.nItems)[0];
    if (nItems > A.itemCap) nItems = A.itemCap;
    uint64_t* stage = s_hit[wave];
    const uint32_t stageAddr = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((LdsU64)stage));
    uint32_t* qs = s_q[wave];
    uint32_t staged = 0;
    uint64_t wTests = 0, wRows = 0, wLeafTests = 0;
    const uint64_t chunk = queue_chunk(nItems, nw);
    const ConstF32 lbox = (ConstF32)A.leafBox;
    const uint64_t rowBytes = static_cast<uint64_t>(D) * 8;
    uint64_t ib, ie;
    while (next_items(A.queue, chunk, nItems, lane, &ib, &ie))
    for (uint64_t item = ib; item < ie; ++item) {
        const uint64_t it = ((ConstU64)A.items)[item];
        const uint64_t c = it >> kSliceBits;
        const uint64_t slice = it & ((1ull << kSliceBits) - 1);
        if (slice % SETS) continue;  // done with the first slice of its group
        // the lane's query of every set (bit s of vbits: set s holds one)
        uint32_t q[SETS];
        uint32_t vbits = 0;
#pragma unroll
        for (int s = 0; s < SETS; ++s) {
            uint64_t qq;
            vbits |= static_cast<uint32_t>(slice_query(A, c, slice + s, lane, &qq)) << s;
            q[s] = static_cast<uint32_t>(qq);
        }
        // every set's float boxes in one batch of loads (invalid lanes read box 0)
        float flo[SETS][DM], fhi[SETS][DM];
#pragma unroll
        for (int s = 0; s < SETS; ++s) {
            const float* fb = A.fboxes + static_cast<uint64_t>(q[s]) * 2 * DM;
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                flo[s][d] = fb[d];
                fhi[s][d] = fb[DM + d];
            }
        }
        const uint64_t jb = ((ConstU64)A.colResPrev)[c], je = ((ConstU64)A.colResPrev)[c + 1];
        if (__ballot(vbits != 0) == 0 || jb >= je) continue;
        wRows += je - jb;
        const uint64_t leaf0 = ((ConstU64)A.leafOff)[c];
        const uint32_t nLeaves = static_cast<uint32_t>((je - jb + kLeafRows - 1) / kLeafRows);
        // one vector load per 64-byte line of the cluster's rows, ids and leaf
        // boxes pulls them into L2 at once, ahead of their scalar loads (the
        // loaded words are consumed at the end of the item)
        uint32_t pf = 0;
        {
            const uint32_t* pr = reinterpret_cast<const uint32_t*>(A.coords + jb * D);
            const uint64_t nlr = ((je - jb) * D * 8 + 63) / 64;
            for (uint64_t l = lane; l < nlr; l += kWave) pf ^= pr[l * 16];
            const uint32_t* pi = reinterpret_cast<const uint32_t*>(A.ids + jb);
            const uint64_t nli = ((je - jb) * 8 + 63) / 64;
            for (uint64_t l = lane; l < nli; l += kWave) pf ^= pi[l * 16];
            const uint32_t* pb = reinterpret_cast<const uint32_t*>(A.leafBox + leaf0 * 2 * D);
            const uint64_t nlb = (static_cast<uint64_t>(nLeaves) * 2 * D * 4 + 63) / 64;
            for (uint64_t l = lane; l < nlb; l += kWave) pf ^= pb[l * 16];
        }
        for (uint32_t lf = 0; lf < nLeaves; ++lf) {
            // queries whose box meets the leaf's box, per set
            float llo[DM], lhi[DM];
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                llo[d] = d < D ? lbox[(leaf0 + lf) * 2 * D + d] : 0.0f;
                lhi[d] = d < D ? lbox[(leaf0 + lf) * 2 * D + D + d] : 0.0f;
            }
            // compact those queries into the wave's list
            uint32_t total = 0;
#pragma unroll
            for (int s = 0; s < SETS; ++s) {
                const uint64_t vm = __ballot((vbits >> s) & 1u);
                if (vm == 0) continue;
                const uint64_t m = aabb_mask<DM>(vm, llo, lhi, flo[s], fhi[s]);
                wLeafTests += __popcll(vm);
                if ((m >> lane) & 1ull) {
                    const uint32_t r = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
                    qs[total + r] = q[s];
                }
                total += static_cast<uint32_t>(__popcll(m));
            }
            if (total == 0) continue;
            wave_lds_sync();
            const uint64_t r0 = jb + static_cast<uint64_t>(lf) * kLeafRows;
            const uint32_t nr = static_cast<uint32_t>(je - r0 < static_cast<uint64_t>(kLeafRows) ? je - r0
                                                                                                   : kLeafRows);
            for (uint32_t k0 = 0; k0 < total; k0 += kWave) {
                const bool valid = k0 + lane < total;
                const uint64_t qq = valid ? qs[k0 + lane] : 0;
                const uint64_t vmask = __ballot(valid);
                double lo[DM], hi[DM];
#pragma unroll
                for (int d = 0; d < DM; ++d) {
                    lo[d] = d < D ? A.boxes[qq * 2 * D + d] : 0.0;
                    hi[d] = d < D ? A.boxes[qq * 2 * D + D + d] : 0.0;
                }
                const uint64_t key = qq << A.idBits;
                const uint32_t klo = static_cast<uint32_t>(key), khi = static_cast<uint32_t>(key >> 32);
                wTests += static_cast<uint64_t>(nr) * static_cast<uint64_t>(__popcll(vmask));
                // wait for the boxes here, once, rather than inside the row loop
                asm volatile("" : : "v"(lo[0]), "v"(hi[DM - 1]));
                // Rows in pairs, ping-pong between the register sets A and B:
                // a pair is compared while the next pair's scalar loads are in
                // flight.  Scalar loads complete out of order, so a wait on the
                // pair in hand would also wait on any load issued after it: the
                // empty asm consumes a word of each row in hand (the wait lands
                // before it) and yields the addresses the next loads use, so
                // they are issued after that wait.  Loads run up to three rows
                // past the leaf (the arrays have spare rows at the end); those
                // are never tested.  A row's hits go straight into the LDS
                // stage, which is flushed before a pair whenever it could not
                // take 128 more.
                uint64_t pa = (uint64_t)((ConstF64)A.coords + r0 * D);
                uint64_t pi = (uint64_t)((ConstU64)A.ids + r0);
                double a0[DM], a1[DM], b0[DM], b1[DM];
                uint64_t ia0, ia1, ib0, ib1;
                auto load2 = [&](double* x0, double* x1, uint64_t& i0, uint64_t& i1) {
                    scalar_row<DM, EXACT>(x0, (ConstF64)pa, 0, D);
                    scalar_row<DM, EXACT>(x1, (ConstF64)(pa + rowBytes), 0, D);
                    i0 = ((ConstU64)pi)[0];
                    i1 = ((ConstU64)pi)[1];
                };
                auto test1 = [&](const double* x, uint64_t id) {
                    const uint64_t m = box_mask<DM>(vmask, x, lo, hi);
                    if (m) {
                        stage_write(m, stageAddr + staged * 8, klo, khi, id);
                        staged += static_cast<uint32_t>(__popcll(m));
                    }
                };
                auto room = [&]() {
                    if (staged > static_cast<uint32_t>(STAGE - 2 * kWave)) {
                        flush_hits_lean(A, stage, staged, lane);
                        staged = 0;
                    }
                };
                load2(a0, a1, ia0, ia1);
                for (uint32_t r = 0; r < nr; r += 4) {
                    pa += 2 * rowBytes;
                    pi += 16;
                    asm volatile("" : "+s"(pa), "+s"(pi) : "s"(a0[0]), "s"(a1[0]), "s"(ia0), "s"(ia1));
                    load2(b0, b1, ib0, ib1);
                    room();
                    test1(a0, ia0);
                    if (r + 1 < nr) test1(a1, ia1);
                    if (r + 2 >= nr) break;
                    pa += 2 * rowBytes;
                    pi += 16;
                    asm volatile("" : "+s"(pa), "+s"(pi) : "s"(b0[0]), "s"(b1[0]), "s"(ib0), "s"(ib1));
                    load2(a0, a1, ia0, ia1);
                    room();
                    test1(b0, ib0);
                    if (r + 3 < nr) test1(b1, ib1);
                }
            }
            wave_lds_sync();  // the list is rewritten by the next leaf
        }
        asm volatile("" : : "v"(pf));
    }
    flush_hits(A, stage, staged, lane);
    // work done by this wave: row x box tests, child rows, leaf x box tests
    if (lane == 0 && A.work) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work), wTests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work + 1), wRows);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work + 2), wLeafTests);
    }
}

// Lanes of m whose row x lies in the box [lo, hi] (all in VGPRs; the box the
// same in every lane) in eight dimensions: the v_cmpx chain of box_mask8 with
// the operands' roles swapped (lo <= x and hi >= x, false for NaN as x >= lo
// && x <= hi).
__device__ __forceinline__ uint64_t rows_in_box8(uint64_t m, const double* x, const double* lo, const double* hi) {
    uint64_t out, save;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        "s_and_b64 exec, exec, %[m]\n\t"
        "v_cmpx_le_f64 vcc, %[l0], %[x0]\n\t"
        "v_cmpx_ge_f64 vcc, %[h0], %[x0]\n\t"
        "v_cmpx_le_f64 vcc, %[l1], %[x1]\n\t"
        "v_cmpx_ge_f64 vcc, %[h1], %[x1]\n\t"
        "v_cmpx_le_f64 vcc, %[l2], %[x2]\n\t"
        "v_cmpx_ge_f64 vcc, %[h2], %[x2]\n\t"
        "v_cmpx_le_f64 vcc, %[l3], %[x3]\n\t"
        "v_cmpx_ge_f64 vcc, %[h3], %[x3]\n\t"
        "v_cmpx_le_f64 vcc, %[l4], %[x4]\n\t"
        "v_cmpx_ge_f64 vcc, %[h4], %[x4]\n\t"
        "v_cmpx_le_f64 vcc, %[l5], %[x5]\n\t"
        "v_cmpx_ge_f64 vcc, %[h5], %[x5]\n\t"
        "v_cmpx_le_f64 vcc, %[l6], %[x6]\n\t"
        "v_cmpx_ge_f64 vcc, %[h6], %[x6]\n\t"
        "v_cmpx_le_f64 vcc, %[l7], %[x7]\n\t"
        "v_cmpx_ge_f64 vcc, %[h7], %[x7]\n\t"
        "s_mov_b64 %[out], exec\n\t"
        "s_mov_b64 exec, %[sv]"
        : [out] "=&s"(out), [sv] "=&s"(save)
        : [m] "s"(m), [x0] "v"(x[0]), [x1] "v"(x[1]), [x2] "v"(x[2]), [x3] "v"(x[3]), [x4] "v"(x[4]),
          [x5] "v"(x[5]), [x6] "v"(x[6]), [x7] "v"(x[7]), [l0] "v"(lo[0]), [l1] "v"(lo[1]), [l2] "v"(lo[2]),
          [l3] "v"(lo[3]), [l4] "v"(lo[4]), [l5] "v"(lo[5]), [l6] "v"(lo[6]), [l7] "v"(lo[7]), [h0] "v"(hi[0]),
          [h1] "v"(hi[1]), [h2] "v"(hi[2]), [h3] "v"(hi[3]), [h4] "v"(hi[4]), [h5] "v"(hi[5]), [h6] "v"(hi[6]),
          [h7] "v"(hi[7])
        : "vcc");
    return out;
}
template <int DM>
__device__ __forceinline__ uint64_t rows_in_box(uint64_t m, const double* x, const double* lo, const double* hi) {
    if constexpr (DM == 4) {
        const double xx[8] = {x[0], x[1], x[2], x[3], 0.0, 0.0, 0.0, 0.0};
        const double ll[8] = {lo[0], lo[1], lo[2], lo[3], 0.0, 0.0, 0.0, 0.0};
        const double hh[8] = {hi[0], hi[1], hi[2], hi[3], 0.0, 0.0, 0.0, 0.0};
        return rows_in_box8(m, xx, ll, hh);
    } else {
#pragma unroll
        for (int d = 0; d < DM; d += 8) m = rows_in_box8(m, x + d, lo + d, hi + d);
        return m;
    }
}

// The exact filter, row-major, for last-level clusters kept by at most rowMax
// queries (the k_range_final kernel skips them).  Lane = query wastes most of
// a wave on such clusters, and their rows are read for few tests, so here one
// wave takes the whole cluster (its one slice of queries): first lane = leaf,
// every query's float box (scalar) against up to 64 leaf boxes, giving each
// query a mask of the leaves that meet it; then for every run of 64 rows (four
// leaves) that meets some query, lane = row, the rows loaded once with vector
// loads and tested against each query whose mask holds one of the four leaves
// (only the rows of those leaves, by the FP64 v_cmpx chain with the box in
// SGPRs).  Rows of leaves no query meets are never read.  Same predicate and
// hits as k_range_final.
// Queries per cluster the row-major filter holds (their FP64 boxes in LDS,
// 4 KB per wave).
constexpr int rows_max_queries(int DM) { return 256 / DM; }

template <int DM, bool EXACT, int STAGE>
__global__ __launch_bounds__(kStepWaves * kWave) void k_range_rows(RangeStepArgs A) {
    constexpr int kMaxQ = rows_max_queries(DM);
    __shared__ uint64_t s_hit[kStepWaves][STAGE];
    __shared__ uint64_t s_lm[kStepWaves][kWave];
    __shared__ double s_box[kStepWaves][kMaxQ][2 * DM];
    __shared__ uint64_t s_hdr[kStepWaves][kRowUnit][5];  // a unit's clusters: list range, rows, first leaf
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kStepWaves;
    const int D = EXACT ? DM : A.D;
    // work: the non-empty units of kRowUnit list slots
    const uint64_t nItems = *((ConstU64)A.sparseUnitCount);
    uint64_t* stage = s_hit[wave];
    uint64_t* lm = s_lm[wave];
    double(*bx)[2 * DM] = s_box[wave];
    uint64_t(*hdr)[5] = s_hdr[wave];
    uint32_t staged = 0;
    uint64_t wTests = 0, wRows = 0, wLeafTests = 0;
    const uint64_t chunk = queue_chunk(nItems, nw);
    uint64_t ib, ie;
    while (next_items(A.queue + 1, chunk, nItems, lane, &ib, &ie))
    for (uint64_t u = ib; u < ie; ++u) {
        const uint64_t s0 = ((ConstU64)A.sparseUnits)[u] * kRowUnit;
        const uint64_t tileEnd = s0 / kScanThreads * kScanThreads + ((ConstU64)A.sparseCnt)[s0 / kScanThreads];
        const int ns = static_cast<int>(s0 + kRowUnit < tileEnd ? kRowUnit : tileEnd - s0);
        // the unit's clusters and their list and row ranges, lane = slot, in a
        // few vector loads (rather than a chain of scalar loads per cluster)
        if (lane < ns) {
            const uint64_t hc = A.sparse[s0 + lane];
            const uint64_t h0 = A.off[hc], h1 = A.off[hc + 1], h2 = A.colResPrev[hc], h3 = A.colResPrev[hc + 1],
                           h4 = A.leafOff[hc];
            hdr[lane][0] = h0;
            hdr[lane][1] = h1;
            hdr[lane][2] = h2;
            hdr[lane][3] = h3;
            hdr[lane][4] = h4;
        }
        wave_lds_sync();
        // queries whose list entry was stored (an overflowed list drops the rest;
        // the batch is then rerun)
        auto stored = [&](uint64_t g0, uint64_t gn) {
            return static_cast<uint32_t>(g0 >= A.qcap ? 0 : (A.qcap - g0 < gn ? A.qcap - g0 : gn));
        };
        // lane k < nQ: query k of the cluster, loaded one cluster ahead
        uint32_t qNext = 0;
        {
            const uint64_t g0 = uniform_u64(hdr[0][0]), gn = uniform_u64(hdr[0][1]) - g0;
            if (static_cast<uint32_t>(lane) < stored(g0, gn) && gn <= A.rowMax) qNext = A.qlist[g0 + lane];
        }
        for (int k = 0; k < ns; ++k) {
            const uint64_t g0 = uniform_u64(hdr[k][0]);
            const uint64_t gn = uniform_u64(hdr[k][1]) - g0;
            const uint64_t jb = uniform_u64(hdr[k][2]), je = uniform_u64(hdr[k][3]);
            const uint32_t qv = qNext;
            if (k + 1 < ns) {
                const uint64_t g0n = uniform_u64(hdr[k + 1][0]), gnn = uniform_u64(hdr[k + 1][1]) - g0n;
                qNext = static_cast<uint32_t>(lane) < stored(g0n, gnn) && gnn <= A.rowMax ? A.qlist[g0n + lane] : 0u;
            }
            if (gn == 0 || gn > A.rowMax) continue;
            const uint32_t nQ = stored(g0, gn);
            if (nQ == 0 || jb >= je) continue;
            // the queries' FP64 boxes, loaded now and written to LDS after the
            // first leaf tests, so that both loads are in flight together
            // (lanes past nQ read query qv = 0's box and keep it)
            double qb[2 * DM];
            {
                const double* b = A.boxes + static_cast<uint64_t>(qv) * 2 * D;
#pragma unroll
                for (int d = 0; d < DM; ++d) {
                    qb[d] = (EXACT || d < D) ? b[d] : 0.0;
                    qb[DM + d] = (EXACT || d < D) ? b[D + d] : 0.0;
                }
            }
            bool boxesStored = false;
            const uint64_t leaf0 = uniform_u64(hdr[k][4]);
            const uint32_t nLeaves = static_cast<uint32_t>((je - jb + kLeafRows - 1) / kLeafRows);
            for (uint32_t lc = 0; lc < nLeaves; lc += kWave) {
                const uint32_t nl = nLeaves - lc < static_cast<uint32_t>(kWave) ? nLeaves - lc : kWave;
                // (query, leaf) pairs, lane = pair: the query's float box against the
                // leaf's (outward-rounded both, so a miss is exact); lane k ends with
                // query k's mask of the chunk's leaves that meet it
                lm[lane] = 0;
                wave_lds_sync();
                const uint32_t pairs = nl * nQ;
                for (uint32_t p0 = 0; p0 < pairs; p0 += kWave) {
                    const uint32_t p = p0 + lane;
                    const bool pv = p < pairs;
                    const uint32_t k = pv ? p / nl : 0, l = pv ? p - k * nl : 0;
                    const uint64_t q = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(k) << 2,
                                                                                          static_cast<int>(qv)));
                    const uint64_t lb = (leaf0 + lc + l) * 2 * D;
                    float llo[DM], lhi[DM], qlo[DM], qhi[DM];
#pragma unroll
                    for (int d = 0; d < DM; ++d) {
                        llo[d] = d < D ? A.leafBox[lb + d] : 0.0f;
                        lhi[d] = d < D ? A.leafBox[lb + D + d] : 0.0f;
                        qlo[d] = A.fboxes[q * 2 * DM + d];
                        qhi[d] = A.fboxes[q * 2 * DM + DM + d];
                    }
                    bool meet = pv;
#pragma unroll
                    for (int d = 0; d < DM; ++d) meet = meet & (qlo[d] <= lhi[d]) & (qhi[d] >= llo[d]);
                    if (meet) atomicOr(reinterpret_cast<unsigned long long*>(&lm[k]), 1ull << l);
                }
                wLeafTests += pairs;
                if (!boxesStored) {
                    if (static_cast<uint32_t>(lane) < nQ) {
#pragma unroll
                        for (int d = 0; d < 2 * DM; ++d) bx[lane][d] = qb[d];
                    }
                    boxesStored = true;
                }
                wave_lds_sync();
                const uint64_t lmv = static_cast<uint32_t>(lane) < nQ ? lm[lane] : 0ull;
                uint64_t uni = lmv;
                for (int o = 1; o < kWave; o <<= 1) uni |= __shfl_xor(uni, o, kWave);
                uni = uniform_u64(uni);
                // runs of 64 rows (four leaves) that meet some query
                for (uint32_t g = 0; 4 * g < nl; ++g) {
                    const uint32_t gm = static_cast<uint32_t>(uni >> (4 * g)) & 15u;
                    if (gm == 0) continue;
                    const uint64_t r0 = jb + (static_cast<uint64_t>(lc) + 4 * g) * kLeafRows;
                    const bool rv = r0 + lane < je;
                    const uint64_t row = rv ? r0 + lane : r0;
                    double x[DM];
#pragma unroll
                    for (int d = 0; d < DM; ++d) x[d] = (EXACT || d < D) ? A.coords[row * D + d] : 0.0;
                    const uint64_t id = A.ids[row];
                    const uint64_t rvm = __ballot(rv);
                    uint64_t leafRows = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if ((gm >> b) & 1u) leafRows |= 0xFFFFull << (16 * b);
                    wRows += __popcll(rvm & leafRows);
                    // queries meeting one of the run's leaves, each one's box read from
                    // LDS (the same address in every lane) and tested
                    const uint32_t qmv = static_cast<uint32_t>(lmv >> (4 * g)) & 15u;
                    for (uint64_t act = __ballot(qmv != 0); act; act &= act - 1) {
                        const int k = __builtin_ctzll(act);
                        const uint32_t qm = __builtin_amdgcn_readlane(qmv, k);
                        const uint64_t q = __builtin_amdgcn_readlane(qv, k);
                        uint64_t m = 0;
#pragma unroll
                        for (int b = 0; b < 4; ++b)
                            if ((qm >> b) & 1u) m |= 0xFFFFull << (16 * b);
                        m &= rvm;
                        double lo[DM], hi[DM];
#pragma unroll
                        for (int d = 0; d < DM; ++d) {
                            lo[d] = bx[k][d];
                            hi[d] = bx[k][DM + d];
                        }
                        if (staged > static_cast<uint32_t>(STAGE - kWave)) {
                            flush_hits(A, stage, staged, lane);
                            staged = 0;
                        }
                        const uint64_t hit = rows_in_box<DM>(m, x, lo, hi);
                        wTests += __popcll(m);
                        if (hit) {
                            if ((hit >> lane) & 1ull) {
                                const uint32_t r =
                                    __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(hit >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(hit), 0u));
                                stage[staged + r] = (q << A.idBits) | id;
                            }
                            staged += static_cast<uint32_t>(__popcll(hit));
                        }
                    }
                }
                wave_lds_sync();  // the leaf masks are rewritten by the next chunk
            }
        }
    }
    flush_hits(A, stage, staged, lane);
    if (lane == 0 && A.work) {
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work), wTests);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work + 1), wRows);
        atomicAdd(reinterpret_cast<unsigned long long*>(A.work + 2), wLeafTests);
    }
}

// Dimensions of the kernel instantiation that handles D (4, 8, 16 or 32).
inline int dm_of(int D) { return D <= 4 ? 4 : (D <= 8 ? 8 : (D <= 16 ? 16 : 32)); }

// Outward-rounded float copy of the boxes for the exact filter's leaf tests:
// out[q * 2DM + d] = rd(lo_d), out[q * 2DM + DM + d] = ru(hi_d), padding 0.
__global__ void k_box_f32(const double* __restrict__ boxes, uint64_t nq, int D, int DM, float* __restrict__ out) {
    const uint64_t n = nq * 2 * DM;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t q = e / (2 * DM);
        const int k = static_cast<int>(e - q * 2 * DM), hi = k >= DM, d = hi ? k - DM : k;
        float v = 0.0f;
        if (d < D) {
            const double x = boxes[q * 2 * D + (hi ? D : 0) + d];
            v = hi ? __double2float_ru(x) : __double2float_rd(x);
        }
        out[e] = v;
    }
}

// Group scan over the clusters of one level: off[i] = exclusive prefix of
// cnt[i], items for every 64-query slice of every cluster, totals.
__global__ __launch_bounds__(kScanThreads) void k_group_scan(const uint64_t* __restrict__ cnt, uint64_t n,
                                                             uint64_t* __restrict__ off, uint64_t* __restrict__ items,
                                                             uint64_t itemCap, uint64_t* nItems, uint64_t* overflow,
                                                             uint64_t* status, ScanCtrl* ctrl, uint64_t rowMax,
                                                             SparseList list) {
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_c[kScanThreads / kWave], s_s[kScanThreads / kWave];
    __shared__ uint32_t s_p[kScanThreads / kWave];
    const uint32_t epoch = launch_epoch(ctrl);
    __shared__ uint64_t s_ex0, s_ex1;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const uint64_t ntiles = n == 0 ? 1 : (n + kScanThreads - 1) / kScanThreads;
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(reinterpret_cast<unsigned long long*>(&ctrl->tileCounter), 1ull);
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= ntiles) break;
        const uint64_t i = tile * kScanThreads + tid;
        const uint64_t c = i < n ? cnt[i] : 0;
        // with a row-major list: groups of 1..rowMax entries get no work
        // items; their ids go to the tile's kScanThreads slots of the list
        // instead (in order), the tile's count and its units to the list
        const bool sp = list.slots && c > 0 && c <= rowMax;
        const uint64_t sl = sp ? 0 : (c + kWave - 1) / kWave;
        const uint64_t spm = __ballot(sp);
        uint64_t xc = c, xs = sl;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint64_t yc = __shfl_up(xc, o, kWave), ys = __shfl_up(xs, o, kWave);
            if (lane >= o) {
                xc += yc;
                xs += ys;
            }
        }
        if (lane == kWave - 1) {
            s_c[wid] = xc;
            s_s[wid] = xs;
            s_p[wid] = static_cast<uint32_t>(__popcll(spm));
        }
        __syncthreads();
        if (list.slots) {
            uint32_t p = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < kScanThreads / kWave; ++w) {
                if (w < wid) p += s_p[w];
                tot += s_p[w];
            }
            if (sp) list.slots[tile * kScanThreads + p + __popcll(spm & ((1ull << lane) - 1))] = i;
            if (tid == 0) {
                list.counts[tile] = tot;
                const uint32_t nu = (tot + kRowUnit - 1) / kRowUnit;
                if (nu) {
                    const uint64_t base =
                        atomicAdd(reinterpret_cast<unsigned long long*>(list.unitCount), static_cast<uint64_t>(nu));
                    for (uint32_t k = 0; k < nu; ++k) list.units[base + k] = tile * (kScanThreads / kRowUnit) + k;
                }
            }
        }
        if (wid == 0) {
            uint64_t tc = lane < kScanThreads / kWave ? s_c[lane] : 0;
            uint64_t ts = lane < kScanThreads / kWave ? s_s[lane] : 0;
            tc = wave_sum_u64(tc);
            ts = wave_sum_u64(ts);
            uint64_t e0, e1;
            tile_prefix2(status, epoch, tile, lane, tc, ts, e0, e1);
            if (lane == 0) {
                s_ex0 = e0;
                s_ex1 = e1;
                if (tile == ntiles - 1) {
                    off[n] = e0 + tc;
                    *nItems = e1 + ts;
                    if (e1 + ts > itemCap) atomicOr(reinterpret_cast<unsigned long long*>(overflow), 2ull);
                }
            }
        }
        __syncthreads();
        uint64_t pc = s_ex0 + xc - c, ps = s_ex1 + xs - sl;
        for (int w = 0; w < wid; ++w) {
            pc += s_c[w];
            ps += s_s[w];
        }
        if (i < n) {
            off[i] = pc;
            for (uint64_t k = 0; k < sl; ++k)
                if (ps + k < itemCap) items[ps + k] = (i << kSliceBits) | k;
        }
        __syncthreads();
    }
    if (tid == 0) scan_exit(ctrl);
}

// Counting-sort scatter of (id, q) pairs into grouped query lists; leaves cnt
// zeroed.  A workgroup takes chunks of kAggChunk pairs; with mesh <= kAggLds it
// ranks the chunk's pairs per id in LDS and reserves each id's run with one
// global atomic (a run's slots count down from cnt[id]), else one atomic per
// pair.  The order inside a list is not fixed.  Slots at or beyond cap (counts
// of pairs that overflowed the pair buffer, a batch that is rerun) are not
// written.
__global__ __launch_bounds__(kAggThreads) void k_group_scatter(const uint64_t* __restrict__ pairs, const uint64_t* dn,
                                                               uint64_t cap, uint64_t mesh,
                                                               uint64_t* __restrict__ cnt,
                                                               const uint64_t* __restrict__ off,
                                                               uint32_t* __restrict__ qlist) {
    __shared__ uint32_t lc[kAggLds];
    const int tid = threadIdx.x;
    uint64_t n = *dn;
    if (n > cap) n = cap;
    if (mesh > kAggLds) {
        for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + tid; i < n;
             i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
            const uint64_t p = pairs[i];
            const uint64_t id = p >> kQueryBits;
            if (id >= mesh) continue;  // sentinel of an overflowed allocation
            const uint32_t q = static_cast<uint32_t>(p & ((1ull << kQueryBits) - 1));
            const uint64_t pos = atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[id]), ~0ull) - 1;
            if (off[id] + pos < cap) qlist[off[id] + pos] = q;
        }
        return;
    }
    for (uint64_t i = tid; i < mesh; i += kAggThreads) lc[i] = 0;
    __syncthreads();
    const uint64_t chunks = (n + kAggChunk - 1) / kAggChunk;
    for (uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        uint64_t p[kAggPer];
        uint32_t rank[kAggPer];
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t i = c * kAggChunk + static_cast<uint64_t>(r) * kAggThreads + tid;
            p[r] = i < n ? pairs[i] : ~0ull;
            const uint64_t id = p[r] >> kQueryBits;
            rank[r] = id < mesh ? atomicAdd(&lc[id], 1u) : 0;
        }
        __syncthreads();
        for (uint64_t id = tid; id < mesh; id += kAggThreads) {
            const uint32_t k = lc[id];
            if (k) lc[id] = static_cast<uint32_t>(
                       atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[id]), 0ull - k) - k);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t id = p[r] >> kQueryBits;
            if (id < mesh && off[id] + lc[id] + rank[r] < cap)
                qlist[off[id] + lc[id] + rank[r]] = static_cast<uint32_t>(p[r] & ((1ull << kQueryBits) - 1));
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kAggPer; ++r) {
            const uint64_t id = p[r] >> kQueryBits;
            if (id < mesh) lc[id] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- radix sort
// Single-sweep LSD radix sort of u64 keys: digits of RB = 8 or 9 bits, tiles of
// 8192 keys, the key count on the device.  A counting kernel reads the keys
// once for the first digit's counts (and, optionally, the keys of every group
// key >> groupShift); each pass is then one kernel that ranks a tile's keys by
// digit, publishes the tile's digit counts, takes every digit's exclusive
// prefix over the earlier tiles by decoupled look-back, lays the tile out in
// digit order in LDS (two halves of 4096 keys) and writes each digit's run
// with consecutive threads on consecutive addresses, while counting the next
// pass's digit from the keys it holds.  Tiles are handed out by a ticket
// counter, so a tile only waits on tiles that resident workgroups already hold.

constexpr int kRsThreads = 512;
constexpr int kRsWaves = kRsThreads / kWave;
constexpr int kRsPer = 16;
constexpr int kRsTile = kRsThreads * kRsPer;
constexpr int kRsHalf = kRsTile / 2;  // keys staged in LDS per round
constexpr int kRsMaxPasses = 8;
constexpr int kRsMaxBins = 512;
constexpr int kRsGroupLds = 8192;  // groups counted in LDS (more: global atomics)
// scratch (u32 words): [0, 64) tile tickets of each pass (and two u64 words of
// the group scan at [32, 36)), then the digit counts of each pass, then two
// look-back regions of tiles x bins words, used by alternate passes, then the
// u64 group counts
constexpr uint64_t kRsCounts0 = 64;
constexpr uint64_t kRsHdrWords = kRsCounts0 + static_cast<uint64_t>(kRsMaxPasses) * kRsMaxBins;
// look-back word: flag (aggregate / inclusive prefix) in the top two bits
constexpr uint32_t kRsAgg = 1u << 30, kRsInc = 2u << 30, kRsVal = (1u << 30) - 1;
constexpr int kRsProbe = 8;  // predecessors read per look-back round

// Lanes of the wave whose RB-bit digit equals this lane's, among the lanes set
// in `valid`: one ballot per digit bit, each folded into the running mask of
// ruled-out lanes by one bitop3 per half (out | (ballot ^ smeared bit): table
// 0xF0 | (0xCC ^ 0xAA) = 0xF6 over the operands (out, ballot, bit)).
template <int RB>
__device__ __forceinline__ void digit_peers(uint32_t dig, uint64_t valid, uint32_t& peersLo, uint32_t& peersHi) {
    uint32_t lo = ~static_cast<uint32_t>(valid), hi = ~static_cast<uint32_t>(valid >> 32);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const uint32_t sm = static_cast<uint32_t>(static_cast<int32_t>(dig << (31 - b)) >> 31);
        const uint64_t bal = __ballot(sm != 0);
        lo = __builtin_amdgcn_bitop3_b32(lo, static_cast<uint32_t>(bal), sm, 0xF6);
        hi = __builtin_amdgcn_bitop3_b32(hi, static_cast<uint32_t>(bal >> 32), sm, 0xF6);
    }
    peersLo = ~lo;
    peersHi = ~hi;
}

// Exclusive scan of the kBins digit values of a workgroup of kRsThreads threads
// (thread tid holds digit tid; threads past kBins hold 0); with barriers.
template <int kBins>
__device__ __forceinline__ uint32_t bins_excl_scan(uint32_t v, uint32_t* wsum) {
    static_assert(kBins <= kRsThreads, "one digit per thread");
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    uint32_t x = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[wid] = x;
    __syncthreads();
    uint32_t p = x - v;
#pragma unroll
    for (int w = 0; w < kRsWaves; ++w)
        if (w < wid) p += wsum[w];
    __syncthreads();
    return p;
}

// Clears the tickets, digit counts and group counts.
__global__ void k_rs_begin(uint32_t* __restrict__ hdr, uint64_t* __restrict__ groupCnt, uint64_t nGroups) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    for (uint64_t i = i0; i < kRsHdrWords; i += stride) hdr[i] = 0;
    if (groupCnt)
        for (uint64_t i = i0; i < nGroups; i += stride) groupCnt[i] = 0;
}

// Counts of the first digit (at firstBit) over all keys and, with groupCnt, of
// every group key >> groupShift (a row whose 64 keys share a group adds once);
// also clears look-back region 0 for the tiles of this sort.
template <int RB>
__global__ __launch_bounds__(kRsThreads) void k_rs_counts(const uint64_t* __restrict__ keys, const uint64_t* dn,
                                                          uint64_t cap, int firstBit, uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ status0,
                                                          uint64_t* __restrict__ groupCnt, int groupShift,
                                                          uint64_t nGroups) {
    constexpr int kBins = 1 << RB;
    __shared__ uint32_t lc[kRsWaves][kBins];  // per-wave counters
    __shared__ uint32_t lg[kRsGroupLds];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const bool ldsGroups = groupCnt && nGroups <= kRsGroupLds;
    for (int i = tid; i < kRsWaves * kBins; i += kRsThreads) (&lc[0][0])[i] = 0;
    if (ldsGroups)
        for (int i = tid; i < static_cast<int>(nGroups); i += kRsThreads) lg[i] = 0;
    __syncthreads();
    uint64_t n = *dn;
    if (n > cap) n = cap;
    const uint64_t ntiles = (n + kRsTile - 1) / kRsTile;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        for (int d = tid; d < kBins; d += kRsThreads) status0[t * kBins + d] = 0;
        uint64_t k[kRsPer];
#pragma unroll
        for (int r = 0; r < kRsPer; ++r) {
            const uint64_t e = t * kRsTile + static_cast<uint64_t>(r) * kRsThreads + tid;
            k[r] = e < n ? keys[e] : 0;
        }
#pragma unroll
        for (int r = 0; r < kRsPer; ++r) {
            const bool ok = t * kRsTile + static_cast<uint64_t>(r) * kRsThreads + tid < n;
            if (ok) atomicAdd(&lc[wid][(k[r] >> firstBit) & (kBins - 1)], 1u);
            if (groupCnt) {
                const uint64_t g = k[r] >> groupShift;
                const uint64_t g0 = readlane_u64(g, 0);
                if (__ballot(ok) == ~0ull && __ballot(g != g0) == 0) {
                    if (lane == 0) {
                        if (ldsGroups) atomicAdd(&lg[g0], static_cast<uint32_t>(kWave));
                        else atomicAdd(reinterpret_cast<unsigned long long*>(groupCnt + g0), 64ull);
                    }
                } else if (ok) {
                    if (ldsGroups) atomicAdd(&lg[g], 1u);
                    else atomicAdd(reinterpret_cast<unsigned long long*>(groupCnt + g), 1ull);
                }
            }
        }
    }
    __syncthreads();
    for (int d = tid; d < kBins; d += kRsThreads) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kRsWaves; ++w) c += lc[w][d];
        if (c) atomicAdd(&counts[d], c);
    }
    if (ldsGroups)
        for (int i = tid; i < static_cast<int>(nGroups); i += kRsThreads)
            if (lg[i]) atomicAdd(reinterpret_cast<unsigned long long*>(groupCnt + i),
                                 static_cast<unsigned long long>(lg[i]));
}

// One stable pass on the digit at `shift`.  Within a tile, wave w owns keys
// [1024 w, 1024 (w + 1)), 64 per round, and ranks them against wave-private
// digit counters in LDS (peers of equal digit found with RB ballots; rank =
// counter + peers below the lane), with no workgroup barrier.  The waves'
// counts give the tile's digit counts (thread d owns digit d), which are
// published (tile 0: as its inclusive prefix) before the look-back; a key's
// place in the tile is its digit's start in the tile + the counts of waves < w
// + its rank, and a digit's run goes to the digit's global start + its count in
// earlier tiles.  out = key & outMask.  NEXT: the next digit's counts (at
// shift + RB) are added into nextCounts.
template <int RB, bool NEXT>
__global__ __launch_bounds__(kRsThreads, 4) void k_rs_onesweep(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                               const uint64_t* dn, uint64_t cap, int shift,
                                                               const uint32_t* __restrict__ counts,
                                                               uint32_t* __restrict__ nextCounts, uint32_t* status,
                                                               uint32_t* __restrict__ statusNext, uint32_t* ticket,
                                                               uint64_t outMask) {
    constexpr int kBins = 1 << RB;
    constexpr uint32_t kPerWave = static_cast<uint32_t>(kRsPer) * kWave;
    __shared__ uint32_t wc[kRsWaves][kBins];  // wave counters, then each wave's place in the tile per digit
    __shared__ uint32_t gofs[kBins];          // output position of a run = gofs[d] + place in the tile
    __shared__ uint32_t side[NEXT ? kBins : 1];
    __shared__ uint64_t sk[kRsHalf];
    __shared__ uint32_t wsum[kRsWaves];
    __shared__ uint32_t s_tile;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const bool owner = tid < kBins;  // thread tid owns digit tid
    uint64_t n = *dn;
    if (n > cap) n = cap;
    const uint64_t ntiles = (n + kRsTile - 1) / kRsTile;
    uint32_t* mycnt = wc[wid];
    if (NEXT)
        for (int i = tid; i < kBins; i += kRsThreads) side[i] = 0;
    // global digit starts (exclusive scan of the digit counts)
    const uint32_t gst = bins_excl_scan<kBins>(owner ? counts[tid] : 0u, wsum);
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint64_t t = s_tile;
        if (t >= ntiles) break;
        const uint64_t t0 = t * kRsTile;
        const uint32_t cnt = static_cast<uint32_t>(n - t0 < static_cast<uint64_t>(kRsTile) ? n - t0 : kRsTile);
        const uint32_t e0 = static_cast<uint32_t>(wid) * kPerWave + static_cast<uint32_t>(lane);
        uint64_t key[kRsPer];
        uint32_t rank[kRsPer];
#pragma unroll
        for (int r = 0; r < kRsPer; ++r) {
            const uint32_t e = e0 + static_cast<uint32_t>(r) * kWave;
            key[r] = e < cnt ? in[t0 + e] : 0;
        }
#pragma unroll
        for (int i = 0; i < (kBins + kWave - 1) / kWave; ++i)
            if (lane + i * kWave < kBins) mycnt[lane + i * kWave] = 0;
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < kRsPer; ++r) {
            const bool ok = e0 + static_cast<uint32_t>(r) * kWave < cnt;
            const uint32_t dig = ok ? static_cast<uint32_t>((key[r] >> shift) & (kBins - 1)) : 0;
            uint32_t plo, phi;
            digit_peers<RB>(dig, __ballot(ok), plo, phi);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
            const uint32_t pre = ok ? mycnt[dig] : 0;
            rank[r] = pre + below;
            wave_lds_sync();
            if (ok && below == 0) mycnt[dig] = pre + static_cast<uint32_t>(__popc(plo) + __popc(phi));
            wave_lds_sync();
            if (NEXT && ok) atomicAdd(&side[(key[r] >> (shift + RB)) & (kBins - 1)], 1u);
        }
        __syncthreads();
        uint32_t v = 0;
        if (owner) {
#pragma unroll
            for (int w = 0; w < kRsWaves; ++w) v += wc[w][tid];
            st_agent_u32(status + t * kBins + tid, (t == 0 ? kRsInc : kRsAgg) | v);
            if (statusNext) statusNext[t * kBins + tid] = 0;
        }
        const uint32_t ex = bins_excl_scan<kBins>(v, wsum);  // the digit's start in the tile
        // look-back: the digit's exclusive prefix over the earlier tiles,
        // reading kRsProbe predecessors per round and summing aggregates up to
        // the nearest inclusive prefix; a predecessor that has not published
        // yet is read again in the next round
        uint32_t pre = 0;
        if (owner && t > 0) {
            int64_t next = static_cast<int64_t>(t) - 1;
            for (;;) {
                uint32_t sw[kRsProbe];
#pragma unroll
                for (int k = 0; k < kRsProbe; ++k) {
                    const int64_t j = next - k;
                    sw[k] = j >= 0 ? ld_agent_u32(status + static_cast<uint64_t>(j) * kBins + tid) : kRsInc;
                }
                bool stop = false, done = false;
                int used = 0;
#pragma unroll
                for (int k = 0; k < kRsProbe; ++k) {
                    if (!stop) {
                        const uint32_t f = sw[k] & ~kRsVal;
                        if (f == 0) {
                            stop = true;
                        } else {
                            pre += sw[k] & kRsVal;
                            ++used;
                            if (f & kRsInc) {
                                stop = true;
                                done = true;
                            }
                        }
                    }
                }
                if (done) break;
                next -= used;
                if (used < kRsProbe) __builtin_amdgcn_s_sleep(1);
            }
            st_agent_u32(status + t * kBins + tid, kRsInc | (pre + v));
        }
        if (owner) {
            uint32_t p = ex;
#pragma unroll
            for (int w = 0; w < kRsWaves; ++w) {
                const uint32_t c = wc[w][tid];
                wc[w][tid] = p;
                p += c;
            }
            gofs[tid] = gst + pre - ex;
        }
        __syncthreads();
        // rank -> place in the tile
#pragma unroll
        for (int r = 0; r < kRsPer; ++r) rank[r] += mycnt[(key[r] >> shift) & (kBins - 1)];
        for (uint32_t h0 = 0; h0 < cnt; h0 += kRsHalf) {
#pragma unroll
            for (int r = 0; r < kRsPer; ++r) {
                const uint32_t e = e0 + static_cast<uint32_t>(r) * kWave;
                if (e < cnt && rank[r] - h0 < static_cast<uint32_t>(kRsHalf)) sk[rank[r] - h0] = key[r];
            }
            __syncthreads();
            const uint32_t nh = cnt - h0 < static_cast<uint32_t>(kRsHalf) ? cnt - h0 : kRsHalf;
            for (uint32_t i = tid; i < nh; i += kRsThreads) {
                const uint64_t k = sk[i];
                out[gofs[(k >> shift) & (kBins - 1)] + h0 + i] = k & outMask;
            }
            __syncthreads();
        }
    }
    if (NEXT)
        for (int d = tid; d < kBins; d += kRsThreads)
            if (side[d]) atomicAdd(&nextCounts[d], side[d]);
}

// off[0..n] = exclusive scan of cnt[0..n) (off[n] = total), one workgroup.
constexpr int kGsThreads = 1024;
__global__ __launch_bounds__(kGsThreads) void k_rs_group_scan(const uint64_t* __restrict__ cnt, uint64_t n,
                                                              uint64_t* __restrict__ off) {
    __shared__ uint64_t part[kGsThreads / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const uint64_t per = (n + kGsThreads - 1) / kGsThreads;
    const uint64_t a = static_cast<uint64_t>(tid) * per;
    const uint64_t b = a + per < n ? a + per : n;
    uint64_t sum = 0;
    for (uint64_t i = a; i < b; ++i) sum += cnt[i];
    uint64_t x = sum;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) part[wid] = x;
    __syncthreads();
    uint64_t p = x - sum, total = 0;
    for (int w = 0; w < kGsThreads / kWave; ++w) {
        if (w < wid) p += part[w];
        total += part[w];
    }
    for (uint64_t i = a; i < b; ++i) {
        const uint64_t c = cnt[i];
        off[i] = p;
        p += c;
    }
    if (tid == 0) off[n] = total;
}

// status[0] = hits, status[1] = overflow flags (1 pairs, 2 items, 4 hits)
__global__ void k_copy_status(const uint64_t* ctl, uint64_t* status) {
    if (threadIdx.x == 0) {
        status[0] = ctl[2];
        status[1] = ctl[3];
        status[2] = ctl[8];
        status[3] = ctl[9];
    }
}

int scan_launch_grid(tblaeq_ctx* ctx, uint64_t tiles) {
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * 4;
    return static_cast<int>(tiles < cap ? (tiles ? tiles : 1) : cap);
}

// Scratch words (u64) radix_sort needs for up to cap keys and `groups` groups.
uint64_t radix_hist_words(uint64_t cap, uint64_t groups) {
    const uint64_t tiles = (cap + kRsTile - 1) / kRsTile;
    return (kRsHdrWords + 2 * tiles * kRsMaxBins + 1) / 2 + groups;
}

template <int RB, bool NEXT>
static void launch_onesweep(tblaeq_ctx* ctx, const uint64_t* src, uint64_t* dst, const uint64_t* dn, uint64_t cap,
                            int shift, const uint32_t* counts, uint32_t* nextCounts, uint32_t* status,
                            uint32_t* statusNext, uint32_t* ticket, uint64_t outMask) {
    const uint64_t capTiles = (cap + kRsTile - 1) / kRsTile;
    const uint64_t scap = static_cast<uint64_t>(ctx->cuCount) *
                          resident_blocks(reinterpret_cast<const void*>(k_rs_onesweep<RB, NEXT>), kRsThreads);
    const int grid = static_cast<int>(capTiles < scap ? (capTiles ? capTiles : 1) : scap);
    hipLaunchKernelGGL((k_rs_onesweep<RB, NEXT>), dim3(grid), dim3(kRsThreads), 0, ctx->stream, src, dst, dn, cap,
                       shift, counts, nextCounts, status, statusNext, ticket, outMask);
}

template <int RB>
static int radix_passes(tblaeq_ctx* ctx, uint64_t* a, uint64_t* b, const uint64_t* dn, uint64_t cap, int firstBit,
                        int passes, uint64_t* scratch, uint64_t** result, const RadixOut& ro, uint64_t blocksPerCu) {
    constexpr int kBins = 1 << RB;
    const uint64_t capTiles = (cap + kRsTile - 1) / kRsTile;
    uint32_t* hdr = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* counts = hdr + kRsCounts0;
    uint32_t* region[2] = {hdr + kRsHdrWords, hdr + kRsHdrWords + capTiles * kRsMaxBins};
    uint64_t* groupCnt = ro.groupOff ? scratch + (kRsHdrWords + 2 * capTiles * kRsMaxBins + 1) / 2 : nullptr;
    const uint64_t gcap = static_cast<uint64_t>(ctx->cuCount) * blocksPerCu;
    const int gridC = static_cast<int>(capTiles < gcap ? (capTiles ? capTiles : 1) : gcap);
    TimedScope ts(ctx, kKBatch0 + 4);
    hipLaunchKernelGGL(k_rs_begin, dim3(16), dim3(256), 0, ctx->stream, hdr, groupCnt, ro.nGroups);
    hipLaunchKernelGGL(k_rs_counts<RB>, dim3(gridC), dim3(kRsThreads), 0, ctx->stream, a, dn, cap, firstBit, counts,
                       region[0], groupCnt, ro.groupShift, ro.nGroups);
    uint64_t* src = a;
    uint64_t* dst = b;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        uint64_t* o = (last && ro.ids) ? ro.ids : dst;
        uint32_t* cp = counts + static_cast<uint64_t>(p) * kBins;
        const int shift = firstBit + RB * p;
        if (!last)
            launch_onesweep<RB, true>(ctx, src, o, dn, cap, shift, cp, cp + kBins, region[p & 1],
                                      region[(p + 1) & 1], hdr + p, ~0ull);
        else
            launch_onesweep<RB, false>(ctx, src, o, dn, cap, shift, cp, nullptr, region[p & 1], nullptr, hdr + p,
                                       ro.idMask);
        src = o;
        dst = (src == a) ? b : a;
    }
    TB_HIP(hipGetLastError());
    if (ro.groupOff) {
        // group starts = exclusive scan of the group counts (total at [nGroups])
        if (ro.nGroups <= (1u << 20)) {
            hipLaunchKernelGGL(k_rs_group_scan, dim3(1), dim3(kGsThreads), 0, ctx->stream, groupCnt, ro.nGroups,
                               ro.groupOff);
            TB_HIP(hipGetLastError());
        } else {
            int rc = scan_counts(ctx, groupCnt, ro.nGroups, ro.groupOff, reinterpret_cast<uint64_t*>(hdr + 32));
            if (rc) return rc;
        }
    }
    *result = src;
    return TBLAEQ_OK;
}

// Stable LSD radix sort of `dn` keys (device count, at most cap < 2^30) on bits
// [firstBit, firstBit + nbits): 9-bit digits when they take fewer passes than
// 8-bit ones, else 8-bit.  Ping-pongs between a and b; *result = the buffer
// holding the result (ro.ids when given).
int radix_sort(tblaeq_ctx* ctx, uint64_t* a, uint64_t* b, const uint64_t* dn, uint64_t cap, int firstBit, int nbits,
               uint64_t* hist, uint64_t** result, const RadixOut& ro, uint64_t blocksPerCu) {
    if (cap > kRsVal) {
        set_error("radix sort: more than 2^30 - 1 keys");
        return TBLAEQ_ERR_CAPACITY;
    }
    if (nbits < 1 || firstBit < 0 || firstBit + nbits > 64) {
        set_error("radix sort: key bits out of range");
        return TBLAEQ_ERR_INVALID;
    }
    const int rb = (nbits + 8) / 9 < (nbits + 7) / 8 ? 9 : 8;
    const int passes = (nbits + rb - 1) / rb;
    if (rb == 9) return radix_passes<9>(ctx, a, b, dn, cap, firstBit, passes, hist, result, ro, blocksPerCu);
    return radix_passes<8>(ctx, a, b, dn, cap, firstBit, passes, hist, result, ro, blocksPerCu);
}

int resident_blocks(const void* kernel, int threads) {
    static std::mutex mu;
    static std::unordered_map<const void*, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    cache.emplace(kernel, n);
    return n;
}

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

template <int MODE, int STAGE>
auto range_kernel(int D) -> void (*)(RangeStepArgs) {
    if constexpr (MODE == kFinal) {
        if (D == 4) return k_range_final<4, true, STAGE>;
        if (D == 8) return k_range_final<8, true, STAGE>;
        if (D == 16) return k_range_final<16, true, STAGE>;
        if (D < 4) return k_range_final<4, false, STAGE>;
        if (D < 8) return k_range_final<8, false, STAGE>;
        if (D < 16) return k_range_final<16, false, STAGE>;
        return k_range_final<32, false, STAGE>;
    } else {
        if (D == 4) return k_range_step<4, MODE, true>;
        if (D == 8) return k_range_step<8, MODE, true>;
        if (D == 16) return k_range_step<16, MODE, true>;
        if (D < 4) return k_range_step<4, MODE, false>;
        if (D < 8) return k_range_step<8, MODE, false>;
        if (D < 16) return k_range_step<16, MODE, false>;
        return k_range_step<32, MODE, false>;
    }
}

template <int STAGE>
auto rows_kernel(int D) -> void (*)(RangeStepArgs) {
    if (D == 4) return k_range_rows<4, true, STAGE>;
    if (D == 8) return k_range_rows<8, true, STAGE>;
    if (D == 16) return k_range_rows<16, true, STAGE>;
    if (D < 4) return k_range_rows<4, false, STAGE>;
    if (D < 8) return k_range_rows<8, false, STAGE>;
    if (D < 16) return k_range_rows<16, false, STAGE>;
    return k_range_rows<32, false, STAGE>;
}

template <int MODE, int STAGE>
void launch_range_dm(tblaeq_ctx* ctx, const RangeStepArgs& A, dim3 g, dim3 b) {
    void (*k)(RangeStepArgs) = range_kernel<MODE, STAGE>(A.D);
    if (MODE == kFinal) {  // work-queue kernel: resident blocks only
        const uint64_t cap =
            static_cast<uint64_t>(ctx->cuCount) *
            static_cast<uint64_t>(resident_blocks(reinterpret_cast<const void*>(k), static_cast<int>(b.x)));
        if (g.x > cap) g.x = static_cast<unsigned>(cap);
    }
    hipLaunchKernelGGL(k, g, b, 0, ctx->stream, A);
}

// Queries the occupancy of this engine's final-filter and radix kernels (outside any capture).
void warm_range_kernels(tblaeq_batch* b) {
    const int t = kStepWaves * kWave;
    resident_blocks(reinterpret_cast<const void*>(range_kernel<kFinal, 256>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(range_kernel<kFinal, 512>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(range_kernel<kFinal, 1024>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(rows_kernel<256>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(rows_kernel<512>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(rows_kernel<1024>(b->D)), t);
    resident_blocks(reinterpret_cast<const void*>(k_rs_onesweep<8, true>), kRsThreads);
    resident_blocks(reinterpret_cast<const void*>(k_rs_onesweep<8, false>), kRsThreads);
    resident_blocks(reinterpret_cast<const void*>(k_rs_onesweep<9, true>), kRsThreads);
    resident_blocks(reinterpret_cast<const void*>(k_rs_onesweep<9, false>), kRsThreads);
}

template <int MODE>
int launch_range_step(tblaeq_batch* bt, const RangeStepArgs& A, uint64_t itemBound) {
    tblaeq_ctx* ctx = bt->ctx;
    uint64_t blocks = (itemBound + kStepWaves - 1) / kStepWaves;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * bt->rangeBlocksPerCu;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    const dim3 g(static_cast<unsigned>(blocks)), b(kStepWaves * kWave);
    TimedScope ts(ctx, MODE == kFinal ? kKBatch0 + 1 : kKBatch0 + 0);
    if (MODE != kFinal) launch_range_dm<MODE, 1>(ctx, A, g, b);
    else if (bt->hitStage == 256) launch_range_dm<MODE, 256>(ctx, A, g, b);
    else if (bt->hitStage == 1024) launch_range_dm<MODE, 1024>(ctx, A, g, b);
    else launch_range_dm<MODE, 512>(ctx, A, g, b);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

template <int STAGE>
static void launch_rows_dm(tblaeq_ctx* ctx, const RangeStepArgs& A, dim3 g, dim3 b) {
    void (*k)(RangeStepArgs) = rows_kernel<STAGE>(A.D);
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) *
                         static_cast<uint64_t>(resident_blocks(reinterpret_cast<const void*>(k), static_cast<int>(b.x)));
    if (g.x > cap) g.x = static_cast<unsigned>(cap);
    hipLaunchKernelGGL(k, g, b, 0, ctx->stream, A);
}

// Row-major exact filter of the clusters kept by at most A.rowMax queries
// (items from queue slot 1, zeroed with the control block).
static int launch_range_rows(tblaeq_batch* bt, const RangeStepArgs& A, uint64_t itemBound) {
    tblaeq_ctx* ctx = bt->ctx;
    uint64_t blocks = (itemBound + kStepWaves - 1) / kStepWaves;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * bt->rangeBlocksPerCu;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    const dim3 g(static_cast<unsigned>(blocks)), b(kStepWaves * kWave);
    TimedScope ts(ctx, kKBatch0 + 12);
    if (bt->hitStage == 256) launch_rows_dm<256>(ctx, A, g, b);
    else if (bt->hitStage == 1024) launch_rows_dm<1024>(ctx, A, g, b);
    else launch_rows_dm<512>(ctx, A, g, b);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

// Bounding box of each last-level position's points (its CSC column at the
// finest level), rounded outward to float: lo[D] | hi[D] per position.  An
// empty cluster gets an empty box (+inf | -inf).
__global__ void k_aabb(const uint64_t* __restrict__ ids, uint64_t n, const uint64_t* __restrict__ colRes,
                       const double* __restrict__ coords, int D, float* __restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t w0 = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
    for (uint64_t j = w0; j < n; j += nw) {
        const uint64_t c = ids[j];
        const uint64_t cb = colRes[c], ce = colRes[c + 1];
        for (int d = 0; d < D; ++d) {
            double mn = __builtin_inf(), mx = -__builtin_inf();
            for (uint64_t e = cb + lane; e < ce; e += kWave) {
                const double v = coords[e * D + d];
                mn = fmin(mn, v);
                mx = fmax(mx, v);
            }
            for (int o = 32; o > 0; o >>= 1) {
                mn = fmin(mn, __shfl_xor(mn, o, kWave));
                mx = fmax(mx, __shfl_xor(mx, o, kWave));
            }
            if (lane == 0) {
                out[j * 2 * D + d] = __double2float_rd(mn);
                out[j * 2 * D + D + d] = __double2float_ru(mx);
            }
        }
    }
}

// Ball and box around each cluster (position j of a level, id ids[j], centre
// coords[j]) that hold the balls of all its children (positions
// colRes[id]..colRes[id+1] of the next level, centres childCoords, radii
// childRadius): R >= |centre - child centre| + child radius for every child,
// rounded up by a relative 1e-12; box lo[D] | hi[D] = the extent of the child
// balls, rounded outward to float.  Written by id; a cluster without children
// gets R = 0 and an empty box (+inf | -inf).
__global__ void k_enclose(const uint64_t* __restrict__ ids, uint64_t n, const uint64_t* __restrict__ colRes,
                          const double* __restrict__ coords, const double* __restrict__ childCoords,
                          const double* __restrict__ childRadius, int D, double* __restrict__ cOut,
                          double* __restrict__ rOut, float* __restrict__ boxOut) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t w0 = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
    for (uint64_t j = w0; j < n; j += nw) {
        const uint64_t c = ids[j];
        const double* ctr = coords + j * D;
        const uint64_t cb = colRes[c], ce = colRes[c + 1];
        double r = 0.0;
        for (uint64_t e = cb + lane; e < ce; e += kWave) {
            const double* x = childCoords + e * D;
            double d2 = 0.0;
            for (int d = 0; d < D; ++d) {
                const double diff = x[d] - ctr[d];
                d2 = __builtin_fma(diff, diff, d2);
            }
            r = fmax(r, __builtin_sqrt(d2) + childRadius[e]);
        }
        r = wave_max_f64(r);
        if (lane == 0) rOut[c] = r + r * 1e-12;
        for (int d = lane; d < D; d += kWave) cOut[c * D + d] = ctr[d];
        for (int d = 0; d < D; ++d) {
            double mn = __builtin_inf(), mx = -__builtin_inf();
            for (uint64_t e = cb + lane; e < ce; e += kWave) {
                const double v = childCoords[e * D + d], cr = childRadius[e];
                mn = fmin(mn, v - cr);
                mx = fmax(mx, v + cr);
            }
            for (int o = 32; o > 0; o >>= 1) {
                mn = fmin(mn, __shfl_xor(mn, o, kWave));
                mx = fmax(mx, __shfl_xor(mx, o, kWave));
            }
            if (lane == 0) {
                boxOut[c * 2 * D + d] = __double2float_rd(mn);
                boxOut[c * 2 * D + D + d] = __double2float_ru(mx);
            }
        }
    }
}

// Outward-rounded float box of every kd leaf: leaf k of cluster c (by id) is
// rows colRes[c] + k*kLeafRows .. (at most kLeafRows, within the column) of the
// finest level; one thread per cluster.
__global__ void k_leaf_box(const uint64_t* __restrict__ colRes, uint64_t cols, const uint64_t* __restrict__ leafOff,
                           const double* __restrict__ coords, int D, float* __restrict__ out) {
    for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < cols;
         c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = colRes[c], e = colRes[c + 1];
        uint64_t leaf = leafOff[c];
        for (uint64_t r0 = b; r0 < e; r0 += kLeafRows, ++leaf) {
            const uint64_t r1 = r0 + kLeafRows < e ? r0 + kLeafRows : e;
            for (int d = 0; d < D; ++d) {
                double mn = __builtin_inf(), mx = -__builtin_inf();
                for (uint64_t r = r0; r < r1; ++r) {
                    const double v = coords[r * D + d];
                    mn = fmin(mn, v);
                    mx = fmax(mx, v);
                }
                out[leaf * 2 * D + d] = __double2float_rd(mn);
                out[leaf * 2 * D + D + d] = __double2float_ru(mx);
            }
        }
    }
}

// kd order of the rows of every column (rows colRes[c]..colRes[c+1] of the
// point array x, D doubles each): a node of more than kLeafRows rows splits
// along the dimension of its widest extent at a multiple of kLeafRows near its
// middle (nth_element), recursively, so that each run of kLeafRows consecutive
// rows is a spatially tight leaf.  perm[i] = the original row now at i.
static void kd_leaf_order(const std::vector<double>& x, int D, const std::vector<uint64_t>& colRes,
                          std::vector<uint64_t>& perm) {
    const uint64_t cols = colRes.size() - 1;
    perm.resize(colRes.back());
    for (uint64_t i = 0; i < perm.size(); ++i) perm[i] = i;
    auto split = [&](auto& self, uint64_t b, uint64_t e) -> void {
        const uint64_t n = e - b;
        if (n <= static_cast<uint64_t>(kLeafRows)) return;
        int best = 0;
        double wbest = -1.0;
        for (int d = 0; d < D; ++d) {
            double mn = x[perm[b] * D + d], mx = mn;
            for (uint64_t i = b + 1; i < e; ++i) {
                const double v = x[perm[i] * D + d];
                mn = std::min(mn, v);
                mx = std::max(mx, v);
            }
            if (mx - mn > wbest) {
                wbest = mx - mn;
                best = d;
            }
        }
        const uint64_t leaves = (n + kLeafRows - 1) / kLeafRows;
        const uint64_t m = b + (leaves / 2) * kLeafRows;
        std::nth_element(perm.begin() + static_cast<long>(b), perm.begin() + static_cast<long>(m),
                         perm.begin() + static_cast<long>(e), [&](uint64_t u, uint64_t v) {
                             const double a = x[u * D + best], c = x[v * D + best];
                             return a < c || (a == c && u < v);
                         });
        self(self, b, m);
        self(self, m, e);
    };
    const unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::atomic<uint64_t> next{0};
    for (unsigned t = 0; t < hw; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const uint64_t c0 = next.fetch_add(1024);
                if (c0 >= cols) return;
                const uint64_t c1 = std::min<uint64_t>(cols, c0 + 1024);
                for (uint64_t c = c0; c < c1; ++c) split(split, colRes[c], colRes[c + 1]);
            }
        });
    for (auto& t : th) t.join();
}

int launch_grouping(tblaeq_batch* b, uint64_t mesh) {
    tblaeq_ctx* ctx = b->ctx;
    uint64_t* ctl = b->ctl;
    const uint64_t tiles = (mesh + kScanThreads - 1) / kScanThreads;
    int rc = scan_prepare(ctx, tiles);
    if (rc) return rc;
    {
        TimedScope ts(ctx, kKBatch0 + 2);
        hipLaunchKernelGGL(k_group_scan, dim3(scan_launch_grid(ctx, tiles)), dim3(kScanThreads), 0, ctx->stream,
                           b->cnt, mesh, b->off, b->items, b->itemCap, ctl + 1, ctl + 3, ctx->status, ctx->ctrl, 0ull,
                           SparseList{});
    }
    {
        TimedScope ts(ctx, kKBatch0 + 3);
        const uint64_t g = std::min<uint64_t>((b->pairCap + kAggChunk - 1) / kAggChunk,
                                              static_cast<uint64_t>(ctx->cuCount) * 4);
        hipLaunchKernelGGL(k_group_scatter, dim3(static_cast<unsigned>(g ? g : 1)), dim3(kAggThreads), 0, ctx->stream,
                           b->pairs, ctl + 0, b->pairCap, mesh, b->cnt, b->off, b->qlist);
    }
    // the counts also cover pairs that overflowed the pair buffer (the batch is
    // then rerun): clear what the scatter did not count down
    launch_fill(b, b->cnt, 0, mesh);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->stream, ctl, 0ull, 1);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

SparseList sparse_list(tblaeq_batch* b, uint64_t n) {
    SparseList l;
    const uint64_t s = sparse_slots(n);
    l.slots = b->sparse;
    l.counts = b->sparse + s;
    l.units = b->sparse + s + s / kScanThreads;
    l.unitCount = b->queue + kSparseUnitQueue;
    return l;
}

int launch_group_scan(tblaeq_batch* b, uint64_t mesh, uint64_t* off, uint64_t* items, uint64_t* nItems,
                      uint64_t rowMax) {
    tblaeq_ctx* ctx = b->ctx;
    const uint64_t tiles = (mesh + kScanThreads - 1) / kScanThreads;
    int rc = scan_prepare(ctx, tiles);
    if (rc) return rc;
    TimedScope ts(ctx, kKBatch0 + 2);
    hipLaunchKernelGGL(k_group_scan, dim3(scan_launch_grid(ctx, tiles)), dim3(kScanThreads), 0, ctx->stream, b->cnt,
                       mesh, off, items, b->itemCap, nItems, b->ctl + 3, ctx->status, ctx->ctrl, rowMax,
                       rowMax ? sparse_list(b, mesh) : SparseList{});
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

int scan_counts(tblaeq_ctx* ctx, const uint64_t* cnt, uint64_t n, uint64_t* off, uint64_t* scratch2) {
    const uint64_t tiles = (n + kScanThreads - 1) / kScanThreads;
    int rc = scan_prepare(ctx, tiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_group_scan, dim3(scan_launch_grid(ctx, tiles)), dim3(kScanThreads), 0, ctx->stream, cnt, n,
                       off, nullptr, 0ull, scratch2, scratch2 + 1, ctx->status, ctx->ctrl, 0ull, SparseList{});
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

int launch_count_scan(tblaeq_batch* b, const uint64_t* cnt, uint64_t n, uint64_t* off) {
    tblaeq_ctx* ctx = b->ctx;
    const uint64_t tiles = (n + kScanThreads - 1) / kScanThreads;
    int rc = scan_prepare(ctx, tiles);
    if (rc) return rc;
    TimedScope ts(ctx, kKBatch0 + 2);
    hipLaunchKernelGGL(k_group_scan, dim3(scan_launch_grid(ctx, tiles)), dim3(kScanThreads), 0, ctx->stream, cnt, n,
                       off, nullptr, 0ull, b->ctl + 6, b->ctl + 7, ctx->status, ctx->ctrl, 0ull, SparseList{});
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

__global__ void k_narrow_ids(const uint64_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[i] = static_cast<uint32_t>(in[i]);
}

hipError_t narrow_ids(hipStream_t st, const uint64_t* in, uint32_t* out, uint64_t n) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const uint64_t g = std::min<uint64_t>((n + 255) / 256, static_cast<uint64_t>(cus) * 8);
    hipLaunchKernelGGL(k_narrow_ids, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

void launch_copy_status(tblaeq_batch* b, uint64_t* status) {
    hipLaunchKernelGGL(k_copy_status, dim3(1), dim3(64), 0, b->ctx->stream, b->ctl, status);
}

void launch_fill(tblaeq_batch* b, uint64_t* p, uint64_t v, uint64_t n) {
    const uint64_t g = std::min<uint64_t>((n + 255) / 256, static_cast<uint64_t>(b->ctx->cuCount) * 8);
    hipLaunchKernelGGL(k_fill_u64, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, b->ctx->stream, p, v, n);
}

// Frees one of the engine's own allocations.
void bfree(tblaeq_batch* b, void* p) {
    auto it = std::find(b->owned.begin(), b->owned.end(), p);
    if (it == b->owned.end()) return;
    (void)hipFree(p);
    b->owned.erase(it);
}

template <class T>
T* balloc(tblaeq_batch* b, size_t n, int* rc) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        *rc = hip_fail(e, "hipMalloc(batch)");
        return nullptr;
    }
    b->owned.push_back(p);
    return static_cast<T*>(p);
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_batch_create(tblaeq_ctx* ctx, const tblaeq_level* levels, uint64_t nlevels,
                                   const uint64_t* coarsest_ids, const double* coarsest_vals, uint64_t ncoarsest,
                                   const double* host_max_radius, tblaeq_batch** out) {
    if (!ctx || !levels || nlevels == 0 || !coarsest_ids || !coarsest_vals || !out || ncoarsest == 0) {
        set_error("tblaeq_batch_create: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t D = levels[0].D;
    if (D == 0 || D > 32) {
        set_error("tblaeq_batch_create: the batched path handles 1 <= D <= 32");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    for (uint64_t l = 0; l < nlevels; ++l)
        if (!levels[l].pvals || !levels[l].radius || !levels[l].col_res || levels[l].D != D) {
            set_error("tblaeq_batch_create: every level needs resident P values, radii and column pointers");
            return TBLAEQ_ERR_INVALID;
        }
    auto* b = new tblaeq_batch();
    b->ctx = ctx;
    b->D = static_cast<int>(D);
    b->L = static_cast<int>(nlevels);
    const int L = b->L;
    b->count.resize(L + 1);
    b->coords.resize(L + 1);
    b->ids.resize(L + 1);
    b->radiusPos.resize(L, nullptr);
    b->parentPos.resize(L, nullptr);
    b->weightPos.resize(L, nullptr);
    b->colRes.resize(L);
    b->nnzPerCol.resize(L);
    b->radiusById.resize(L);
    b->emptyCols.resize(L);
    b->maxRadius.assign(host_max_radius ? host_max_radius : nullptr, host_max_radius ? host_max_radius + L : nullptr);
    int rc = TBLAEQ_OK;
    auto fail = [&](int code) {
        tblaeq_batch_destroy(b);
        return code;
    };
    b->count[0] = ncoarsest;
    b->coords[0] = const_cast<double*>(coarsest_vals);
    b->ids[0] = const_cast<uint64_t*>(coarsest_ids);
    for (int l = 0; l < L; ++l) {
        b->colRes[l] = levels[l].col_res;
        b->nnzPerCol[l] = levels[l].nnz_per_col;
        b->radiusById[l] = levels[l].radius;
        b->emptyCols[l] = levels[l].empty_cols;
        b->count[l + 1] = levels[l].nnz;
        b->maxMesh = std::max<uint64_t>(b->maxMesh, levels[l].cols);
    }
    // Whole-hierarchy prolongation: level l+1 positions from all level-l positions.
    uint64_t* dn = balloc<uint64_t>(b, 1, &rc);
    if (rc) return fail(rc);
    for (int l = 0; l < L; ++l) {
        const uint64_t n = b->count[l];
        uint64_t* offp = balloc<uint64_t>(b, n, &rc);
        if (rc) return fail(rc);
        b->coords[l + 1] = balloc<double>(b, b->count[l + 1] * D, &rc);
        if (rc) return fail(rc);
        b->ids[l + 1] = balloc<uint64_t>(b, b->count[l + 1], &rc);
        if (rc) return fail(rc);
        b->radiusPos[l] = balloc<double>(b, n, &rc);
        if (rc) return fail(rc);
        b->weightPos[l] = balloc<uint64_t>(b, n, &rc);
        if (rc) return fail(rc);
        const unsigned g = static_cast<unsigned>((n + 255) / 256);
        hipLaunchKernelGGL(k_gather_off, dim3(g), dim3(256), 0, ctx->stream, b->ids[l], n, levels[l].col_res, offp);
        hipLaunchKernelGGL(k_gather_f64, dim3(g), dim3(256), 0, ctx->stream, b->ids[l], n, levels[l].radius,
                           b->radiusPos[l]);
        hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, ctx->stream, b->ids[l], n, levels[l].nnz_per_col,
                           b->weightPos[l]);
        if (hipMemcpyAsync(dn, &n, 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return fail(TBLAEQ_ERR_HIP);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(TBLAEQ_ERR_HIP);
        rc = tblaeq_prolong_refactor(ctx, &levels[l], b->ids[l], b->coords[l], offp, dn, n, b->ids[l + 1],
                                     b->coords[l + 1]);
        if (rc) return fail(rc);
        if (l >= 1) {
            b->parentPos[l] = balloc<uint64_t>(b, n, &rc);
            if (rc) return fail(rc);
            hipLaunchKernelGGL(k_fill_parent, dim3(static_cast<unsigned>((levels[l - 1].cols + 255) / 256)), dim3(256),
                               0, ctx->stream, levels[l - 1].col_res, levels[l - 1].cols, b->parentPos[l]);
        }
    }
    {
        // kd order of every last-level cluster's rows and the leaves' boxes
        const uint64_t cols = levels[L - 1].cols, nf = b->count[L];
        std::vector<uint64_t> hcr(cols + 1), perm;
        std::vector<double> hx(nf * D);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
            hipMemcpy(hcr.data(), levels[L - 1].col_res, (cols + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hx.data(), b->coords[L], nf * D * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "batch create: kd order copy"));
        kd_leaf_order(hx, static_cast<int>(D), hcr, perm);
        std::vector<uint64_t> hlo(cols + 1, 0);
        for (uint64_t c = 0; c < cols; ++c) {
            const uint64_t nl = (hcr[c + 1] - hcr[c] + kLeafRows - 1) / kLeafRows;
            hlo[c + 1] = hlo[c] + nl;
            b->maxLeaves = std::max(b->maxLeaves, nl);
        }
        b->nLeaves = hlo[cols];
        uint64_t* dperm = balloc<uint64_t>(b, nf, &rc);
        if (rc) return fail(rc);
        // kFilterPadRows spare rows after the last one: the exact filter loads
        // rows in pairs ahead of their tests and may read up to three rows
        // past a leaf's end
        double* xk = balloc<double>(b, (nf + kFilterPadRows) * D, &rc);
        if (rc) return fail(rc);
        uint64_t* ik = balloc<uint64_t>(b, nf + kFilterPadRows, &rc);
        if (rc) return fail(rc);
        TB_HIP(hipMemsetAsync(xk + nf * D, 0, kFilterPadRows * D * 8, ctx->stream));
        TB_HIP(hipMemsetAsync(ik + nf, 0, kFilterPadRows * 8, ctx->stream));
        b->leafOff = balloc<uint64_t>(b, cols + 1, &rc);
        if (rc) return fail(rc);
        b->leafBox = balloc<float>(b, b->nLeaves * 2 * D, &rc);
        if (rc) return fail(rc);
        {
            const uint64_t m = std::max<uint64_t>(cols, b->count[L - 1]);
            b->sparse = balloc<uint64_t>(b, sparse_words(m), &rc);
        }
        if (rc) return fail(rc);
        if (hipMemcpy(dperm, perm.data(), nf * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->leafOff, hlo.data(), (cols + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "batch create: kd order upload"));
        move_rows(ctx, false, b->coords[L], dperm, nf, D, xk);
        move_rows(ctx, false, b->ids[L], dperm, nf, 1, ik);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "kd gather"));
        bfree(b, b->coords[L]);
        bfree(b, b->ids[L]);
        bfree(b, dperm);
        b->coords[L] = xk;
        b->ids[L] = ik;
        const uint64_t g = std::min<uint64_t>((cols + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 16);
        hipLaunchKernelGGL(k_leaf_box, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream,
                           levels[L - 1].col_res, cols, b->leafOff, b->coords[L], static_cast<int>(D), b->leafBox);
    }
    {
        const uint64_t n = b->count[L - 1];
        b->aabb = balloc<float>(b, n * 2 * D, &rc);
        if (rc) return fail(rc);
        const uint64_t g = std::min<uint64_t>((n * kWave + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 16);
        hipLaunchKernelGGL(k_aabb, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream, b->ids[L - 1], n,
                           levels[L - 1].col_res, b->coords[L], static_cast<int>(D), b->aabb);
    }
    if (L >= 2) {
        const uint64_t n = b->count[L - 2], cols = std::max<uint64_t>(n, levels[L - 2].cols);
        b->enclC = balloc<double>(b, cols * D, &rc);
        if (rc) return fail(rc);
        b->enclR = balloc<double>(b, cols, &rc);
        if (rc) return fail(rc);
        b->enclBox = balloc<float>(b, cols * 2 * D, &rc);
        if (rc) return fail(rc);
        const uint64_t g = std::min<uint64_t>((n * kWave + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 16);
        hipLaunchKernelGGL(k_enclose, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream,
                           b->ids[L - 2], n, levels[L - 2].col_res, b->coords[L - 2], b->coords[L - 1],
                           b->radiusPos[L - 1], static_cast<int>(D), b->enclC, b->enclR, b->enclBox);
    }
    b->ctl = balloc<uint64_t>(b, kCtlWords + kQueueSlots, &rc);
    if (rc) return fail(rc);
    b->queue = b->ctl + kCtlWords;
    launch_fill(b, b->ctl, 0, kCtlWords + kQueueSlots);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "batch create"));
    *out = b;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_destroy(tblaeq_batch* b) {
    if (!b) return TBLAEQ_OK;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    if (b->rangeExec) (void)hipGraphExecDestroy(b->rangeExec);
    if (b->rangeGraph) (void)hipGraphDestroy(b->rangeGraph);
    if (b->knnExec) (void)hipGraphExecDestroy(b->knnExec);
    if (b->knnGraph) (void)hipGraphDestroy(b->knnGraph);
    for (void* p : b->owned) (void)hipFree(p);
    delete b;
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_reserve(tblaeq_batch* b, uint64_t max_queries, uint64_t pair_cap, uint64_t hit_cap) {
    if (!b || max_queries == 0 || max_queries >= (1ull << kQueryBits)) {
        set_error("tblaeq_batch_reserve: invalid argument (queries per batch must be below 2^24)");
        return TBLAEQ_ERR_INVALID;
    }
    int rc = TBLAEQ_OK;
    if (pair_cap > b->pairCap) {
        b->pairs = balloc<uint64_t>(b, pair_cap, &rc);
        if (rc) return rc;
        b->qlist = balloc<uint32_t>(b, pair_cap, &rc);
        if (rc) return rc;
        b->qlist2 = balloc<uint32_t>(b, pair_cap, &rc);
        if (rc) return rc;
        b->itemCap = pair_cap / kWave + b->maxMesh + 1;
        b->items = balloc<uint64_t>(b, b->itemCap, &rc);
        if (rc) return rc;
        b->items2 = balloc<uint64_t>(b, b->itemCap, &rc);
        if (rc) return rc;
        b->maskOff = balloc<uint64_t>(b, b->itemCap + 1, &rc);
        if (rc) return rc;
        b->pairCap = pair_cap;
    }
    if (!b->cnt) {
        b->cnt = balloc<uint64_t>(b, b->maxMesh, &rc);
        if (rc) return rc;
        b->off = balloc<uint64_t>(b, b->maxMesh + 1, &rc);
        if (rc) return rc;
        b->off2 = balloc<uint64_t>(b, b->maxMesh + 1, &rc);
        if (rc) return rc;
        TB_HIP(hipMemsetAsync(b->cnt, 0, b->maxMesh * 8, b->ctx->stream));
    }
    if (hit_cap > b->hitCap) {
        b->hits = balloc<uint64_t>(b, hit_cap, &rc);
        if (rc) return rc;
        b->hitsTmp = balloc<uint64_t>(b, hit_cap, &rc);
        if (rc) return rc;
        b->recAux = balloc<uint64_t>(b, hit_cap, &rc);
        if (rc) return rc;
        b->hitCap = hit_cap;
    }
    if (max_queries > b->maxQ || !b->fbox) {
        b->fbox = balloc<float>(b, std::max(max_queries, b->maxQ) * 2 * dm_of(b->D), &rc);
        if (rc) return rc;
    }
    b->maxQ = std::max(b->maxQ, max_queries);
    if (radix_hist_words(b->hitCap, b->maxQ) > b->histCap) {
        b->histCap = radix_hist_words(b->hitCap, b->maxQ);
        b->hist = balloc<uint64_t>(b, b->histCap, &rc);
        if (rc) return rc;
    }
    // size the look-back status array for every scan a batch runs, so that a
    // captured batch never allocates
    const uint64_t histTiles = (b->histCap + kScanThreads - 1) / kScanThreads;
    const uint64_t meshTiles = (b->maxMesh + kScanThreads - 1) / kScanThreads;
    if ((rc = scan_prepare(b->ctx, std::max(histTiles, meshTiles)))) return rc;
    TB_HIP(hipStreamSynchronize(b->ctx->stream));
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_range(tblaeq_batch* b, const double* d_boxes, uint64_t nq, uint64_t* d_out_off,
                                  uint64_t* d_out_ids, uint64_t out_cap, uint64_t* d_status,
                                  uint64_t* d_level_surv, uint64_t* d_fine_count) {
    if (!b || !d_boxes || nq == 0 || !d_out_off || !d_out_ids || !d_status || nq > b->maxQ || !b->pairs ||
        !b->hits) {
        set_error("tblaeq_batch_range: invalid argument or engine not reserved");
        return TBLAEQ_ERR_INVALID;
    }
    tblaeq_ctx* ctx = b->ctx;
    const int L = b->L;
    if (kMaskQueue0 + L > kQueueSlots) {
        set_error("tblaeq_batch_range: too many levels");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    uint64_t* ctl = b->ctl;
    launch_fill(b, ctl, 0, kCtlWords + kQueueSlots);
    RangeStepArgs A{};
    A.D = b->D;
    A.nq = nq;
    A.boxes = d_boxes;
    A.cnt = b->cnt;
    A.overflow = ctl + 3;
    A.L = L;
    A.itemCap = b->itemCap;
    A.qcap = b->pairCap;
    // two group sets: the prune of level l reads one and writes the other
    uint64_t* offS[2] = {b->off, b->off2};
    uint64_t* itemS[2] = {b->items, b->items2};
    uint32_t* qS[2] = {b->qlist, b->qlist2};
    uint64_t* nS[2] = {ctl + 1, ctl + 4};
    const uint64_t rowMax = std::min<uint64_t>(b->rangeRowMax, static_cast<uint64_t>(rows_max_queries(dm_of(b->D))));
    int cur = 0;
    int rc;
    for (int l = 0; l < L; ++l) {
        const int nxt = cur ^ 1;
        A.level0 = l == 0;
        A.level = l;
        A.items = itemS[cur];
        A.nItems = nS[cur];
        A.off = offS[cur];
        A.qlist = qS[cur];
        A.colResPrev = l > 0 ? b->colRes[l - 1] : nullptr;
        A.count0 = b->count[0];
        A.coords = b->coords[l];
        A.radiusPos = b->radiusPos[l];
        A.ids = b->ids[l];
        A.aabb = l == L - 1 ? b->aabb : nullptr;
        A.ballStats = (d_level_surv || d_fine_count) ? 1 : 0;
        A.weightPos = (l == L - 1 && d_fine_count) ? b->weightPos[l] : nullptr;
        A.offNext = offS[nxt];
        A.qlistNext = qS[nxt];
        A.cap = b->pairCap;
        A.survStat = d_level_surv;
        A.fineStat = d_fine_count;
        const uint64_t bound = l == 0 ? (nq + kWave - 1) / kWave : b->itemCap;
        // keep masks of the count pass, reused by the write pass
        A.masks = b->hits;
        A.maskPos = b->hitsTmp;
        A.maskCap = b->hitCap;
        A.maskOff = b->maskOff;
        A.maskNext = b->queue + kMaskQueue0 + l;
        if ((rc = launch_range_step<kCount>(b, A, bound))) return rc;
        // group the kept (cluster, query) pairs of level l by cluster
        const uint64_t mesh = b->count[l];
        const uint64_t tiles = (mesh + kScanThreads - 1) / kScanThreads;
        if ((rc = scan_prepare(ctx, tiles))) return rc;
        {
            TimedScope ts(ctx, kKBatch0 + 2);
            // the last prune's groups feed the exact filter: clusters kept by at
            // most rowMax queries go to the row-major filter's list instead
            const bool split = l == L - 1 && rowMax > 0;
            hipLaunchKernelGGL(k_group_scan, dim3(scan_launch_grid(ctx, tiles)), dim3(kScanThreads), 0, ctx->stream,
                               b->cnt, mesh, offS[nxt], itemS[nxt], b->itemCap, nS[nxt], ctl + 3, ctx->status,
                               ctx->ctrl, split ? rowMax : 0ull, split ? sparse_list(b, mesh) : SparseList{});
        }
        A.survStat = nullptr;
        A.fineStat = nullptr;
        if ((rc = launch_range_step<kWrite>(b, A, bound))) return rc;
        // the writes count the pairs back to zero only when they take their
        // slots by atomics
        launch_fill(b, b->cnt, 0, mesh);
        cur = nxt;
    }
    // final step: exact filter of the finest candidates
    const uint64_t maxId = b->count[L];
    const int idBits = bits_for(maxId - 1);  // ids < count[L]
    const int qBits = bits_for(nq - 1);      // queries < nq
    if (idBits + qBits > 64) {
        set_error("tblaeq_batch_range: id and query bits exceed 64");
        return TBLAEQ_ERR_UNSUPPORTED;
    }
    A.level0 = 0;
    A.level = L;
    A.items = itemS[cur];
    A.nItems = nS[cur];
    A.off = offS[cur];
    A.qlist = qS[cur];
    A.colResPrev = b->colRes[L - 1];
    A.coords = b->coords[L];
    A.radiusPos = nullptr;
    A.aabb = nullptr;
    A.weightPos = nullptr;
    A.ids = b->ids[L];
    A.out = b->hits;
    A.cap = std::min(b->hitCap, out_cap);
    A.outCount = ctl + 2;
    A.idBits = idBits;
    A.work = ctl + 8;
    A.leafOff = b->leafOff;
    A.leafBox = b->leafBox;
    A.fboxes = b->fbox;
    {
        const uint64_t n = nq * 2 * dm_of(b->D);
        const uint64_t g = std::min<uint64_t>((n + 255) / 256, static_cast<uint64_t>(ctx->cuCount) * 4);
        hipLaunchKernelGGL(k_box_f32, dim3(static_cast<unsigned>(g ? g : 1)), dim3(256), 0, ctx->stream, d_boxes, nq,
                           b->D, dm_of(b->D), b->fbox);
    }
    A.queue = b->queue;
    A.masks = nullptr;
    A.rowMax = rowMax;
    {
        const SparseList sl = sparse_list(b, b->count[L - 1]);
        A.sparse = sl.slots;
        A.sparseCnt = sl.counts;
        A.sparseUnits = sl.units;
        A.sparseUnitCount = sl.unitCount;
        A.sparseTiles = (b->count[L - 1] + kScanThreads - 1) / kScanThreads;
    }
    if ((rc = launch_range_step<kFinal>(b, A, b->itemCap))) return rc;
    if (A.rowMax && (rc = launch_range_rows(b, A, b->count[L - 1]))) return rc;
    // sort the (query, id) keys; the last pass writes the ids to the caller's
    // buffer and the first position of every query to d_out_off
    uint64_t* sorted = nullptr;
    RadixOut ro;
    ro.ids = d_out_ids;
    ro.idMask = (idBits >= 64) ? ~0ull : ((1ull << idBits) - 1);
    ro.groupOff = d_out_off;
    ro.groupShift = idBits;
    ro.nGroups = nq;
    if ((rc = radix_sort(ctx, b->hits, b->hitsTmp, ctl + 2, A.cap, 0, idBits + qBits, b->hist, &sorted, ro,
                         b->radixBlocksPerCu)))
        return rc;
    hipLaunchKernelGGL(k_copy_status, dim3(1), dim3(64), 0, ctx->stream, ctl, d_status);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_leaf_tests(tblaeq_batch* b, uint64_t* out) {
    if (!b || !out) {
        set_error("tblaeq_batch_leaf_tests: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    TB_HIP(hipStreamSynchronize(b->ctx->stream));
    TB_HIP(hipMemcpy(out, b->ctl + 10, 8, hipMemcpyDeviceToHost));
    return TBLAEQ_OK;
}

extern "C" int tblaeq_batch_range_graph(tblaeq_batch* b, const double* d_boxes, uint64_t nq, uint64_t* d_out_off,
                                        uint64_t* d_out_ids, uint64_t out_cap, uint64_t* d_status) {
    if (!b) return TBLAEQ_ERR_INVALID;
    tblaeq_ctx* ctx = b->ctx;
    const bool same = b->rangeExec && b->gParamVersion == b->paramVersion && b->gBoxes == d_boxes && b->gNq == nq &&
                      b->gOff == d_out_off && b->gIds == d_out_ids && b->gCap == out_cap && b->gStatus == d_status &&
                      b->gCtxStatus == ctx->status && b->gCtxWs == ctx->ws;
    if (!same) {
        if (b->rangeExec) TB_HIP(hipGraphExecDestroy(b->rangeExec));
        if (b->rangeGraph) TB_HIP(hipGraphDestroy(b->rangeGraph));
        b->rangeExec = nullptr;
        b->rangeGraph = nullptr;
        const bool timing = ctx->timing;
        ctx->timing = false;  // no event records inside the graph
        const uint64_t before = ctx->scanLaunches;
        warm_range_kernels(b);
        TB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = tblaeq_batch_range(b, d_boxes, nq, d_out_off, d_out_ids, out_cap, d_status, nullptr, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->timing = timing;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        TB_HIP(e);
        b->rangeGraph = g;
        TB_HIP(hipGraphInstantiate(&b->rangeExec, g, nullptr, nullptr, 0));
        b->gScanLaunches = ctx->scanLaunches - before;
        ctx->scanLaunches = before;
        b->gBoxes = d_boxes;
        b->gParamVersion = b->paramVersion;
        b->gNq = nq;
        b->gOff = d_out_off;
        b->gIds = d_out_ids;
        b->gCap = out_cap;
        b->gStatus = d_status;
        b->gCtxStatus = ctx->status;
        b->gCtxWs = ctx->ws;
    }
    // keep the host count of scan launches (epoch wrap protection) in step with replays
    for (uint64_t i = 0; i < b->gScanLaunches; ++i) {
        int rc = scan_prepare(ctx, 1);
        if (rc) return rc;
    }
    TimedScope ts(ctx, kKBatch0 + 6);
    TB_HIP(hipGraphLaunch(b->rangeExec, ctx->stream));
    return TBLAEQ_OK;
}
