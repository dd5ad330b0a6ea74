// Device state of the batched query engine (shared by the range and KNN kernels).
//
// The engine keeps, for every level l = 0..L, the level's points in "position"
// order: level 0 in coarsest-mesh order, level l >= 1 in the CSC row order of
// P[l-1] (at the finest level, the rows of each column reordered into kd
// order).  Their reconstructed coordinates are produced once, at creation, by
// running the prolongation over the whole hierarchy (every cluster a survivor),
// so they are bit-identical to what the per-query prolongation produces.  A
// cluster c of level l owns the contiguous positions colRes_l[c]..colRes_l[c+1]
// of level l+1.
#pragma once
#include <vector>

#include "ctx.hpp"

struct tblaeq_batch {
    tblaeq_ctx* ctx = nullptr;
    int D = 0;
    int L = 0;                              // intervals
    std::vector<uint64_t> count;            // [L+1] positions per level
    std::vector<double*> coords;            // [L+1] count*D, reconstructed
    std::vector<uint64_t*> ids;             // [L+1] level id of each position
    std::vector<double*> radiusPos;         // [L]   radius of each position's point
    std::vector<uint64_t*> parentPos;       // [L]   (l>=1) level-(l-1) id of the parent
    std::vector<uint64_t*> weightPos;       // [L]   children count of each position's point
    std::vector<const uint64_t*> colRes;    // [L]   P[l] column pointers (by level-l id)
    std::vector<const uint64_t*> nnzPerCol; // [L]
    std::vector<const double*> radiusById;  // [L]
    std::vector<uint64_t> emptyCols;        // [L]
    std::vector<double> maxRadius;          // [L]   host: largest radius of the level
    std::vector<void*> owned;               // device allocations to free

    // range scratch
    uint64_t maxQ = 0, pairCap = 0, hitCap = 0, itemCap = 0, maxMesh = 0;
    uint64_t* pairs = nullptr;   // (id << 24 | q)
    uint32_t* qlist = nullptr;   // grouped query lists
    uint32_t* qlist2 = nullptr;  // second set of groups (range levels alternate)
    uint64_t* off2 = nullptr;
    uint64_t* items2 = nullptr;
    float* aabb = nullptr;       // [count[L-1] * 2D] bounding box of each last-level position's points
    uint64_t* leafOff = nullptr; // [cols of P[L-1] + 1] first kd leaf of every last-level cluster, by id
    float* leafBox = nullptr;    // [leaves * 2D] outward-rounded float box lo[D] | hi[D] of every kd leaf
    uint64_t maxLeaves = 0;      // most kd leaves of one last-level cluster
    uint64_t nLeaves = 0;
    float* fbox = nullptr;       // [maxQ * 2 * dm] outward float boxes of the current range batch
    uint64_t* sparse = nullptr;  // last-level clusters of the row-major exact filter, per scan tile, and counts
    double* enclC = nullptr;     // [count[L-2] * D] (L >= 2) centre of each level-(L-2) cluster, by id
    double* enclR = nullptr;     // [count[L-2]] radius of a ball around it holding its children's balls
    float* enclBox = nullptr;    // [count[L-2] * 2D] box lo[D] | hi[D] holding its children's balls, by id
    uint64_t* maskOff = nullptr; // [itemCap + 1] first keep-mask row of each work item of a level
    uint64_t* cnt = nullptr;     // [maxMesh]  per-cluster pair counts (kept zeroed between steps)
    uint64_t* off = nullptr;     // [maxMesh+1]
    uint64_t* items = nullptr;   // (id << 20 | slice)
    uint64_t* hits = nullptr;    // (q << idBits | id)
    uint64_t* hitsTmp = nullptr;
    uint64_t* recAux = nullptr;  // [hitCap] grouped KNN record ids
    uint64_t* hist = nullptr;    // radix histograms
    uint64_t histCap = 0;
    uint64_t* ctl = nullptr;     // [0] pairs [1] items [2] hits [3] overflow [4..] spare; kCtlWords
    uint64_t* queue = nullptr;   // [kQueueSlots] work-queue counters of the grid-stride kernels (ctl + kCtlWords)
    // captured range batch (replayed while the arguments are unchanged)
    hipGraph_t rangeGraph = nullptr;
    hipGraphExec_t rangeExec = nullptr;
    // captured KNN batch and the arguments it was captured with
    hipGraph_t knnGraph = nullptr;
    hipGraphExec_t knnExec = nullptr;
    const void* kPoints = nullptr;
    const void* kK = nullptr;
    uint64_t kNq = 0, kKmax = 0, kScanLaunches = 0, kParamVersion = 0;
    void* kIds = nullptr;
    void* kD2 = nullptr;
    void* kN = nullptr;
    void* kStatus = nullptr;
    // context buffers baked into the captured graphs (a reallocation recaptures)
    const void* gCtxStatus = nullptr;
    const void* gCtxWs = nullptr;
    const void* kCtxStatus = nullptr;
    const void* kCtxWs = nullptr;
    const void* gBoxes = nullptr;
    uint64_t gNq = 0, gCap = 0;
    void* gOff = nullptr;
    void* gIds = nullptr;
    void* gStatus = nullptr;
    uint64_t gScanLaunches = 0;
    // knn scratch
    uint8_t* flags = nullptr;    // per query survivor flags of a middle level
    uint64_t flagCap = 0;
    // launch parameters (tblaeq_batch_set_param)
    uint64_t knnSeed = 64;       // clusters expanded before the certified radius is taken
    uint64_t knnWgPerCu = 4;     // KNN workgroups per CU (grid = min(groups, CUs * this))
    uint64_t knnMode = 0;        // 0 cluster-major refine, 1 query-major refine
    uint64_t rangeBlocksPerCu = 16;   // range step kernels (4 waves per block)
    uint64_t stepBlocksPerCu = 16;    // cluster-major KNN passes and refine
    uint64_t radixBlocksPerCu = 4;    // hit radix sort
    uint64_t hitStage = 512;          // LDS hits staged per wave by the final filter (256, 512, 1024)
    uint64_t rangeRowMax = 16;        // final filter: clusters kept by at most this many queries go row-major
    uint64_t knnRowMax = 4;           // KNN refine: clusters refined by at most this many queries go row-major
    uint64_t knnLeafRefine = 0;       // KNN refine: work items through kd-leaf tests (1) or all rows (0)
    uint64_t paramVersion = 0;        // bumped by every parameter change (captured graphs re-capture)
    uint64_t gParamVersion = 0;
};

namespace tblaeq {
constexpr int kSliceBits = 20;
constexpr int kQueryBits = 24;
constexpr int kStepWaves = 4;  // waves per workgroup of the cluster-major kernels
// The finest rows of every last-level cluster are stored in kd order (engine
// creation): each run of kLeafRows consecutive rows of a cluster is a leaf with
// its own bounding box, tested before its rows by the range exact filter.
constexpr int kLeafRows = 16;
constexpr int kFilterPadRows = 4;  // spare rows after the finest level's last row

// chunked counting scatters: kAggPer elements per thread, targets ranked in
// LDS when there are at most kAggLds of them
constexpr int kAggThreads = 256;
constexpr int kAggPer = 16;
constexpr int kAggChunk = kAggThreads * kAggPer;
constexpr int kAggLds = 8192;

constexpr int kCtlWords = 16;
constexpr int kQueueSlots = 64;
constexpr int kMaskQueue0 = 8;  // queue slots kMaskQueue0 + l: keep-mask rows reserved at level l
// ctl slots: [0] pairs [1] items [2] hits / records [3] overflow flags
// (1 pairs, 2 items, 4 hits / records) [4] items of the second group set
// [6], [7] scratch of count-only scans.

// Groups the (cluster id << kQueryBits | q) pairs counted in b->cnt (mesh
// clusters) by cluster: b->off, work items (cluster, 64-query slice) in
// b->items with their count in ctl[1], query lists in b->qlist.  Leaves
// b->cnt zeroed and resets the pair count ctl[0].
int launch_grouping(tblaeq_batch* b, uint64_t mesh);
// Group scan of b->cnt over mesh clusters into off / items (count in *nItems).
// rowMax > 0: clusters of 1..rowMax pairs get no work items; they go to the
// row-major list in b->sparse instead (sparse_list).
int launch_group_scan(tblaeq_batch* b, uint64_t mesh, uint64_t* off, uint64_t* items, uint64_t* nItems,
                      uint64_t rowMax = 0);
// Row-major cluster list of a level of n clusters, in b->sparse: kScanThreads
// slots per group-scan tile (the tile's listed clusters first, in order), then
// every tile's count, then the non-empty units of kRowUnit slots that the
// row-major kernels take as work (unit u: slots u * kRowUnit ..), their number
// in queue slot kSparseUnitQueue (zeroed with the control block).
constexpr int kRowUnit = 16;
constexpr int kSparseUnitQueue = 2;
constexpr uint64_t sparse_slots(uint64_t n) { return (n + kScanThreads - 1) / kScanThreads * kScanThreads; }
constexpr uint64_t sparse_words(uint64_t n) {
    return sparse_slots(n) + sparse_slots(n) / kScanThreads + sparse_slots(n) / kRowUnit;
}
struct SparseList {
    uint64_t* slots = nullptr;
    uint64_t* counts = nullptr;
    uint64_t* units = nullptr;
    uint64_t* unitCount = nullptr;
};
// The row-major list of a level of n clusters in b->sparse.
SparseList sparse_list(tblaeq_batch* b, uint64_t n);
// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total).
int launch_count_scan(tblaeq_batch* b, const uint64_t* cnt, uint64_t n, uint64_t* off);
// status[0] = ctl[2], status[1] = ctl[3].
void launch_copy_status(tblaeq_batch* b, uint64_t* status);
// Fills n words with v (a kernel, so that captured batches stay replayable).
void launch_fill(tblaeq_batch* b, uint64_t* p, uint64_t v, uint64_t n);
// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total) on ctx's stream;
// scratch2 = two device words.
int scan_counts(tblaeq_ctx* ctx, const uint64_t* cnt, uint64_t n, uint64_t* off, uint64_t* scratch2);
// Outputs of the last radix pass: ids (optional) receives key & idMask instead
// of the ping-pong buffer; groupOff (optional, nGroups + 1 words) receives the
// first sorted position of every group key >> groupShift (groups without keys:
// the next group's start; groupOff[nGroups] = key count; every key's group is
// below nGroups).
struct RadixOut {
    uint64_t* ids = nullptr;
    uint64_t idMask = ~0ull;
    uint64_t* groupOff = nullptr;
    int groupShift = 0;
    uint64_t nGroups = 0;
};
// Stable LSD radix sort of *dn (<= cap < 2^30) u64 keys on bits [firstBit,
// firstBit + nbits) between a and b (8- or 9-bit digits); *result = the buffer
// holding the sorted keys (ro.ids when given).  hist needs
// radix_hist_words(cap, ro.nGroups) words.
int radix_sort(tblaeq_ctx* ctx, uint64_t* a, uint64_t* b, const uint64_t* dn, uint64_t cap, int firstBit, int nbits,
               uint64_t* hist, uint64_t** result, const RadixOut& ro, uint64_t blocksPerCu);
uint64_t radix_hist_words(uint64_t cap, uint64_t groups);
// Bits needed to represent v (at least 1).
int bits_for(uint64_t v);
// Blocks of `kernel` (with `threads` threads) that fit on one CU at once,
// queried once per kernel.  Work-queue kernels launch at most this many per CU:
// every wave is resident from the start and takes items until the queue is dry.
int resident_blocks(const void* kernel, int threads);
// Moves n rows of `width` 8-byte words on ctx's stream: scatter ? dst[idx[i]] =
// src[i] : dst[i] = src[idx[i]] (idx device, distinct for a scatter).
void move_rows(tblaeq_ctx* ctx, bool scatter, const void* src, const uint64_t* idx, uint64_t n, uint64_t width,
               void* dst);
}  // namespace tblaeq
