/*
 * tblaeq.h -- C ABI of the MI355X-native T-BLAEQ query path (gfx950 HIP kernels).
 *
 * The reference (Mgepahmge/T-BLAEQ) has no plugin or FFI layer: its operator
 * interface is the C++ class set QueryHandler / QueryEngine / IQueryStrategy.
 * This header is the thin C layer under the MI355X-native re-implementation of
 * those classes; each entry point names the reference code it replaces.
 *
 * Conventions
 *  - Every call returns an int status (TBLAEQ_OK = 0) and never exits the
 *    process; tblaeq_last_error() gives a message for the calling thread.
 *  - Device pointers are plain HIP device addresses owned by the caller.
 *  - Element counts that are produced on the device stay on the device
 *    (uint64_t* d_count).  Calls that consume such a count also take a host-side
 *    upper bound (`*_bound`) that sizes the launch; the device count selects the
 *    work, so no host synchronisation is needed between levels.
 *  - Work is issued on the context's stream, in order; calls are not
 *    thread-safe on one context.
 *  - Coordinates are FP64, ids and counts are 64-bit.
 */
#ifndef TBLAEQ_H
#define TBLAEQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBLAEQ_OK 0
#define TBLAEQ_ERR_INVALID 1     /* bad argument (null pointer, zero dimension, ...) */
#define TBLAEQ_ERR_HIP 2         /* HIP runtime error */
#define TBLAEQ_ERR_UNSUPPORTED 3 /* shape outside what the kernels handle */
#define TBLAEQ_ERR_CAPACITY 4    /* an output or scratch buffer is too small */
#define TBLAEQ_ERR_IO 5          /* index file missing or malformed */
#define TBLAEQ_ERR_NOMEM 6       /* device or host allocation failed */
#define TBLAEQ_ERR_TIMEOUT 7     /* a wait ended before the work was done */
#define TBLAEQ_ERR_INTERNAL 8    /* unexpected internal error (message in tblaeq_last_error) */

typedef struct tblaeq_ctx tblaeq_ctx;

/* One interval of the hierarchy, resident on the device.  Mirrors the device
 * half of SparseTensorCscFormat + MaxRadius + Map (Data_Structures.cuh:299-390,
 * IndexData.cuh:89-243).  child_ids[j] = map[l+1][rowIds[j]] folds the
 * reference's refactor pass (Refactor.cu:10-16) into the index. */
typedef struct tblaeq_level {
    uint64_t D;                  /* dimensionality */
    uint64_t cols;               /* points of the coarse level l */
    uint64_t nnz;                /* points of the fine level l+1 */
    const uint64_t* col_res;     /* [cols+1] CSC column pointers */
    const uint64_t* nnz_per_col; /* [cols]   children per cluster (KNN STEP weights) */
    const double* pvals;         /* [nnz*D]  P-tensor ratios, AoS */
    const double* radius;        /* [cols]   max child distance per cluster */
    const uint64_t* child_ids;   /* [nnz]    fine-level id of CSC row j */
    uint64_t empty_cols;         /* clusters with nnz_per_col == 0 (bounds KNN STEP phase 1) */
} tblaeq_level;

/* --- context --------------------------------------------------------------
 * Replaces the implicit CUDA default-stream state of the reference: owns a
 * stream (or adopts `stream` when non-null), the look-back status words and a
 * growable scratch area. */
int tblaeq_ctx_create(int device, void* stream, tblaeq_ctx** out);
int tblaeq_ctx_destroy(tblaeq_ctx* ctx);
int tblaeq_ctx_sync(tblaeq_ctx* ctx);
void* tblaeq_ctx_stream(tblaeq_ctx* ctx);
const char* tblaeq_status_string(int status);
/* Kernel timing with HIP events recorded on the context stream (no reference
 * counterpart; the reference times whole queries with steady_clock).  Kinds:
 * 0 range compaction, 1 KNN compaction, 2 box filter, 3 prolongation, 4 tile
 * sort, 5 merge pass, 6 STEP, 7 KNN bounds, 8.. batched-path kernels.  `tag`
 * labels subsequent launches (e.g. the level); tag -1 in a query sums all. */
int tblaeq_ctx_set_timing(tblaeq_ctx* ctx, int enable);
int tblaeq_ctx_set_tag(tblaeq_ctx* ctx, int tag);
int tblaeq_ctx_timing(tblaeq_ctx* ctx, int kind, int tag, uint64_t* launches, double* ms);
int tblaeq_ctx_timing_reset(tblaeq_ctx* ctx);
const char* tblaeq_last_error(void);
/* Progress messages of long-running calls (the GPU index build reports its
 * seeding rounds, seeding picks and Lloyd iterations): fn(message, user) is
 * called on the calling thread; fn = NULL turns them off (the default).
 * Process-wide.  No reference counterpart. */
typedef void (*tblaeq_progress_fn)(const char* message, void* user);
void tblaeq_set_progress(tblaeq_progress_fn fn, void* user);

/* --- plain per-query level kernels ---------------------------------------- */

/* Range prune + order-preserving compaction, fused.
 * Replaces rangePruningKernel + countKernel + prefixSumKernel +
 * prefixSumFinalScanKernel + gridCompactPrefixKernel and the host mask count
 * (RangePruning.cu:5-48, compact.cuh:34-268, GridCompact.cuh:41-116,
 * L0Strategy.cu:83-101,160-172).
 * Keeps grid point i iff boxdist^2(vals_i, [lo,hi]) <= radius[ids_i]^2.
 * Writes survivors (ids, vals) in grid order, surv_off[s] = exclusive prefix of
 * the survivors' CSC column lengths (their output offset in the prolongation),
 * *d_surv_count and *d_out_count (= next level's grid size). */
int tblaeq_range_prune_compact(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids, const double* vals,
                               const uint64_t* d_count, uint64_t count_bound, const double* d_lo, const double* d_hi,
                               uint64_t* surv_ids, double* surv_vals, uint64_t* surv_off, uint64_t* d_surv_count,
                               uint64_t* d_out_count);

/* Allocation-free variant of tblaeq_range_prune_compact.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernel, arguments and results, but it takes the caller's stream (a
 * hipStream_t, null = the default stream) instead of a context and keeps its
 * look-back status words in a fixed device array of the library (512 KiB per
 * device, held until the process ends), so it calls no hipMalloc, hipFree or
 * allocating library routine.  count_bound may be at most 16 777 216 grid points
 * (else TBLAEQ_ERR_CAPACITY).  All calls on one device share that array:
 * calls that overlap on one device (concurrent streams or threads) race and
 * must be serialised by the caller. */
int tblaeq_range_prune_compact_noalloc(void* stream, const tblaeq_level* lv, const uint64_t* ids, const double* vals,
                                       const uint64_t* d_count, uint64_t count_bound, const double* d_lo,
                                       const double* d_hi, uint64_t* surv_ids, double* surv_vals, uint64_t* surv_off,
                                       uint64_t* d_surv_count, uint64_t* d_out_count);

/* KNN lower bound, selection of the closest clusters and STEP, then compaction.
 * Replaces calculateClusterDistanceKernel + gpuSort + host STEP loop + mask
 * upload + compaction (KNN.cu:9-110, MergeSort.cuh:628-696,
 * L0Strategy.cu:103-158).  lb_i = sqrt_rd(|q - c_i|^2) - r_i; clusters are
 * ordered by (lb, id); STEP takes clusters while the running child count is
 * below K, raising ub = max(ub, lb + 2r), then keeps every cluster with lb <= ub.
 * Outputs as tblaeq_range_prune_compact. */
int tblaeq_knn_lower_bound_sort_step(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                     const double* vals, const uint64_t* d_count, uint64_t count_bound,
                                     const double* d_q, uint64_t K, uint64_t* surv_ids, double* surv_vals,
                                     uint64_t* surv_off, uint64_t* d_surv_count, uint64_t* d_out_count);

/* SpTSpM prolongation fused with the id refactor.
 * Replaces SpTSpMMultiplication_v3{,_L0_nb} + SpMSpVKernelAOS_v2 + refactor
 * (SpMSpV.cu:20-105,200-254, SpMSpV.cuh:47-67, Refactor.cu:10-27).  For every
 * survivor s (cluster c) and CSC row j of column c, in survivor order then
 * column order: out_ids[o] = child_ids[j], out_vals[o,d] = pvals[j,d] * surv_vals[s,d]
 * with o = surv_off[s] + (j - col_res[c]).  No host-built index arrays. */
int tblaeq_prolong_refactor(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* surv_ids,
                            const double* surv_vals, const uint64_t* surv_off, const uint64_t* d_surv_count,
                            uint64_t surv_bound, uint64_t* out_ids, double* out_vals);

/* Prolongation from P columns packed in survivor order: survivor s's P rows
 * (and, when packed_child is non-null, its child ids) start at packed row
 * surv_off[s].  Replaces SpTSpMMultiplication_v3_L2 (SpMSpV.cu:107-198), used by
 * the streaming policies that do not keep P resident (L2, L3).  lv->pvals is
 * not read; lv->col_res gives the column lengths. */
int tblaeq_prolong_refactor_packed(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* surv_ids,
                                   const double* surv_vals, const uint64_t* surv_off, const uint64_t* d_surv_count,
                                   uint64_t surv_bound, const double* packed_pvals, const uint64_t* packed_child,
                                   uint64_t* out_ids, double* out_vals);

/* KNN lower bounds lb_i = sqrt_rd(|q - c_i|^2) - radius[ids_i] for `count`
 * grid points (host count), written to out_lb.  Used by the tiled policy (L3)
 * to run STEP over a grid that does not fit on the device
 * (calculateClusterDistanceKernelL3, L3Kernels.cu:28-47). */
int tblaeq_knn_lower_bounds(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals, uint64_t count,
                            const double* d_q, const double* radius, double* out_lb);

/* Range prune + compaction over a tile whose element count is host-known and
 * whose radius is indexed by position (pre-gathered), as
 * rangePruningKernelL3 (L3Kernels.cu:3-26).  Same outputs as
 * tblaeq_range_prune_compact, with surv_off relative to the tile. */
int tblaeq_range_prune_compact_tile(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids,
                                    const double* vals, uint64_t count, const double* d_lo, const double* d_hi,
                                    uint64_t* surv_ids, double* surv_vals, uint64_t* surv_off,
                                    uint64_t* d_surv_count, uint64_t* d_out_count);

/* --- exact refine (no reference counterpart: the reference returns the
 * candidate superset; these give the exact answer over the reconstructed
 * coordinates of the candidates) ------------------------------------------ */

/* Ids of candidates inside [lo,hi] (inclusive), ascending.  `count` is the
 * host-known candidate count.  Returns the hit count in *n_out (host). */
int tblaeq_refine_range(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals, uint64_t count,
                        const double* d_lo, const double* d_hi, uint64_t* out_ids, uint64_t out_cap,
                        uint64_t* n_out);

/* Top-K candidates by (squared distance, id), ascending; d2 = fma chain of (q-v)^2. */
int tblaeq_refine_knn_topk(tblaeq_ctx* ctx, uint64_t D, const uint64_t* ids, const double* vals, uint64_t count,
                           const double* d_q, uint64_t K, uint64_t* out_ids, double* out_d2, uint64_t* n_out);

/* Learned KNN radius filter (no reference counterpart): the KNN compaction of
 * a grid with the bound *d_radius in place of the STEP bound -- keeps the
 * points with sqrt_rd(|q - c|^2) - r_c <= *d_radius in grid order, with their
 * prolongation offsets; *d_out_count = kept points, *d_out_nnz = their
 * children.  Used on the last level's STEP survivors so that only clusters
 * within a proposed radius are prolonged. */
int tblaeq_knn_radius_compact(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* ids, const double* vals,
                              const uint64_t* d_count, uint64_t count_bound, const double* d_q, const double* d_radius,
                              uint64_t* out_ids, double* out_vals, uint64_t* out_off, uint64_t* d_out_count,
                              uint64_t* d_out_nnz);

/* --- sort ------------------------------------------------------------------
 * Block merge sort of (key, value) pairs, lexicographic, ascending.  LDS
 * bitonic tiles of 2048 pairs, then merge-path passes.  Replaces
 * blockMergeSortKernel + hostSerialMergeSort (MergeSort.cuh:240-444,628-696);
 * also used to build CSC P-tensors from (cluster, point) keys. */
int tblaeq_sort_pairs(tblaeq_ctx* ctx, uint64_t* keys, uint64_t* vals, uint64_t n);

/* --- batched engine ----------------------------------------------------------
 * No reference counterpart (the reference runs one query at a time through
 * QueryEngine::run, QueryEngine.cu:151-214).  Thousands of queries share one
 * launch per level; results equal the per-query path's exact refine. */
typedef struct tblaeq_batch tblaeq_batch;

/* Builds the engine over resident levels (P values, radii, column pointers on
 * the device): reconstructs every level's coordinates once with the
 * prolongation kernel.  host_max_radius[l] = largest radius of level l (or
 * NULL).  D must be <= 32. */
int tblaeq_batch_create(tblaeq_ctx* ctx, const tblaeq_level* levels, uint64_t nlevels, const uint64_t* coarsest_ids,
                        const double* coarsest_vals, uint64_t ncoarsest, const double* host_max_radius,
                        tblaeq_batch** out);
int tblaeq_batch_destroy(tblaeq_batch* b);
/* Scratch for up to max_queries per batch, pair_cap (cluster, query) survivor
 * pairs per level and hit_cap range hits per batch (KNN batches use
 * min(pair_cap, hit_cap) candidate records).  Allocates; call before
 * timing or graph capture. */
int tblaeq_batch_reserve(tblaeq_batch* b, uint64_t max_queries, uint64_t pair_cap, uint64_t hit_cap);
/* nq range boxes (device, rows lo|hi).  Per-query hit ids, ascending, in
 * d_out_ids[d_out_off[q] .. d_out_off[q+1]).  d_status[0] = total hits,
 * d_status[1] = overflow flags (non-zero: capacity exceeded, results invalid;
 * split the batch), d_status[2] = box tests of the final filter (candidate
 * rows x queries), d_status[3] = candidate rows it scanned (d_status holds 4
 * words).  Optional d_level_surv[q*L + l] (survivors per level) and
 * d_fine_count[q] (candidate superset size) must be zeroed by the caller.
 * Asynchronous on the context stream; no host synchronisation inside. */
int tblaeq_batch_range(tblaeq_batch* b, const double* d_boxes, uint64_t nq, uint64_t* d_out_off, uint64_t* d_out_ids,
                       uint64_t out_cap, uint64_t* d_status, uint64_t* d_level_surv, uint64_t* d_fine_count);
/* Leaf x box tests of the exact filter in the engine's last range batch (the
 * float tests of the kd leaves' boxes that decide which rows are tested);
 * synchronises the context stream.  No reference counterpart (measurement). */
int tblaeq_batch_leaf_tests(tblaeq_batch* b, uint64_t* out);
/* tblaeq_batch_range as one hipGraph launch: the first call with a given set
 * of arguments captures the whole batch (level loop, grouping, hit sort) on
 * the context stream; later calls with the same arguments replay it (the
 * buffers' contents may change between calls).  No statistics outputs. */
int tblaeq_batch_range_graph(tblaeq_batch* b, const double* d_boxes, uint64_t nq, uint64_t* d_out_off,
                             uint64_t* d_out_ids, uint64_t out_cap, uint64_t* d_status);
/* nq KNN points (device, rows of D), K per query in d_K (each <= Kmax).
 * Top-K ids and squared distances by (distance, id) in rows of Kmax;
 * d_out_n[q] = number written (min(K, candidates)).  Same traversal (STEP per
 * level) as the per-query path.  d_status[0] = candidate records refined
 * cluster-major, d_status[1] = overflow flags (non-zero: a capacity of
 * tblaeq_batch_reserve was exceeded, results invalid; split the batch),
 * d_status[2] = distance evaluations of the refine (cluster-major: child
 * rows times the work item's queries; row-major: for each query, the rows of
 * the kd leaves its box test keeps),
 * d_status[3] = distance and bound evaluations of the query-major expansions
 * (seed clusters, last-level bound) (d_status holds 4 words).
 * Requires Kmax + empty clusters + 1 <= 256 and at most 8 levels (else
 * TBLAEQ_ERR_UNSUPPORTED).  d_level_surv / d_fine_count as for
 * tblaeq_batch_range (written, not accumulated). */
int tblaeq_batch_knn(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, uint64_t Kmax,
                     uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n, uint64_t* d_status,
                     uint64_t* d_level_surv, uint64_t* d_fine_count);
/* Largest K the batched KNN kernels take for this index (256 - empty
 * clusters - 1); the handler and the server route queries with a larger K to
 * the per-query traversal. */
int tblaeq_batch_knn_max_k(const tblaeq_batch* b, uint64_t* out);
/* tblaeq_batch_knn as one hipGraph launch: the first call with a given set of
 * arguments captures the whole batch (levels, grouping, last-level passes,
 * selection, refine, merge) on the context stream; later calls with the same
 * arguments replay it (the buffers' contents may change between calls).  No
 * statistics outputs. */
int tblaeq_batch_knn_graph(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, uint64_t Kmax,
                           uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n, uint64_t* d_status);

/* Launch parameters of the batched engine (no effect on results, only on
 * speed); tblaeq_batch_get_param reads the current value.
 *   TBLAEQ_PARAM_KNN_SEED           last-level clusters expanded first to fix
 *                                   the certified radius, 1..256 (default 64)
 *   TBLAEQ_PARAM_KNN_WG_PER_CU      query-major KNN workgroups per CU, 1..16 (4)
 *   TBLAEQ_PARAM_KNN_MODE           0: the last level's surviving clusters are
 *                                   grouped and refined cluster-major (default);
 *                                   1: each workgroup refines its own queries
 *   TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU  range level kernels, 1..64 (16)
 *   TBLAEQ_PARAM_STEP_BLOCKS_PER_CU   cluster-major KNN kernels, 1..64 (16)
 *   TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU  hit radix sort, 1..32 (4)
 *   TBLAEQ_PARAM_HIT_STAGE          LDS hit stage per wave of the final range
 *                                   filter: 256, 512 or 1024 (512)
 *   TBLAEQ_PARAM_RANGE_ROW_QUERIES  last-level clusters kept by at most this
 *                                   many queries are filtered row-major (lane =
 *                                   point, one query at a time), the others
 *                                   query-major, 0..64 (16; 0: all query-major;
 *                                   values above 256 / dm take effect as 256 /
 *                                   dm, dm = 4, 8, 16 or 32 the instantiated
 *                                   dimension at or above D)
 *   TBLAEQ_PARAM_KNN_ROW_QUERIES    last-level clusters refined by at most this
 *                                   many KNN queries are refined row-major
 *                                   (lane = (cluster, query) pair for the leaf
 *                                   tests, then lane = point of a kept leaf),
 *                                   the others query-major, 0..4 (4; 0: all
 *                                   query-major; only with more than one level)
 *   TBLAEQ_PARAM_KNN_LEAF_REFINE    1: the other clusters' work items (a cluster
 *                                   and up to 64 queries) are refined through
 *                                   per (query, kd leaf) box tests too, then the
 *                                   kept leaves' rows, four leaves a pass; 0 (the
 *                                   default): every row against every query of
 *                                   the item (pays where the boxes prune little) */
enum {
    TBLAEQ_PARAM_KNN_SEED = 1,
    TBLAEQ_PARAM_KNN_WG_PER_CU = 2,
    TBLAEQ_PARAM_KNN_MODE = 3,
    TBLAEQ_PARAM_RANGE_BLOCKS_PER_CU = 4,
    TBLAEQ_PARAM_STEP_BLOCKS_PER_CU = 5,
    TBLAEQ_PARAM_RADIX_BLOCKS_PER_CU = 6,
    TBLAEQ_PARAM_HIT_STAGE = 7,
    TBLAEQ_PARAM_RANGE_ROW_QUERIES = 8,
    TBLAEQ_PARAM_KNN_ROW_QUERIES = 9,
    TBLAEQ_PARAM_KNN_LEAF_REFINE = 10
};
int tblaeq_batch_set_param(tblaeq_batch* b, int param, uint64_t value);
int tblaeq_batch_get_param(const tblaeq_batch* b, int param, uint64_t* value);

/* Autotuner: times candidate values of the launch parameters on a sample
 * batch (n_boxes range boxes and / or n_points KNN points with K in d_K,
 * device memory) with HIP events, one parameter at a time, and keeps the
 * fastest; every candidate's results are hashed on the device and must equal
 * the starting configuration's (else TBLAEQ_ERR_INVALID).  Allocates.
 * out_ms (optional, 4 doubles): range ms before / after, KNN ms before / after
 * (median of reps runs).  The sample must fit the reserved capacities. */
int tblaeq_batch_autotune(tblaeq_batch* b, const double* d_boxes, uint64_t n_boxes, const double* d_points,
                          const uint64_t* d_K, uint64_t n_points, uint64_t Kmax, int reps, double* out_ms);
/* Tuning table: one "name value" line per launch parameter, with a comment
 * header naming the index shape; load applies each line through
 * tblaeq_batch_set_param. */
int tblaeq_batch_save_params(const tblaeq_batch* b, const char* path);
int tblaeq_batch_load_params(tblaeq_batch* b, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* TBLAEQ_H */
