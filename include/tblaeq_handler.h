/*
 * tblaeq_handler.h -- C facade over the MI355X-native QueryHandler.
 *
 * The reference's public interface is the C++ class QueryHandler
 * (src/core/QueryHandler.cuh:50-247); these functions expose the same
 * operations to callers that bind C (ctypes, cgo, JNI).  All return an int
 * status from tblaeq.h; tblaeq_last_error() describes failures.
 */
#ifndef TBLAEQ_HANDLER_H
#define TBLAEQ_HANDLER_H

#include <stdint.h>

#include "tblaeq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tblaeq_handler tblaeq_handler;
typedef struct tblaeq_result tblaeq_result;

/* QueryHandler(N, D, valMin, valMax, isInt, height, ratios, seed, sigmaDivisor)
 * (QueryHandler.cuh:95-101, QueryHandler.cu:101-124): synthetic hierarchy. */
int tblaeq_handler_create_random(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, uint64_t height,
                                 const uint64_t* ratios, uint64_t seed, double sigmaDivisor, tblaeq_handler** out);
/* No reference counterpart: the generator's hierarchy shape (coarsest layer
 * uniform in [vmin, vmax]^D, child i of a layer under parent i % parents,
 * Gaussian with sigma = spacing / sigmaDivisor, clamped, rounded when isInt)
 * built on `threads` host threads (0 = all): each chunk of 65536 values or
 * children draws from its own mt19937_64 seeded by SplitMix64 of (seed, level,
 * chunk), so the index does not depend on the thread count, and each P-tensor
 * column lists its rows in ascending order.  For indexes too large for the
 * single-stream generator above (hundreds of millions of points); its points
 * are not the reference's sequence.  Dataset name "Syn<N>D<D>". */
int tblaeq_handler_create_synthetic(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, uint64_t height,
                                    const uint64_t* ratios, uint64_t seed, double sigmaDivisor, uint64_t threads,
                                    tblaeq_handler** out);
/* QueryHandler(forceCPU, PointCloud, name, height, ratios) (QueryHandler.cu:79-99):
 * k-means build of an in-memory dataset (data: N*D doubles, AoS). */
int tblaeq_handler_create_from_data(const double* data, uint64_t N, uint64_t D, uint64_t height,
                                    const uint64_t* ratios, const char* name, tblaeq_handler** out);
/* QueryHandler(indexPath, true) (QueryHandler.cu:126-137). */
int tblaeq_handler_load(const char* path, tblaeq_handler** out);
/* saveIndex (QueryHandler.cu:139-145), byte-compatible files. */
int tblaeq_handler_save(tblaeq_handler* h, const char* path);
int tblaeq_handler_destroy(tblaeq_handler* h);
/* meta = {D, N, height, intervals}; meshSizes finest first (up to cap). */
int tblaeq_handler_meta(tblaeq_handler* h, uint64_t* meta, uint64_t* meshSizes, uint64_t cap);
/* Finest generated data layer (generator-built indexes only). */
int tblaeq_handler_data(tblaeq_handler* h, double* out, uint64_t cap);
/* Rows rows[0..n) of the generated finest data (n*D doubles into out); for
 * callers that need a sample of a large index without copying all of it.  No
 * reference counterpart. */
int tblaeq_handler_data_rows(tblaeq_handler* h, const uint64_t* rows, uint64_t n, double* out);
/* prepareForQuery (QueryHandler.cu:160-182): levels != NULL forces a per-level
 * policy (0..3 = L0..L3); otherwise the scheduler picks one from the device's
 * free memory, or from freeMemOverride when it is non-zero. */
int tblaeq_handler_prepare(tblaeq_handler* h, const uint8_t* levels, uint64_t n, uint64_t freeMemOverride);
int tblaeq_handler_policy(tblaeq_handler* h, uint8_t* out, uint64_t cap);
/* PolicyScheduler::recommend with an injected free-memory value (MemoryPolicy.cu:100-177). */
int tblaeq_handler_recommend(tblaeq_handler* h, uint64_t freeMem, int countWorkBuffers, uint8_t* out, uint64_t cap);
/* The handler's execution context (stream, timing) for the tblaeq.h calls. */
int tblaeq_handler_ctx(tblaeq_handler* h, tblaeq_ctx** out);
/* Testing hook: fixed device byte budget for L3 tiles (0 = from free memory). */
int tblaeq_handler_set_l3_tile_bytes(uint64_t bytes);
/* performQuery over an in-memory batch (QueryHandler.cu:198-248): type 0 =
 * range (rows lo|hi, 2*D values), 1 = KNN (rows of D values, K neighbours).
 * refine = 1 adds the exact answer (copied to the host), 2 computes it on the
 * device and reports only its size; saveFineMesh keeps host copies of the final
 * candidate grids. */
int tblaeq_handler_query(tblaeq_handler* h, int type, const double* queries, uint64_t nq, uint64_t K, int refine,
                         int saveFineMesh, tblaeq_result** out);
/* QueryHandler::performQuery(queryPath, qType, saveFineMesh, maxQueryCount, K)
 * (QueryHandler.cu:198-218): the query file's first max_queries queries
 * (type 0 = range, 1 = KNN) through the per-query traversal, timed the
 * reference's way (QueryEngine.cu:168-177: min and max dropped when more than
 * two queries run); no refine.  A file whose dimension differs from the index
 * gives TBLAEQ_ERR_INVALID (the reference returns a result with errorCode 1);
 * a missing or malformed file TBLAEQ_ERR_IO. */
int tblaeq_handler_query_file(tblaeq_handler* h, const char* path, int type, int save_fine_mesh, int max_queries,
                              uint64_t K, tblaeq_result** out);
/* Dataset text file (loadFromFile, File.cu:8-50): a line "N D", then N rows of
 * D numbers.  *n and *dim receive the sizes; min(N*D, cap) values are copied to
 * out (out may be NULL).  Missing or malformed file: TBLAEQ_ERR_IO. */
int tblaeq_dataset_load(const char* path, uint64_t* n, uint64_t* dim, double* out, uint64_t cap);
/* Query text file (loadQueryPointFromFile / loadQueryRangeFromFile,
 * File.cu:139-216): a line "N D", then N rows of D numbers (type 1, points) or
 * of 2D numbers, the lower corner then the upper one (type 0, ranges).  *n,
 * *dim and out as for tblaeq_dataset_load (out receives N rows); for a range
 * file, info (info_cap bytes, NUL-terminated) receives the label its name
 * gives in the CSV's "Query Parameter" column (Utils.cu:18-35). */
int tblaeq_query_file_load(const char* path, int type, uint64_t* n, uint64_t* dim, double* out, uint64_t cap,
                           char* info, uint64_t info_cap);
int tblaeq_result_destroy(tblaeq_result* r);
int tblaeq_result_info(tblaeq_result* r, uint64_t* nq, uint64_t* levels, double* totalUs);
int tblaeq_result_query(tblaeq_result* r, uint64_t i, uint64_t* levelSizes, double* us, uint64_t* fineSize,
                        uint64_t* refinedCount);
/* Per level of query i: clusters kept by the prune and points produced by the prolongation. */
int tblaeq_result_levels(tblaeq_result* r, uint64_t i, uint64_t* survivors, uint64_t* outputs);
int tblaeq_result_fine_mesh(tblaeq_result* r, uint64_t i, uint64_t* ids, double* vals);
int tblaeq_result_refined(tblaeq_result* r, uint64_t i, uint64_t* ids, double* d2);
/* saveQueryResult (QueryEngine.cu:254-305), same CSV columns. */
int tblaeq_result_save_csv(tblaeq_result* r, const char* path);

/* --- batched engine (tblaeq_batch_* in tblaeq.h) over the handler's index ---
 * batch_init prepares the index (L0 if not yet prepared), builds the engine and
 * reserves scratch for maxQueries queries, pairCap survivor pairs per level and
 * hitCap range hits.  set_queries uploads a batch (type 0 = range rows lo|hi,
 * 1 = KNN rows with K per query).  batch_run executes it on the device (range
 * batches that exceed a capacity are split and rerun); *total = number of hits
 * or neighbours.  Results and per-query statistics stay on the device until
 * batch_results / batch_stats copy them out. */
/* Batch slots (no reference counterpart): the handler holds up to two batched
 * engines, each on its own HIP stream; batch_slot selects the one every
 * tblaeq_handler_batch_* call acts on (slot 0 runs on the index's stream and is
 * the default; batch_init on slot 1 creates its own context).  run_slots runs
 * the uploaded batches of the slots in mask (bit i = slot i) concurrently --
 * e.g. a range batch in slot 0 beside a KNN batch in slot 1 -- each replayed as
 * its captured graph, and waits for all; totals[i] = hits or neighbours of slot
 * i.  A slot whose batch exceeds its capacities (or holds K above the batched
 * KNN limit, or has graphs turned off) is then run as batch_run runs it.
 * Results as for a batch_run of each slot. */
int tblaeq_handler_batch_slot(tblaeq_handler* h, int slot);
int tblaeq_handler_batch_run_slots(tblaeq_handler* h, uint64_t mask, uint64_t* totals);
int tblaeq_handler_batch_init(tblaeq_handler* h, uint64_t maxQueries, uint64_t pairCap, uint64_t hitCap);
int tblaeq_handler_batch_set_queries(tblaeq_handler* h, int type, const double* queries, const uint64_t* Ks,
                                     uint64_t nq);
int tblaeq_handler_batch_run(tblaeq_handler* h, int stats, uint64_t* total);
/* range: off[nq+1] and ids[off[nq]]; KNN: n[nq], ids and d2 in rows of Kmax. */
int tblaeq_handler_batch_results(tblaeq_handler* h, uint64_t* off, uint64_t* ids, double* d2, uint64_t* n);
/* levelSurv[q*L + l] survivors per level, fine[q] candidate superset size (when run with stats). */
int tblaeq_handler_batch_stats(tblaeq_handler* h, uint64_t* levelSurv, uint64_t* fine);
/* Range batches run as a captured hipGraph (default on) or as plain launches. */
int tblaeq_handler_batch_use_graph(tblaeq_handler* h, int on);
/* Work of the last batch_run's final stage (summed over split sub-batches):
 * work[0] = tests (range: candidate rows x boxes of the exact filter; KNN:
 * distance evaluations of the cluster-major refine), work[1] = candidate rows
 * scanned. */
int tblaeq_handler_batch_work(tblaeq_handler* h, uint64_t* work);
/* tblaeq_batch_leaf_tests of the handler's engine (the last range sub-batch). */
int tblaeq_handler_batch_leaf_tests(tblaeq_handler* h, uint64_t* out);
/* tblaeq_batch_autotune on the handler's engine with host sample queries
 * (boxes rows lo|hi, KNN points and K per point); out_ms as there. */
int tblaeq_handler_batch_autotune(tblaeq_handler* h, const double* boxes, uint64_t nBoxes, const double* points,
                                  const uint64_t* Ks, uint64_t nPoints, int reps, double* outMs);
int tblaeq_handler_batch_save_params(tblaeq_handler* h, const char* path);
int tblaeq_handler_batch_load_params(tblaeq_handler* h, const char* path);
/* tblaeq_batch_set_param on the handler's engine (TBLAEQ_PARAM_* of tblaeq.h). */
int tblaeq_handler_batch_set_param(tblaeq_handler* h, int param, uint64_t value);
int tblaeq_handler_batch_get_param(tblaeq_handler* h, int param, uint64_t* value);
int tblaeq_handler_batch_free(tblaeq_handler* h);

#ifdef __cplusplus
}
#endif
#endif
