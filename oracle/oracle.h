/*
 * CPU oracle for the T-BLAEQ hierarchical multigrid query path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may use it, and only as the checker (or, for the baseline, as
 * the thing timed on the host cores).
 *
 * It restates the reference's algorithms in plain C++ (file:line cited per
 * function in oracle.cpp) so that the HIP path can be compared bit-for-bit:
 *   - synthetic hierarchy generator      src/Kmeans/RandomKmeans.cuh:176-250
 *   - P-tensor / map / radius build       src/Setup/Setup.cuh:87-171, func.hpp:124-137,202-213
 *   - build plan truncation               src/core/QueryHandler.cu:23-45
 *   - on-disk index format                src/core/IndexSerializer.cu:11-259
 *   - per-level traversal (L0 semantics)  src/core/strategies/L0Strategy.cu:50-191
 *       range prune                       src/Query/RangePruning.cu:5-25
 *       KNN lower bound + STEP            src/Query/KNN.cu:9-24,75-96
 *       order-preserving compaction       src/Query/GridCompact.cuh:41-100
 *       SpTSpM prolongation               src/Kernel/SpMSpV.cuh:47-67, SpMSpV.cu:200-254
 *       refactor                          src/Query/Refactor.cu:10-16
 *   - greedy memory-policy scheduler      src/core/MemoryPolicy.cu:100-177
 *   - Lloyd k-means with k-means++        src/Kmeans/CPUKmeans.cuh:67-210
 * plus an exact refine (range filter, KNN top-K ordered by (squared distance,
 * id)) and a brute-force scan over the index-reconstructed finest coordinates.
 *
 * Parity pins (each checked in tests/test_oracle.py):
 *   generator/build  -> byte layout + closed-form invariants (ratios times
 *                       parents reproduce children within 1 ulp-ish, radius
 *                       equals max child distance, map is a stable-by-label
 *                       permutation) and committed golden vectors
 *   range prune      -> brute force on tiny inputs, r=0 reduces to point-in-box
 *   sqrt_rd          -> hardware sqrt under fesetround(FE_DOWNWARD)
 *   STEP             -> literal sequential restatement of KNN.cu:75-96 vs the
 *                       closed-form selection used here
 *   traversal        -> with all radii set to +inf the candidate set is the
 *                       whole dataset; last-level candidates contain every
 *                       brute-force hit whose ancestors survive
 *   scheduler        -> hand-computed byte formulas
 *   k-means          -> inertia non-increasing, fixed point of Lloyd step
 * Parity unpinned: tie order of exactly equal KNN lower bounds (the reference
 * sorts with an unstable bitonic/merge network; here ties break by point id),
 * and whether nvcc contracted the reference's distance loop into FMA (here the
 * loop is an explicit fma chain, on the CPU and on the GPU).
 */
#ifndef TBLAEQ_ORACLE_H
#define TBLAEQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OIndex OIndex;
typedef struct OResult OResult;

/* --- index construction ------------------------------------------------ */
OIndex* oracle_gen_index(uint64_t N, uint64_t D, double vmin, double vmax, int isInt, const uint64_t* ratios,
                         uint64_t nratios, uint64_t seed, double sigmaDivisor);
/* Build from a dataset with the CPU k-means (kmeans++ seeded mt19937_64(42)). */
OIndex* oracle_build_index(const double* data, uint64_t N, uint64_t D, const uint64_t* ratios, uint64_t nratios,
                           uint64_t maxIter);
OIndex* oracle_load_index(const char* dir);
int oracle_save_index(const OIndex* idx, const char* dir);
void oracle_free_index(OIndex* idx);

/* meta[0]=D meta[1]=N meta[2]=height meta[3]=intervals; meshSizes (finest first) copied to ms (cap entries). */
uint64_t oracle_index_meta(const OIndex* idx, uint64_t* meta, uint64_t* ms, uint64_t cap);
/* Copy the finest generated data layer (only for generator-built indexes). Returns N*D copied or 0. */
uint64_t oracle_index_data(const OIndex* idx, double* out, uint64_t cap);
/* Generator layers (finest first) and labels (layer i -> i+1), generator-built indexes only. */
uint64_t oracle_index_layer(const OIndex* idx, uint64_t i, double* out);
uint64_t oracle_index_labels(const OIndex* idx, uint64_t i, uint64_t* out);
/* Arrays of interval l: which = 0 colRes, 1 nnzPerCol, 2 rowIds, 3 map[l+1], 4 P values, 5 radius. */
uint64_t oracle_index_array(const OIndex* idx, uint64_t l, int which, void* out);
/* Radius of level l with a scale factor (used by tests to disable pruning with +inf). */
void oracle_index_scale_radius(OIndex* idx, double factor);

/* --- per-query traversal ----------------------------------------------- */
OResult* oracle_range(const OIndex* idx, const double* lo, const double* hi);
OResult* oracle_knn(const OIndex* idx, const double* q, uint64_t K);
void oracle_free_result(OResult* r);
uint64_t oracle_result_size(const OResult* r);
void oracle_result_copy(const OResult* r, uint64_t* ids, double* vals);
/* Per-level bookkeeping: cand[l] = grid size entering level l, surv[l] = survivors, out[l] = prolonged size. */
void oracle_result_levels(const OResult* r, uint64_t* cand, uint64_t* surv, uint64_t* out);
/* Survivor ids (in grid order) of level l. Returns count, copies up to cap. */
uint64_t oracle_result_survivors(const OResult* r, uint64_t l, uint64_t* ids, uint64_t cap);

/* --- exact refine and brute force --------------------------------------- */
/* Range: ids of candidates inside [lo,hi] (inclusive), ascending. Returns count. */
uint64_t oracle_refine_range(const OResult* r, const double* lo, const double* hi, uint64_t* out, uint64_t cap);
/* KNN: top-K of the candidates by (squared distance, id). Returns count (<= K). */
uint64_t oracle_refine_knn(const OResult* r, const double* q, uint64_t K, uint64_t* ids, double* d2);
/* Reconstructed coordinates of every finest point, indexed by original id (N*D). */
void oracle_recon_finest(const OIndex* idx, double* out);
/* Reconstructed coordinates of level `level` (0 = coarsest, intervals = finest) by level id; returns the count. */
uint64_t oracle_recon_level(const OIndex* idx, uint64_t level, double* out);
uint64_t oracle_brute_range(const double* pts, uint64_t N, uint64_t D, const double* lo, const double* hi,
                            uint64_t* out, uint64_t cap);
uint64_t oracle_brute_knn(const double* pts, uint64_t N, uint64_t D, const double* q, uint64_t K, uint64_t* ids,
                          double* d2);

/* --- arithmetic primitives ---------------------------------------------- */
double oracle_sqrt_rd(double x);
double oracle_sqrt_rd_fenv(double x); /* hardware sqrt with the FPU in round-down mode (pin) */
double oracle_range_dist2(const double* c, const double* lo, const double* hi, uint64_t D);
double oracle_knn_lb(const double* c, const double* q, double r, uint64_t D);
/* KNN STEP over an already sorted list (literal sequential restatement of KNN.cu:75-96). */
uint64_t oracle_step_sequential(const double* lb_sorted, const uint64_t* w_sorted, const double* r_sorted,
                                uint64_t p, uint64_t K, uint8_t* sel);

/* --- memory-policy scheduler (reference byte formulas) ------------------ */
/* nnz[l], cols[l] per level; returns policies (0..3) into out. */
void oracle_policy_recommend(const uint64_t* nnz, const uint64_t* cols, uint64_t intervals, uint64_t D,
                             uint64_t coarsestBytes, uint64_t mapsBytes, uint64_t freeMem, uint8_t* out);

/* --- OpenMP CPU traversal baseline --------------------------------------- */
/* Runs nq queries (type 0 = range with boxes[nq*2D] as lo|hi rows, 1 = KNN with pts[nq*D])
 * over all host threads (nthreads<=0: omp default) with traversal + exact refine.
 * Writes per-query microseconds to us[nq] and the refined result count to cnt[nq].
 * Returns wall seconds of the whole batch; *threads_used receives the thread count. */
double oracle_cpu_batch(const OIndex* idx, int type, const double* qs, uint64_t nq, uint64_t K, int nthreads,
                        double* us, uint64_t* cnt, int* threads_used);

/* --- CPU k-means primitives (for the build tests) ------------------------ */
/* Full CPU k-means of CPUKmeans.cuh; labels[N], centroids[k*D]; returns iterations run. */
uint64_t oracle_kmeans(const double* data, uint64_t N, uint64_t D, uint64_t k, uint64_t maxIter, uint64_t* labels,
                       double* centroids);
uint64_t oracle_kmeans_seeded(const double* data, uint64_t N, uint64_t D, uint64_t k, uint64_t maxIter,
                              uint64_t streamSeed, uint64_t* labels, double* centroids);
/* Lloyd assignment (strict <, lowest index wins) and its inertia. */
double oracle_kmeans_assign(const double* data, uint64_t N, uint64_t D, const double* cent, uint64_t k,
                            uint64_t* labels);

#ifdef __cplusplus
}
#endif
#endif
