/*
 * GPU index build: k-means levels and the CSC P-tensor of each level.
 *
 * Replaces the reference's GPU build path, cuVS k-means (CUDAKmeans.cu:7-195,
 * called from IndexBuilder.cu:44-151 when forceUseCPU is false), and the host
 * P-tensor / map / radius construction (Setup.cuh:87-171, func.hpp:124-137).
 *
 * tblaeq_kmeans follows the CPU build's k-means (CPUKmeans.cuh:67-205):
 *   - k-means++ seeding from mt19937_64(seed): the first center is
 *     uniform_int(0, n-1), each further one is the first point whose
 *     cumulative min squared distance reaches uniform_real(0, total); the
 *     random draws are the same stream, the cumulative sums are a parallel
 *     scan (a pick can differ from the sequential sum only when the threshold
 *     falls within rounding of a boundary);
 *   - Lloyd iterations until no label changes or max_iter: assignment by the
 *     smallest squared distance, strict < (lowest index on ties).  The
 *     distance ranking is computed on the FP32 matrix cores as
 *     ||c||^2 - 2 x.c (x, c centred on the data's mid-range), and every point
 *     whose best and second-best FP32 scores are within twice a proven error
 *     bound is re-scored exactly in FP64 with the CPU build's arithmetic, so
 *     the labels equal the CPU build's;
 *   - centroid update: per-cluster FP64 sums in point order divided by the
 *     count (bit-identical to the CPU build's sums); an empty cluster keeps
 *     its centroid.
 *
 * Device pointers, stream-ordered on ctx; the call synchronises once per Lloyd
 * iteration (the stopping test).  D <= 64.
 */
#ifndef TBLAEQ_BUILD_H
#define TBLAEQ_BUILD_H

#include <stdint.h>

#include "tblaeq.h"

#ifdef __cplusplus
extern "C" {
#endif

/* d_x [n x D] FP64 (AoS) -> d_centers [k x D], d_labels [n].  stats (host,
 * optional, 4 words): [0] Lloyd iterations run, [1] points re-scored in FP64
 * (all iterations), [2] FP32 matrix-core score evaluations (points x centers,
 * all iterations), [3] milliseconds in the assignment kernels. */
int tblaeq_kmeans(tblaeq_ctx* ctx, const double* d_x, uint64_t n, uint64_t D, uint64_t k, uint64_t max_iter,
                  uint64_t seed, double* d_centers, uint64_t* d_labels, uint64_t* stats);

/* Options of tblaeq_kmeans_ex and of the builds that take them.
 * seeding 0 is the CPU build's k-means++ (above).  seeding 1 is k-means||
 * (Bahmani et al., 2012; no reference counterpart: the reference's CPU build
 * seeds with k-means++, CPUKmeans.cuh:67-118, its GPU path with cuVS's
 * k-means++ initialisation, CUDAKmeans.cu:77): the first centre as
 * in k-means++, then `rounds` rounds in which every point is drawn
 * independently with probability oversampling * k * d^2 / phi (d^2 its
 * squared distance to the nearest candidate so far, phi their sum; the draws
 * come from SplitMix64 of (seed, round, point), so they do not depend on the
 * launch), each round's draws appended in point order; every point's nearest
 * candidate is tracked (FP32 matrix-core ranking, FP64 re-score of near ties),
 * the candidates are weighted by the number of points nearest to them, and
 * weighted k-means++ (the same random stream as seeding 0) picks the k
 * centres among them.  With no more than k candidates it falls back to
 * seeding 0.  Its centres are not the CPU build's. */
typedef struct {
    uint64_t max_iter;    /* Lloyd iterations at most (default 7, the CPU build's) */
    uint64_t seed;        /* the random stream (default 42, the CPU build's) */
    int seeding;          /* 0 = k-means++, 1 = k-means|| (default 0) */
    double oversampling;  /* k-means||: expected draws per round, times k (default 0.5) */
    uint64_t rounds;      /* k-means|| rounds (default 4) */
} tblaeq_kmeans_options;
void tblaeq_kmeans_options_init(tblaeq_kmeans_options* opt);

/* tblaeq_kmeans with options; stats (host, optional, 6 words): [0..3] as
 * tblaeq_kmeans, [4] the inertia (sum over points of the squared FP64
 * distance to their centroid, as the bits of a double: per 1024-point tile,
 * then over the tiles in a fixed order), [5] k-means|| candidates (0 when
 * seeded by k-means++).  Sends progress messages (tblaeq_set_progress) for the
 * seeding rounds and picks and the Lloyd iterations of large problems. */
int tblaeq_kmeans_ex(tblaeq_ctx* ctx, const double* d_x, uint64_t n, uint64_t D, uint64_t k,
                     const tblaeq_kmeans_options* opt, double* d_centers, uint64_t* d_labels, uint64_t* stats);

/* d_out[i] = (double)d_in[i] (exact) on the context stream: FP32 datasets enter
 * the FP64 build this way. */
int tblaeq_convert_f32_f64(tblaeq_ctx* ctx, const float* d_in, uint64_t n, double* d_out);

/* tblaeq_kmeans on host arrays (x [n x D] in; centers [k x D], labels [n] out),
 * through device buffers of the call. */
int tblaeq_kmeans_host(tblaeq_ctx* ctx, const double* x, uint64_t n, uint64_t D, uint64_t k, uint64_t max_iter,
                       uint64_t seed, double* centers, uint64_t* labels, uint64_t* stats);

/* CSC P-tensor of one level from its fine points, centroids and labels
 * (Setup.cuh:87-146): columns = clusters, rows grouped by cluster in point
 * order (a stable sort of (label, point) keys; the CPU build's std::sort may
 * order the rows of a column differently), d_map[pos] = point index of row
 * pos, d_vals[pos*D + d] = fine[map[pos]][d] / centers[col][d], and
 * d_radius[c] = max over the cluster's points of sqrt(sum_d (c_d - f_d)^2)
 * (Setup.cuh:148-171).  d_colRes [k+1], d_nnzPerCol [k], d_map [n],
 * d_vals [n x D], d_radius [k]. */
int tblaeq_build_level(tblaeq_ctx* ctx, const double* d_fine, uint64_t n, uint64_t D, const double* d_centers,
                       uint64_t k, const uint64_t* d_labels, uint64_t* d_colRes, uint64_t* d_nnzPerCol,
                       uint64_t* d_map, double* d_vals, double* d_radius);

/* QueryHandler(forceCPU, PointCloud, name, height, ratios) (QueryHandler.cu:79-99,
 * IndexBuilder.cu:44-151): k-means build of an in-memory dataset (data: N*D
 * doubles, AoS).  forceCPU = 0 builds on the GPU (tblaeq_kmeans and
 * tblaeq_build_level per level, k-means++ from seed 42, at most 7 Lloyd
 * iterations, as the CPU build); forceCPU = 1 is the host build. */
struct tblaeq_handler;
int tblaeq_handler_build(const double* data, uint64_t N, uint64_t D, uint64_t height, const uint64_t* ratios,
                         const char* name, int forceCPU, struct tblaeq_handler** out);
/* tblaeq_handler_build with a dataset of FP64 (dtype 0) or FP32 (dtype 1)
 * values and k-means options (NULL = the defaults above: 7 iterations, seed
 * 42, k-means++).  forceCPU = 1 runs the host build with max_iter and seed
 * (seeding 1 is GPU-only: TBLAEQ_ERR_UNSUPPORTED).  FP32 values are converted
 * to FP64 exactly (on the device for the GPU build); the index holds FP64.
 * No reference counterpart for the options (the reference fixes them). */
int tblaeq_handler_build_ex(const void* data, int dtype, uint64_t N, uint64_t D, uint64_t height,
                            const uint64_t* ratios, const char* name, int forceCPU, const tblaeq_kmeans_options* opt,
                            struct tblaeq_handler** out);
/* Per level of a build (finest first), six words: the tblaeq_kmeans_ex stats
 * (the host build fills [0] iterations and [4] inertia); *n = words available
 * (0 for indexes that were not built here); copies min(*n, cap). */
int tblaeq_handler_build_stats(struct tblaeq_handler* h, uint64_t* out, uint64_t cap, uint64_t* n);

#ifdef __cplusplus
}
#endif
#endif /* TBLAEQ_BUILD_H */
