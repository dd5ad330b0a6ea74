/*
 * Learned KNN radius proposal.
 *
 * No reference counterpart: the reference finds each query's KNN candidate
 * set by the STEP rule alone (KNN.cu:75-96).  Here a small MLP (48 -> 64 ->
 * 64 -> 4, ReLU, BF16 weights and activations on the MFMA cores with FP32
 * accumulation, FP32 master weights, Adam) proposes each query's K-th-NN
 * distance from the query point and coarse-level cluster bounds.  The batched
 * engine refines the surviving clusters within the proposed radius and
 * certifies the answer exactly (at least K candidates within the radius); a
 * query that fails certification falls back to the seed method, so results
 * never change.
 *
 * Features (48 per query, FP32): the query point as (x - vmin) / scale in 16
 * slots (zero-padded; the first 16 dimensions when D > 16), the 31 smallest
 * lb = |q - c| - r_c over the coarsest clusters divided by scale, ascending,
 * and log(K) / log(256).  Targets (4 per query): log(d_K / scale), then
 * log(1 + survivors at level l) / 10 for l = 0, 1, 2 (survivor budgets).
 */
#ifndef TBLAEQ_LEARN_H
#define TBLAEQ_LEARN_H

#include <stdint.h>

#include "tblaeq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tblaeq_mlp tblaeq_mlp;

/* Number of FP32 parameters: W1 [48 x 64] | b1 [64] | W2 [64 x 64] | b2 [64] |
 * W3 [64 x 16] (columns 4..15 zero) | b3 [16] (entries 4..15 zero). */
uint64_t tblaeq_mlp_param_count(void);
/* Network with its own context (stream) on ctx's device, He-uniform weights
 * from `seed`.  Independent of ctx's lifetime. */
int tblaeq_mlp_create(tblaeq_ctx* ctx, uint64_t seed, tblaeq_mlp** out);
int tblaeq_mlp_destroy(tblaeq_mlp* m);
/* Host FP32 parameters in the packed layout; set also resets Adam's moments. */
int tblaeq_mlp_set_params(tblaeq_mlp* m, const float* params);
int tblaeq_mlp_get_params(tblaeq_mlp* m, float* params);
/* Feature normalisation stored with the network (domain origin and extent). */
int tblaeq_mlp_set_norm(tblaeq_mlp* m, double vmin, double scale);
int tblaeq_mlp_get_norm(const tblaeq_mlp* m, double* vmin, double* scale);
/* Per-output weights of the squared-error loss (default 1, 1, 1, 1). */
int tblaeq_mlp_set_loss_weights(tblaeq_mlp* m, const float* w4);
/* Forward pass, host X [n x 48] -> host Y [n x 4]. */
int tblaeq_mlp_forward(tblaeq_mlp* m, const float* X, uint64_t n, float* Y);
/* Forward pass on device arrays: d_X [n x 48] -> d_Y16 [n x 16] (outputs in columns 0..3). */
int tblaeq_mlp_forward_device(tblaeq_mlp* m, const float* d_X, uint64_t n, float* d_Y16);
/* Gradients of the weighted mean squared error over n rows (host X [n x 48],
 * T [n x 4]) into grads [param_count] (host); *loss = the loss.  Weights
 * unchanged. */
int tblaeq_mlp_gradients(tblaeq_mlp* m, const float* X, const float* T, uint64_t n, float* grads, float* loss);
/* Adam (beta 0.9 / 0.999) over mini-batches of `batch` rows (multiple of 16)
 * in row order, `epochs` passes; loss_per_epoch (optional) = mean mini-batch
 * loss of each pass. */
int tblaeq_mlp_train(tblaeq_mlp* m, const float* X, const float* T, uint64_t n, uint64_t epochs, uint64_t batch,
                     float lr, float* loss_per_epoch);
/* Features of nq KNN queries (device points [nq x D], K [nq]) on the batched
 * engine's coarsest level into d_X [nq x 48]. */
int tblaeq_knn_features(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, double vmin,
                        double scale, float* d_X);
/* Proposed radii: d_radius[q] = scale * exp(y0(q)) * factor. */
int tblaeq_mlp_propose_radius(tblaeq_mlp* m, const float* d_X, uint64_t n, double scale, double factor,
                              double* d_radius);

/* tblaeq_batch_knn with a proposed search radius per query (device): the last
 * level refines the surviving clusters within d_radius[q] and no seeds are
 * expanded.  d_fallback[q] = 1 when fewer than K candidates lie within the
 * radius: that query's row is then not the answer and must be rerun with
 * tblaeq_batch_knn.  Other arguments as there; with one level or
 * TBLAEQ_PARAM_KNN_MODE = 1 the radii are ignored (seed method, no fallbacks). */
int tblaeq_batch_knn_radius(tblaeq_batch* b, const double* d_points, const uint64_t* d_K, uint64_t nq, uint64_t Kmax,
                            const double* d_radius, uint64_t* d_out_ids, double* d_out_d2, uint64_t* d_out_n,
                            uint8_t* d_fallback, uint64_t* d_status, uint64_t* d_level_surv,
                            uint64_t* d_fine_count);

/* Handler level (tblaeq_handler.h): KNN batch uploaded with
 * tblaeq_handler_batch_set_queries, run with radii proposed by `m` (times
 * factor), certified, with fallbacks rerun by the seed method; results as for
 * tblaeq_handler_batch_run.  *fallbacks = queries that fell back. */
struct tblaeq_handler;
int tblaeq_handler_batch_knn_learned(struct tblaeq_handler* h, tblaeq_mlp* m, double factor, int stats,
                                     uint64_t* fallbacks);
/* Per-query (plain) path with a learned radius per query (host radii [nq],
 * e.g. scale * exp(y0) * factor from tblaeq_mlp_forward): the last level
 * prolongs only the STEP survivors within the radius; the exact top-K is
 * certified when K results lie within the radius, else the radius grows by
 * 1.25 (at most three tries) before a full traversal.  Result as
 * tblaeq_handler_query with refine on; *fallbacks = retries (misses). */
struct tblaeq_result;
int tblaeq_handler_query_learned(struct tblaeq_handler* h, const double* points, uint64_t nq, uint64_t K,
                                 const double* radii, uint64_t* fallbacks, struct tblaeq_result** out);
/* Features of host KNN queries (points [n x D], K [n]) into host X [n x 48]. */
int tblaeq_handler_knn_features(struct tblaeq_handler* h, const double* points, const uint64_t* Ks, uint64_t n,
                                double vmin, double scale, float* X);

#ifdef __cplusplus
}
#endif
#endif /* TBLAEQ_LEARN_H */
