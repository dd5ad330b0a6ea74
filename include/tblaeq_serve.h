/*
 * Serving loop over a handler's index: continuous batching of range and KNN
 * requests on the batched engine.
 *
 * No reference counterpart: the reference serves one request at a time
 * (HTTPServer.cu:414-553 calls performSingleRangeQuery / performSingleKNNQuery
 * under a lock).  Here requests are submitted from any thread and return a
 * ticket at once.  One worker per request kind (range, KNN) owns a batched
 * engine on its own stream: while its batch runs on the GPU, new requests
 * queue; when the batch completes, the worker admits everything waiting (up to
 * max_batch and the HBM budget) into the next batch.  Range and KNN batches run
 * concurrently on the device.  Each request's result is the same as when it
 * runs alone (the batched engine's results do not depend on the batch).
 *
 * Admission by free HBM: at creation the server measures free device memory
 * (hipMemGetInfo) after its engines and buffers are allocated, and keeps
 * hbm_fraction of it as the in-flight budget; a request is admitted into a
 * batch only while the estimated bytes of the requests in flight (admitted,
 * not yet completed) stay within the budget, so requests wait in the queue
 * until memory frees up.  At least one request is always admitted so the
 * queue drains.
 *
 * Launch parameters: the server's engines take those of the handler's batched
 * engine (batch slot 0) when it has one, e.g. after
 * tblaeq_handler_batch_autotune on a sample of the expected requests;
 * otherwise the engine defaults.
 *
 * Destroy the server before the handler it was created from.
 */
#ifndef TBLAEQ_SERVE_H
#define TBLAEQ_SERVE_H

#include <stdint.h>

#include "tblaeq_handler.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tblaeq_server tblaeq_server;

typedef struct {
    uint64_t max_batch;          /* requests per batch and kind (default 4096) */
    uint64_t max_k;              /* largest K served by the batched engine (default and upper limit:
                                    256 - empty clusters - 1); requests with a larger K (any K is
                                    accepted) run through the per-query traversal on the KNN worker */
    uint64_t pair_cap;           /* (cluster, query) pairs per level, 0 = sized from free HBM (<= 2^28) */
    uint64_t hit_cap;            /* range hits per batch, 0 = sized from free HBM (<= 2^27); range
                                    batches are sized from the running mean of hits per request so
                                    that they fit, and a batch that overflows is split (into parts
                                    sized from its hit count) and rerun; after an overflow of
                                    either kind, the worker's batches start at the size that fit
                                    and grow by a quarter per batch that fits */
    double hbm_fraction;         /* share of free HBM used as the in-flight budget (default 0.5) */
    uint64_t range_request_bytes; /* admission estimate per range request (default 4 MiB) */
    uint64_t knn_request_bytes;   /* admission estimate per KNN request (default 64 KiB) */
    double max_wait_us;          /* a worker waits up to this long for a batch to fill (default 0) */
} tblaeq_server_config;

/* Fills cfg with the defaults. */
void tblaeq_server_config_init(tblaeq_server_config* cfg);
int tblaeq_server_create(tblaeq_handler* h, const tblaeq_server_config* cfg, tblaeq_server** out);
/* Finishes every admitted and queued request, then stops the workers. */
int tblaeq_server_destroy(tblaeq_server* s);
/* type 0 = range (q = lo[D] | hi[D]), 1 = KNN (q = point[D], K).  Thread-safe. */
int tblaeq_server_submit(tblaeq_server* s, int type, const double* q, uint64_t K, uint64_t* ticket);
/* Waits until the request is done; TBLAEQ_ERR_TIMEOUT after timeout_s seconds
 * (timeout_s < 0 waits without limit). */
int tblaeq_server_wait(tblaeq_server* s, uint64_t ticket, double timeout_s);
/* Result of a finished request: *n = number of ids (range: hits ascending;
 * KNN: top-K by (squared distance, id)); copies min(*n, cap) ids (and KNN
 * squared distances when d2 != NULL).  With release != 0 the ticket is
 * forgotten afterwards. */
int tblaeq_server_result(tblaeq_server* s, uint64_t ticket, uint64_t* ids, uint64_t cap, double* d2, uint64_t* n,
                         int release);
/* out[0] completed requests, [1] p50 latency us, [2] p99 latency us,
 * [3] batches, [4] mean requests per batch, [5] requests deferred by the HBM
 * budget (admissions postponed), [6] peak in-flight bytes, [7] budget bytes. */
int tblaeq_server_stats(tblaeq_server* s, double* out);
int tblaeq_server_reset_stats(tblaeq_server* s);
/* Per request kind (type 0 range, 1 KNN) since the last reset: out[0]
 * completed requests, [1] batches, [2] engine runs repeated after a capacity
 * overflow, [3] requests run through the per-query traversal, [4] worker busy
 * time in us (batch formation excluded), [5] running mean of hits per range
 * request (the batch-size estimate). */
int tblaeq_server_kind_stats(tblaeq_server* s, int type, double* out);

/* The reference's POST /query request and response (HTTPServer.cu:414-553)
 * over this server, without the HTTP transport: request is the JSON body
 * ("dataset_name", "query_type" KNN|RANGE (any case), "query_point" and "k"
 * (default 10) or "lower_bound" and "upper_bound"), checked as the reference
 * checks it.  The request is queued and admitted like any other.  By default
 * it runs the per-query traversal with the fine mesh kept, as the reference
 * does, and the response carries the reference's fields: "success",
 * "dataset_name", "query_type", "result_count", "result" (fine-mesh point ids
 * in traversal order), "memory_policy", "level_original_sizes",
 * "level_pruned_sizes" (points per level) and "query_runtime" ("query_time_us",
 * "query_time_ms", "total_time_us", "avg_time_us", "query_count").  With the
 * extra field "exact": true it runs on the batched engine instead: "result"
 * is the exact answer (range hits ascending; KNN top-K by (squared distance,
 * id) with "squared_distances"), "query_runtime" is the request's latency
 * through the server and "level_pruned_sizes" is absent.  "dataset_name" must
 * name the served index (its directory name up to the first '_').
 * *response receives a malloc'd NUL-terminated JSON text (release it with
 * tblaeq_free_string) and *http_status the status the reference's front end
 * would send: 200, 400 (a rejected request; {"success": false, "message"}),
 * 500 (the query failed), 503 (the server is stopping) or 504 (not finished
 * within timeout_s seconds; the request still completes and is then dropped;
 * timeout_s < 0 waits without limit).  Returns TBLAEQ_OK whenever a response
 * was produced.  Thread-safe. */
int tblaeq_server_query_json(tblaeq_server* s, const char* request, double timeout_s, char** response,
                             int* http_status);
void tblaeq_free_string(char* p);

#ifdef __cplusplus
}
#endif
#endif /* TBLAEQ_SERVE_H */
