/*
 * MI355X extensions of the AOCL-DA C API. These entry points have no counterpart in the
 * reference; they exist for callers that keep data in GPU memory (HBM) and for
 * measurement. All pointers named d_* are device pointers (hipMalloc'ed or obtained from
 * da_amd_malloc). Work is enqueued on the handle's HIP stream; da_amd_synchronize waits
 * for it.
 */
#ifndef AOCLDA_AMD_H
#define AOCLDA_AMD_H

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_metrics.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory helpers (thin wrappers over hipMalloc/hipFree/hipMemcpy). */
da_status da_amd_malloc(size_t bytes, void **d_ptr);
da_status da_amd_free(void *d_ptr);
da_status da_amd_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
da_status da_amd_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);
da_status da_amd_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes);
da_status da_amd_device_synchronize(void);
/* Name of the GPU architecture the library runs on (e.g. "gfx950"). */
da_status da_amd_device_name(char *name, da_int len);

/* Waits for all work enqueued on the handle's stream. */
da_status da_amd_synchronize(da_handle handle);

/*
 * k-NN query with device-resident queries and outputs. X_test is row-major with leading
 * dimension ldx_test (>= n_features); n_ind and n_dist are row-major n_queries x k device
 * arrays. The call returns once the work is enqueued; results are valid after
 * da_amd_synchronize. n_dist may be NULL when return_distance is 0.
 */
da_status da_nn_kneighbors_device_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *d_X_test, da_int ldx_test,
                                    da_int *d_n_ind, float *d_n_dist, da_int k,
                                    da_int return_distance);
da_status da_nn_kneighbors_device_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *d_X_test, da_int ldx_test,
                                    da_int *d_n_ind, double *d_n_dist, da_int k,
                                    da_int return_distance);

/*
 * Pairwise distances that allocate nothing. No counterpart in the reference: this is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Arguments, checks, metrics and results are those
 * of da_pairwise_distances_s, with X, Y (may be NULL) and D in device memory; device memory
 * for compact row-major copies of the operands (column-major input, or leading dimensions
 * above k) and for row norms (cosine and the *_gemm metrics) comes from a fixed array of
 * 32 MiB per GPU, held until the process ends: inputs whose copies and norms exceed it
 * return da_status_invalid_input, host pointers return da_status_invalid_pointer. The work
 * runs on the null stream and the call returns when it is done. Calls on one GPU share the
 * array, so calls that overlap on one GPU (from several host threads) race.
 */
da_status da_pairwise_distances_device_noalloc_s(da_order order, da_int m, da_int n, da_int k,
                                                 const float *d_X, da_int ldx,
                                                 const float *d_Y, da_int ldy, float *d_D,
                                                 da_int ldd, float p, da_metric metric);

/*
 * Per-kernel timing. When enabled, every kernel launch of the handle is bracketed by HIP
 * events on the handle's stream; da_amd_profile_get returns the accumulated duration (ms)
 * and launch count of the kernel whose name is given (names as listed by
 * da_amd_profile_names). da_amd_profile_reset clears the counters.
 */
da_status da_amd_profile_enable(da_handle handle, da_int on);
da_status da_amd_profile_reset(da_handle handle);
da_status da_amd_profile_get(da_handle handle, const char *kernel, double *total_ms,
                             da_int *count);
da_status da_amd_profile_names(da_handle handle, char *names, da_int *lnames);

/*
 * Per-kernel timing of the stateless da_pairwise_distances_? calls (no counterpart in the
 * reference): HIP events around the distance tiles on their stream, kernels
 * "pairwise_mfma" (the fp64 matrix-core GEMM form) and "pairwise_direct".
 */
da_status da_amd_pairwise_profile_enable(da_int on);
da_status da_amd_pairwise_profile_reset(void);
da_status da_amd_pairwise_profile_get(const char *kernel, double *total_ms, da_int *count);

/*
 * Diagnostics of the last k-NN query of the handle: number of queries whose candidate
 * set from the first pass of the fused tile could not be certified exact (they are searched
 * again: a second seeded pass, then the exact fallback for any still uncertified), and the
 * candidates kept per query by that first pass.
 */
da_status da_nn_last_query_stats(da_handle handle, da_int *n_fallback, da_int *n_candidates);

/*
 * Diagnostics of the last da_kmeans_compute_? of a k-means handle: number of samples whose
 * two nearest centres were within the error bound of the MFMA cross term and were therefore
 * re-assigned with exact double distances (summed over the iterations). No counterpart in
 * the reference.
 */
da_status da_kmeans_last_compute_stats(da_handle handle, da_int *n_rechecked);

/*
 * k-NN serving loop over the resident training set of a kNN handle (da_nn_set_data_? done;
 * the handle must outlive the server and not be used by other calls while it runs: the
 * serving thread searches through it). Requests are
 * queued by da_nn_server_submit_? (the queries are copied, the call does not block) and
 * served by a worker thread that coalesces queued requests into batches of at most
 * max_batch_queries queries whose device buffers fit in max_device_bytes (admission by
 * memory; a request larger than a batch is served in slices). Two batches are in flight on
 * two HIP streams, so a batch is prepared and admitted while the previous one runs.
 * Every query's result is the exact k nearest neighbours, identical whether the request is
 * served alone or together with others. da_nn_server_wait_? blocks until a request is
 * complete and returns its n_queries x k results (row-major) and its latency (submission to
 * completion, milliseconds); a request's result goes to one waiter, and a second wait on
 * the same request id returns da_status_invalid_input.
 */
typedef struct _da_nn_server *da_nn_server;
da_status da_nn_server_create(da_handle handle, da_int max_batch_queries,
                              size_t max_device_bytes, da_nn_server *server);
da_status da_nn_server_submit_s(da_nn_server server, da_int n_queries, da_int n_features,
                                const float *X, da_int ldx, da_int k, da_int *request_id);
da_status da_nn_server_submit_d(da_nn_server server, da_int n_queries, da_int n_features,
                                const double *X, da_int ldx, da_int k, da_int *request_id);
da_status da_nn_server_wait_s(da_nn_server server, da_int request_id, da_int *n_ind,
                              float *n_dist, double *latency_ms);
da_status da_nn_server_wait_d(da_nn_server server, da_int request_id, da_int *n_ind,
                              double *n_dist, double *latency_ms);
/* stats[0..7] = completed requests, completed queries, batches, latency p50 (ms), latency
 * p99 (ms), mean queries per batch, seconds since creation, fallback queries. *n_stats is
 * the array size on entry (>= 8) and the number written on exit. */
da_status da_nn_server_stats(da_nn_server server, double *stats, da_int *n_stats);
void da_nn_server_destroy(da_nn_server *server);

/*
 * Batched Lloyd k-means over many independent instances whose data are resident in device
 * memory. No counterpart in the reference (its k-means handles one instance per handle,
 * kmeans.cpp:351-783); each instance runs the reference's Lloyd iteration with supplied
 * initial centres (convergence test, final relabel and inertia of kmeans.cpp:996-1124;
 * empty clusters go to the origin as with "empty clusters" = "ignore").
 *
 * Instance i has n[i] samples of d[i] <= 64 features and k[i] <= min(64, n[i]) clusters.
 * Device arrays, instances packed one after another: X holds the rows (row-major, d[i]
 * columns), C holds the initial centres (k[i] x d[i], row-major) on entry and the final
 * centres on return, labels receives n[i] labels per instance. Host arrays n_iter and
 * inertia receive the iteration count and inertia of each instance. cross_term selects the
 * MFMA operand type of the assignment: 0 = fp32, 1 = bf16, 2 = bf16x3 (each value as hi + lo
 * in bf16, three products); in every form samples whose two nearest centres are within the
 * error bound are re-checked in double, so the labels are exact.
 * The first call for a batch layout tunes the tile shape and captures the run in a HIP graph;
 * later calls with the same sizes, pointers, max_iter and tol replay it.
 */
da_status da_kmeans_batch_device_s(da_int n_inst, const da_int *n, const da_int *d,
                                   const da_int *k, const float *X, float *C, da_int *labels,
                                   da_int max_iter, float tol, da_int cross_term,
                                   da_int *n_iter, float *inertia);
/* Samples per assignment tile chosen by the tuner for the last batch run (0 if none). */
da_status da_kmeans_batch_tile_shape(da_int *samples_per_tile);
/* Releases the device buffers and graphs kept for batch layouts. */
da_status da_kmeans_batch_release(void);

#ifdef __cplusplus
}
#endif

#endif
