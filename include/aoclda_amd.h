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
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory helpers (thin wrappers over hipMalloc/hipFree/hipMemcpy). */
da_status da_amd_malloc(size_t bytes, void **d_ptr);
da_status da_amd_free(void *d_ptr);
da_status da_amd_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
da_status da_amd_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);
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
 * Diagnostics of the last k-NN query of the handle: number of queries whose candidate
 * set could not be certified exact and were recomputed by the exact fallback pass.
 */
da_status da_nn_last_query_stats(da_handle handle, da_int *n_fallback, da_int *n_candidates);

#ifdef __cplusplus
}
#endif

#endif
