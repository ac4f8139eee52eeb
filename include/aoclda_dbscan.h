/*
 * DBSCAN clustering: the entry points of source/include/aoclda_dbscan.h of the reference
 * (:55-94), same names, signatures and semantics. Set the data with da_dbscan_set_data_?
 * (copied into GPU memory), the options "min samples", "eps", "metric", "power",
 * "algorithm", "leaf size" and "storage order", then call da_dbscan_compute_?. Results:
 * da_handle_get_result_int with da_dbscan_labels (n_samples; -1 is noise),
 * da_dbscan_core_sample_indices, da_dbscan_n_clusters, da_dbscan_n_core_samples, and
 * da_handle_get_result_? with da_rinfo (9 values: n_samples, n_features, lda, eps,
 * min samples, leaf size, power, n_core_samples, n_clusters).
 */
#ifndef AOCLDA_DBSCAN_H
#define AOCLDA_DBSCAN_H

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces da_dbscan_set_data_? (aoclda_dbscan.h:55-59). */
da_status da_dbscan_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                               const double *A, da_int lda);
da_status da_dbscan_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                               const float *A, da_int lda);

/* Replaces da_dbscan_compute_? (aoclda_dbscan.h:92-94). */
da_status da_dbscan_compute_d(da_handle handle);
da_status da_dbscan_compute_s(da_handle handle);

#ifdef __cplusplus
}
#endif

#endif
