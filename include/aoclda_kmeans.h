/*
 * k-means clustering (Lloyd and Elkan). Replaces source/include/aoclda_kmeans.h:56-178
 * of the reference. Semantic difference: da_kmeans_set_data_? copies the data matrix
 * into GPU memory (HBM); the reference stores the caller's pointer.
 */
#ifndef AOCLDA_KMEANS_H
#define AOCLDA_KMEANS_H

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* aoclda_kmeans.h:56-60 */
da_status da_kmeans_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                               const double *A, da_int lda);
da_status da_kmeans_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                               const float *A, da_int lda);

/* aoclda_kmeans.h:81-83 */
da_status da_kmeans_set_init_centres_d(da_handle handle, const double *C, da_int ldc);
da_status da_kmeans_set_init_centres_s(da_handle handle, const float *C, da_int ldc);

/* aoclda_kmeans.h:110-112 */
da_status da_kmeans_compute_d(da_handle handle);
da_status da_kmeans_compute_s(da_handle handle);

/* aoclda_kmeans.h:134-140 */
da_status da_kmeans_transform_d(da_handle handle, da_int m_samples, da_int m_features,
                                const double *X, da_int ldx, double *X_transform,
                                da_int ldx_transform);
da_status da_kmeans_transform_s(da_handle handle, da_int m_samples, da_int m_features,
                                const float *X, da_int ldx, float *X_transform,
                                da_int ldx_transform);

/* aoclda_kmeans.h:172-177 */
da_status da_kmeans_predict_d(da_handle handle, da_int k_samples, da_int k_features,
                              const double *Y, da_int ldy, da_int *Y_labels);
da_status da_kmeans_predict_s(da_handle handle, da_int k_samples, da_int k_features,
                              const float *Y, da_int ldy, da_int *Y_labels);

#ifdef __cplusplus
}
#endif

#endif
