/*
 * Approximate nearest neighbours (IVF-Flat): the entry points of
 * source/include/aoclda_approximate_neighbors.h of the reference (:64-191), same names,
 * signatures and semantics. Options: "n_list", "n_probe", "k-means_iter", "seed",
 * "train fraction", "number of neighbors", "algorithm" (ivfflat | auto), "metric"
 * (euclidean | sqeuclidean; inner product and cosine return da_status_not_implemented),
 * "storage order". Results: da_rinfo (n_list, n_index, n_features, k-means iterations),
 * da_approx_nn_cluster_centroids, da_approx_nn_list_sizes.
 */
#ifndef AOCLDA_APPROXIMATE_NEIGHBORS_H
#define AOCLDA_APPROXIMATE_NEIGHBORS_H

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces da_approx_nn_set_training_data_? (aoclda_approximate_neighbors.h:64-70): the
 * pointer is kept until training, as in the reference. */
da_status da_approx_nn_set_training_data_d(da_handle handle, da_int n_samples,
                                           da_int n_features, const double *X_train,
                                           da_int ldx_train);
da_status da_approx_nn_set_training_data_s(da_handle handle, da_int n_samples,
                                           da_int n_features, const float *X_train,
                                           da_int ldx_train);

/* Replaces da_approx_nn_train_? (:98-100): k-means centroids, trained on the GPU. */
da_status da_approx_nn_train_d(da_handle handle);
da_status da_approx_nn_train_s(da_handle handle);

/* Replaces da_approx_nn_add_? (:127-132): vectors join the list of their nearest centroid;
 * they are copied into GPU memory. */
da_status da_approx_nn_add_d(da_handle handle, da_int n_samples_add, da_int n_features,
                             const double *X_add, da_int ldx_add);
da_status da_approx_nn_add_s(da_handle handle, da_int n_samples_add, da_int n_features,
                             const float *X_add, da_int ldx_add);

/* Replaces da_approx_nn_train_and_add_? (:152-154). */
da_status da_approx_nn_train_and_add_d(da_handle handle);
da_status da_approx_nn_train_and_add_s(da_handle handle);

/* Replaces da_approx_nn_kneighbors_? (:184-191): the k nearest vectors among the n_probe
 * lists nearest to each query, n_queries x k in the handle's storage order. */
da_status da_approx_nn_kneighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *X_test, da_int ldx_test, da_int *n_ind,
                                    double *n_dist, da_int k, da_int return_distance);
da_status da_approx_nn_kneighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *X_test, da_int ldx_test, da_int *n_ind,
                                    float *n_dist, da_int k, da_int return_distance);

#ifdef __cplusplus
}
#endif

#endif
