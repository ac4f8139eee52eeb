/*
 * Result extraction. Enum values follow source/include/aoclda_result.h:43-111 and the
 * functions :139-146 of the reference. If *dim is too small it is overwritten with the
 * required size and da_status_invalid_array_dimension is returned.
 */
#ifndef AOCLDA_RESULT
#define AOCLDA_RESULT

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum da_result_ {
    da_rinfo = 1,
    da_trained = 2,
    da_linmod_coef = 101,
    da_linmod_dual_coef,
    da_pca_scores = 201,
    da_pca_variance,
    da_pca_principal_components,
    da_pca_total_variance,
    da_pca_column_means,
    da_pca_column_sdevs,
    da_pca_u,
    da_pca_sigma,
    da_pca_vt,
    da_kernel_pca_n_components,
    da_kernel_pca_eigenvalues,
    da_kernel_pca_eigenvectors,
    da_kernel_pca_scores,
    da_kernel_pca_dual_coef,
    da_kernel_pca_gamma,
    da_kernel_pca_X_fit,
    da_kmeans_cluster_centres = 501,
    da_kmeans_labels,
    da_dbscan_labels,
    da_dbscan_n_clusters,
    da_dbscan_n_core_samples,
    da_dbscan_core_sample_indices,
    da_nn_radius_neighbors_count = 601,
    da_nn_radius_neighbors_offsets,
    da_nn_radius_neighbors_indices,
    da_nn_radius_neighbors_distances,
    da_nn_radius_neighbors_indices_index,
    da_nn_radius_neighbors_distances_index,
    da_svm_n_support_vectors = 701,
    da_svm_n_support_vectors_per_class,
    da_svm_idx_support_vectors,
    da_svm_n_iterations,
    da_svm_support_vectors,
    da_svm_bias,
    da_svm_probaA,
    da_svm_probaB,
    da_svm_dual_coef,
    da_svm_lp_n_iterations,
    da_approx_nn_cluster_centroids = 801,
    da_approx_nn_list_sizes,
    da_cubic_spline_coefficients = 901,
    da_tsne_embedding = 1001,
};
typedef enum da_result_ da_result;

da_status da_handle_get_result_d(const da_handle handle, da_result query, da_int *dim,
                                 double *result);
da_status da_handle_get_result_s(const da_handle handle, da_result query, da_int *dim,
                                 float *result);
da_status da_handle_get_result_int(const da_handle handle, da_result query, da_int *dim,
                                   da_int *result);

#ifdef __cplusplus
}
#endif

#endif
