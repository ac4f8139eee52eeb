/*
 * k-nearest neighbours (brute force). Replaces the kNN entry points of
 * source/include/aoclda_nearest_neighbors.h:75-78 (set_data) and :144-149 (kneighbors)
 * of the reference. Semantic difference: da_nn_set_data_? copies the training data into
 * GPU memory (HBM) where it stays resident until the handle is destroyed or new data
 * is set; the reference stores the caller's pointer instead.
 */
#ifndef AOCLDA_NEAREST_NEIGHBORS_OPTIONS
#define AOCLDA_NEAREST_NEIGHBORS_OPTIONS

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum da_nn_search_mode_ { knn_search_mode = 0, radius_search_mode = 1 };
typedef enum da_nn_search_mode_ da_nn_search_mode;

/* aoclda_nearest_neighbors.h:75-78 */
da_status da_nn_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                           const double *X_train, da_int ldx_train);
da_status da_nn_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                           const float *X_train, da_int ldx_train);

/* aoclda_nearest_neighbors.h:144-149 */
da_status da_nn_kneighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                             const double *X_test, da_int ldx_test, da_int *n_ind,
                             double *n_dist, da_int k, da_int return_distance);
da_status da_nn_kneighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                             const float *X_test, da_int ldx_test, da_int *n_ind,
                             float *n_dist, da_int k, da_int return_distance);

/* aoclda_nearest_neighbors.h:188-193. Results through da_handle_get_result_?:
 * da_nn_radius_neighbors_count (n_queries + 1, total last), _offsets (n_queries + 1, -1 for
 * a query without neighbours, total last), _indices / _distances (all queries' neighbours
 * in query order), _indices_index / _distances_index (one query, its index in result[0]).
 * Neighbours of a query are in ascending training index, or ascending distance (ties by
 * index) with sort_results = 1, which requires return_distance = 1. radius < 0 takes the
 * "radius" option. */
da_status da_nn_radius_neighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                                   const double *X_test, da_int ldx_test, double radius,
                                   da_int return_distance, da_int sort_results);
da_status da_nn_radius_neighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                                   const float *X_test, da_int ldx_test, float radius,
                                   da_int return_distance, da_int sort_results);

#ifdef __cplusplus
}
#endif

#endif
