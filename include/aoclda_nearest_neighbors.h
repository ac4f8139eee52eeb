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

/* Classification labels / regression targets of the training rows, copied into GPU
 * memory (the reference keeps the pointer). Replace da_nn_set_labels_? and
 * da_nn_set_targets_? (aoclda_nearest_neighbors.h:96-97, :115-116). */
da_status da_nn_set_labels_s(da_handle handle, da_int n_samples, const da_int *y_train);
da_status da_nn_set_labels_d(da_handle handle, da_int n_samples, const da_int *y_train);
da_status da_nn_set_targets_d(da_handle handle, da_int n_samples, const double *y_train);
da_status da_nn_set_targets_s(da_handle handle, da_int n_samples, const float *y_train);

/* Distinct labels, ascending; *n_classes <= 0 queries their number. Replaces da_nn_classes_?
 * (aoclda_nearest_neighbors.h:214-215). */
da_status da_nn_classes_d(da_handle handle, da_int *n_classes, da_int *classes);
da_status da_nn_classes_s(da_handle handle, da_int *n_classes, da_int *classes);

/* Class probabilities (n_queries x n_classes in the handle's storage order), labels, or
 * regression targets of the queries from their k nearest neighbours (knn_search_mode, the
 * "number of neighbors" option) or their neighbours within the "radius" option
 * (radius_search_mode, queries without one handled by "outlier handling"); votes weighted by
 * the "weights" option. Replace da_nn_classifier_predict_proba_?, da_nn_classifier_predict_?
 * and da_nn_regressor_predict_? (aoclda_nearest_neighbors.h:249-256, :290-297, :330-335). */
da_status da_nn_classifier_predict_proba_d(da_handle handle, da_int n_queries,
                                           da_int n_features, const double *X_test,
                                           da_int ldx_test, double *proba,
                                           da_nn_search_mode search_mode);
da_status da_nn_classifier_predict_proba_s(da_handle handle, da_int n_queries,
                                           da_int n_features, const float *X_test,
                                           da_int ldx_test, float *proba,
                                           da_nn_search_mode search_mode);
da_status da_nn_classifier_predict_d(da_handle handle, da_int n_queries, da_int n_features,
                                     const double *X_test, da_int ldx_test, da_int *y_test,
                                     da_nn_search_mode search_mode);
da_status da_nn_classifier_predict_s(da_handle handle, da_int n_queries, da_int n_features,
                                     const float *X_test, da_int ldx_test, da_int *y_test,
                                     da_nn_search_mode search_mode);
da_status da_nn_regressor_predict_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *X_test, da_int ldx_test, double *y_test,
                                    da_nn_search_mode search_mode);
da_status da_nn_regressor_predict_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *X_test, da_int ldx_test, float *y_test,
                                    da_nn_search_mode search_mode);

#ifdef __cplusplus
}
#endif

#endif
