// C entry points of the kNN chapter (aoclda_nearest_neighbors.h) and the device-resident
// extensions (aoclda_amd.h). Wrapping as in
// source/core/nearest_neighbors/nearest_neighbors_public.cpp:35-93 of the reference:
// null handle, log clearing, precision check, then handle-type check.
#include "aoclda_amd.h"
#include "aoclda_nearest_neighbors.h"
#include "knn.hpp"

using namespace dagpu;

template <typename T> static NNHandle<T> *nn_of(da_handle h, da_status &s) {
    s = h->check_precision<T>();
    if (s != da_status_success)
        return nullptr;
    auto *nn = dynamic_cast<NNHandle<T> *>(h->alg.get());
    if (!nn)
        s = h->err.error(da_status_invalid_handle_type,
                         "handle was not initialized with handle_type=da_handle_nn or handle is "
                         "invalid.");
    return nn;
}

template <typename T>
static da_status set_data(da_handle h, da_int n_samples, da_int n_features, const T *X,
                          da_int ldx) {
    if (!h)
        return da_status_handle_not_initialized;
    h->err.clear();
    da_status s;
    NNHandle<T> *nn = nn_of<T>(h, s);
    if (!nn)
        return s;
    return nn->set_data(n_samples, n_features, X, ldx);
}

template <typename T>
static da_status kneighbors(da_handle h, da_int nq, da_int nf, const T *X, da_int ldx,
                            da_int *ind, T *dist, da_int k, da_int rd) {
    if (!h)
        return da_status_handle_not_initialized;
    h->err.clear();
    da_status s;
    NNHandle<T> *nn = nn_of<T>(h, s);
    if (!nn)
        return s;
    return nn->kneighbors(nq, nf, X, ldx, ind, dist, k, rd);
}

template <typename T>
static da_status kneighbors_dev(da_handle h, da_int nq, da_int nf, const T *X, da_int ldx,
                                da_int *ind, T *dist, da_int k, da_int rd) {
    if (!h)
        return da_status_handle_not_initialized;
    h->err.clear();
    da_status s;
    NNHandle<T> *nn = nn_of<T>(h, s);
    if (!nn)
        return s;
    return nn->kneighbors_device(nq, nf, X, ldx, ind, dist, k, rd);
}

template <typename T>
static da_status radius_search(da_handle h, da_int nq, da_int nf, const T *X, da_int ldx, T r,
                        da_int rd, da_int sort) {
    if (!h)
        return da_status_handle_not_initialized;
    h->err.clear();
    da_status s;
    NNHandle<T> *nn = nn_of<T>(h, s);
    if (!nn)
        return s;
    return nn->radius_neighbors(nq, nf, X, ldx, r, rd, sort);
}

extern "C" {

da_status da_nn_radius_neighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                                   const double *X_test, da_int ldx_test, double radius,
                                   da_int return_distance, da_int sort_results) {
    return radius_search<double>(handle, n_queries, n_features, X_test, ldx_test, radius,
                                 return_distance, sort_results);
}
da_status da_nn_radius_neighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                                   const float *X_test, da_int ldx_test, float radius,
                                   da_int return_distance, da_int sort_results) {
    return radius_search<float>(handle, n_queries, n_features, X_test, ldx_test, radius,
                                return_distance, sort_results);
}

da_status da_nn_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                           const double *X_train, da_int ldx_train) {
    return set_data<double>(handle, n_samples, n_features, X_train, ldx_train);
}
da_status da_nn_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                           const float *X_train, da_int ldx_train) {
    return set_data<float>(handle, n_samples, n_features, X_train, ldx_train);
}
da_status da_nn_kneighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                             const double *X_test, da_int ldx_test, da_int *n_ind,
                             double *n_dist, da_int k, da_int return_distance) {
    return kneighbors<double>(handle, n_queries, n_features, X_test, ldx_test, n_ind, n_dist, k,
                              return_distance);
}
da_status da_nn_kneighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                             const float *X_test, da_int ldx_test, da_int *n_ind, float *n_dist,
                             da_int k, da_int return_distance) {
    return kneighbors<float>(handle, n_queries, n_features, X_test, ldx_test, n_ind, n_dist, k,
                             return_distance);
}
da_status da_nn_kneighbors_device_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *d_X_test, da_int ldx_test, da_int *d_n_ind,
                                    float *d_n_dist, da_int k, da_int return_distance) {
    return kneighbors_dev<float>(handle, n_queries, n_features, d_X_test, ldx_test, d_n_ind,
                                 d_n_dist, k, return_distance);
}
da_status da_nn_kneighbors_device_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *d_X_test, da_int ldx_test, da_int *d_n_ind,
                                    double *d_n_dist, da_int k, da_int return_distance) {
    return kneighbors_dev<double>(handle, n_queries, n_features, d_X_test, ldx_test, d_n_ind,
                                  d_n_dist, k, return_distance);
}

da_status da_nn_last_query_stats(da_handle handle, da_int *n_fallback, da_int *n_candidates) {
    if (!handle || !n_fallback || !n_candidates)
        return da_status_invalid_pointer;
    if (auto *a = dynamic_cast<NNHandle<float> *>(handle->alg.get())) {
        *n_fallback = a->last_fallback;
        *n_candidates = a->last_candidates;
        return da_status_success;
    }
    if (auto *a = dynamic_cast<NNHandle<double> *>(handle->alg.get())) {
        *n_fallback = a->last_fallback;
        *n_candidates = a->last_candidates;
        return da_status_success;
    }
    return da_status_invalid_handle_type;
}

} // extern "C"

// ---- classification and regression (nearest_neighbors_public.cpp of the reference) ----

#define DA_NN_PROLOGUE(T)                                                                        \
    if (!handle)                                                                                 \
        return da_status_handle_not_initialized;                                                 \
    handle->err.clear();                                                                         \
    da_status s_;                                                                                \
    NNHandle<T> *nn = nn_of<T>(handle, s_);                                                      \
    if (!nn)                                                                                     \
        return s_;

extern "C" {

da_status da_nn_set_labels_s(da_handle handle, da_int n_samples, const da_int *y_train) {
    DA_NN_PROLOGUE(float)
    return nn->set_labels(n_samples, y_train);
}
da_status da_nn_set_labels_d(da_handle handle, da_int n_samples, const da_int *y_train) {
    DA_NN_PROLOGUE(double)
    return nn->set_labels(n_samples, y_train);
}
da_status da_nn_set_targets_d(da_handle handle, da_int n_samples, const double *y_train) {
    DA_NN_PROLOGUE(double)
    return nn->set_targets(n_samples, y_train);
}
da_status da_nn_set_targets_s(da_handle handle, da_int n_samples, const float *y_train) {
    DA_NN_PROLOGUE(float)
    return nn->set_targets(n_samples, y_train);
}
da_status da_nn_classes_d(da_handle handle, da_int *n_classes, da_int *classes) {
    DA_NN_PROLOGUE(double)
    return nn->classes(n_classes, classes);
}
da_status da_nn_classes_s(da_handle handle, da_int *n_classes, da_int *classes) {
    DA_NN_PROLOGUE(float)
    return nn->classes(n_classes, classes);
}
da_status da_nn_classifier_predict_proba_d(da_handle handle, da_int n_queries,
                                           da_int n_features, const double *X_test,
                                           da_int ldx_test, double *proba,
                                           da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(double)
    return nn->predict_proba(n_queries, n_features, X_test, ldx_test, proba, search_mode);
}
da_status da_nn_classifier_predict_proba_s(da_handle handle, da_int n_queries,
                                           da_int n_features, const float *X_test,
                                           da_int ldx_test, float *proba,
                                           da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(float)
    return nn->predict_proba(n_queries, n_features, X_test, ldx_test, proba, search_mode);
}
da_status da_nn_classifier_predict_d(da_handle handle, da_int n_queries, da_int n_features,
                                     const double *X_test, da_int ldx_test, da_int *y_test,
                                     da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(double)
    return nn->predict(n_queries, n_features, X_test, ldx_test, y_test, search_mode);
}
da_status da_nn_classifier_predict_s(da_handle handle, da_int n_queries, da_int n_features,
                                     const float *X_test, da_int ldx_test, da_int *y_test,
                                     da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(float)
    return nn->predict(n_queries, n_features, X_test, ldx_test, y_test, search_mode);
}
da_status da_nn_regressor_predict_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *X_test, da_int ldx_test, double *y_test,
                                    da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(double)
    return nn->predict(n_queries, n_features, X_test, ldx_test, y_test, search_mode);
}
da_status da_nn_regressor_predict_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *X_test, da_int ldx_test, float *y_test,
                                    da_nn_search_mode search_mode) {
    DA_NN_PROLOGUE(float)
    return nn->predict(n_queries, n_features, X_test, ldx_test, y_test, search_mode);
}

} // extern "C"
