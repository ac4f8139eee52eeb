// C entry points of the approximate-neighbours chapter (aoclda_approximate_neighbors.h),
// wrapped as in source/core/approximate_neighbors/approximate_neighbors_public.cpp:30-127 of
// the reference: null handle, log clearing, precision check, handle-type check.
#include "aoclda_approximate_neighbors.h"
#include "ivf.hpp"

using namespace dagpu;

template <typename T> static ApproxNNHandle<T> *ann_of(da_handle h, da_status &s) {
    s = h->check_precision<T>();
    if (s != da_status_success)
        return nullptr;
    auto *a = dynamic_cast<ApproxNNHandle<T> *>(h->alg.get());
    if (!a)
        s = h->err.error(da_status_invalid_handle_type,
                         "handle was not initialized with handle_type=da_handle_approx_nn or "
                         "handle is invalid.");
    return a;
}

#define DA_ANN_PROLOGUE(T)                                                                       \
    if (!handle)                                                                                 \
        return da_status_handle_not_initialized;                                                 \
    handle->err.clear();                                                                         \
    da_status s_;                                                                                \
    ApproxNNHandle<T> *a = ann_of<T>(handle, s_);                                                \
    if (!a)                                                                                      \
        return s_;

extern "C" {

da_status da_approx_nn_set_training_data_d(da_handle handle, da_int n_samples,
                                           da_int n_features, const double *X_train,
                                           da_int ldx_train) {
    DA_ANN_PROLOGUE(double)
    return a->set_training_data(n_samples, n_features, X_train, ldx_train);
}
da_status da_approx_nn_set_training_data_s(da_handle handle, da_int n_samples,
                                           da_int n_features, const float *X_train,
                                           da_int ldx_train) {
    DA_ANN_PROLOGUE(float)
    return a->set_training_data(n_samples, n_features, X_train, ldx_train);
}
da_status da_approx_nn_train_d(da_handle handle) {
    DA_ANN_PROLOGUE(double)
    return a->train();
}
da_status da_approx_nn_train_s(da_handle handle) {
    DA_ANN_PROLOGUE(float)
    return a->train();
}
da_status da_approx_nn_add_d(da_handle handle, da_int n_samples_add, da_int n_features,
                             const double *X_add, da_int ldx_add) {
    DA_ANN_PROLOGUE(double)
    return a->add(n_samples_add, n_features, X_add, ldx_add);
}
da_status da_approx_nn_add_s(da_handle handle, da_int n_samples_add, da_int n_features,
                             const float *X_add, da_int ldx_add) {
    DA_ANN_PROLOGUE(float)
    return a->add(n_samples_add, n_features, X_add, ldx_add);
}
da_status da_approx_nn_train_and_add_d(da_handle handle) {
    DA_ANN_PROLOGUE(double)
    return a->train_and_add();
}
da_status da_approx_nn_train_and_add_s(da_handle handle) {
    DA_ANN_PROLOGUE(float)
    return a->train_and_add();
}
da_status da_approx_nn_kneighbors_d(da_handle handle, da_int n_queries, da_int n_features,
                                    const double *X_test, da_int ldx_test, da_int *n_ind,
                                    double *n_dist, da_int k, da_int return_distance) {
    DA_ANN_PROLOGUE(double)
    return a->kneighbors(n_queries, n_features, X_test, ldx_test, n_ind, n_dist, k,
                         return_distance);
}
da_status da_approx_nn_kneighbors_s(da_handle handle, da_int n_queries, da_int n_features,
                                    const float *X_test, da_int ldx_test, da_int *n_ind,
                                    float *n_dist, da_int k, da_int return_distance) {
    DA_ANN_PROLOGUE(float)
    return a->kneighbors(n_queries, n_features, X_test, ldx_test, n_ind, n_dist, k,
                         return_distance);
}

} // extern "C"
