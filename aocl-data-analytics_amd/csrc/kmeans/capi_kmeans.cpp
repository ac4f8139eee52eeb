// C entry points of the k-means chapter (aoclda_kmeans.h), wrapped as in
// source/core/clustering/kmeans/kmeans_public.cpp of the reference: null handle, log
// clearing, precision check, handle-type check.
#include <algorithm>
#include <climits>
#include "aoclda_kmeans.h"
#include "kmeans.hpp"

using namespace dagpu;

template <typename T> static KMeansHandle<T> *km_of(da_handle h, da_status &s) {
    s = h->check_precision<T>();
    if (s != da_status_success)
        return nullptr;
    auto *km = dynamic_cast<KMeansHandle<T> *>(h->alg.get());
    if (!km)
        s = h->err.error(da_status_invalid_handle_type,
                         "handle was not initialized with handle_type=da_handle_kmeans or handle "
                         "is invalid.");
    return km;
}

#define DA_KM_PROLOGUE(T)                                                                        \
    if (!handle)                                                                                 \
        return da_status_handle_not_initialized;                                                 \
    handle->err.clear();                                                                         \
    da_status s_;                                                                                \
    KMeansHandle<T> *km = km_of<T>(handle, s_);                                                  \
    if (!km)                                                                                     \
        return s_;

extern "C" {

da_status da_kmeans_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                               const double *A, da_int lda) {
    DA_KM_PROLOGUE(double)
    return km->set_data(n_samples, n_features, A, lda);
}
da_status da_kmeans_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                               const float *A, da_int lda) {
    DA_KM_PROLOGUE(float)
    return km->set_data(n_samples, n_features, A, lda);
}
da_status da_kmeans_set_init_centres_d(da_handle handle, const double *C, da_int ldc) {
    DA_KM_PROLOGUE(double)
    return km->set_init_centres(C, ldc);
}
da_status da_kmeans_set_init_centres_s(da_handle handle, const float *C, da_int ldc) {
    DA_KM_PROLOGUE(float)
    return km->set_init_centres(C, ldc);
}
da_status da_kmeans_compute_d(da_handle handle) {
    DA_KM_PROLOGUE(double)
    return km->compute();
}
da_status da_kmeans_compute_s(da_handle handle) {
    DA_KM_PROLOGUE(float)
    return km->compute();
}
da_status da_kmeans_transform_d(da_handle handle, da_int m_samples, da_int m_features,
                                const double *X, da_int ldx, double *X_transform,
                                da_int ldx_transform) {
    DA_KM_PROLOGUE(double)
    return km->transform(m_samples, m_features, X, ldx, X_transform, ldx_transform);
}
da_status da_kmeans_transform_s(da_handle handle, da_int m_samples, da_int m_features,
                                const float *X, da_int ldx, float *X_transform,
                                da_int ldx_transform) {
    DA_KM_PROLOGUE(float)
    return km->transform(m_samples, m_features, X, ldx, X_transform, ldx_transform);
}
da_status da_kmeans_predict_d(da_handle handle, da_int k_samples, da_int k_features,
                              const double *Y, da_int ldy, da_int *Y_labels) {
    DA_KM_PROLOGUE(double)
    return km->predict(k_samples, k_features, Y, ldy, Y_labels);
}
da_status da_kmeans_predict_s(da_handle handle, da_int k_samples, da_int k_features,
                              const float *Y, da_int ldy, da_int *Y_labels) {
    DA_KM_PROLOGUE(float)
    return km->predict(k_samples, k_features, Y, ldy, Y_labels);
}

da_status da_kmeans_last_compute_stats(da_handle handle, da_int *n_rechecked) {
    if (!handle || !n_rechecked)
        return da_status_invalid_pointer;
    if (auto *a = dynamic_cast<KMeansHandle<float> *>(handle->alg.get())) {
        *n_rechecked = (da_int)std::min<long long>(a->rechecked_, INT32_MAX);
        return da_status_success;
    }
    if (auto *a = dynamic_cast<KMeansHandle<double> *>(handle->alg.get())) {
        *n_rechecked = (da_int)std::min<long long>(a->rechecked_, INT32_MAX);
        return da_status_success;
    }
    return da_status_invalid_handle_type;
}

} // extern "C"
