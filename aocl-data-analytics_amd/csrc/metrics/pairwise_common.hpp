// Argument checks and metric switch shared by the da_pairwise_distances entry points:
// source/core/metrics/pairwise_distances.cpp:79-114 (checks, in the reference's order) and
// :36-75 (minkowski p == 2 -> euclidean, p == 1 -> manhattan).
#pragma once
#include "../kernels/launch.hpp"
#include "aoclda_metrics.h"

namespace dagpu {

template <typename T>
inline da_status pairwise_check(da_order order, da_int m, da_int n, da_int k, const T *X,
                                da_int ldx, const T *Y, da_int ldy, const T *D, da_int ldd,
                                T p, da_metric metric) {
    if (m < 1 || k < 1)
        return da_status_invalid_array_dimension;
    if (order == column_major && ldx < m)
        return da_status_invalid_leading_dimension;
    if (order == row_major && ldx < k)
        return da_status_invalid_leading_dimension;
    if (X == nullptr || D == nullptr)
        return da_status_invalid_pointer;
    if (Y != nullptr) {
        if (n < 1)
            return da_status_invalid_array_dimension;
        if (order == column_major) {
            if (ldy < n || ldd < m)
                return da_status_invalid_leading_dimension;
        } else {
            if (ldy < k || ldd < n)
                return da_status_invalid_leading_dimension;
        }
    } else {
        if (ldd < m)
            return da_status_invalid_leading_dimension;
    }
    if (metric == da_minkowski && p <= 0)
        return da_status_invalid_input;
    return da_status_success;
}

// The distance kernels' metric code, or -1 for a metric the build does not provide.
template <typename T> inline int pairwise_kmetric(da_metric metric, T p) {
    switch (metric) {
    case da_euclidean:
        return KM_EUCLIDEAN;
    case da_sqeuclidean:
        return KM_SQEUCLIDEAN;
    case da_manhattan:
        return KM_MANHATTAN;
    case da_cosine:
        return KM_COSINE;
    case da_minkowski:
        return (p == (T)2.0) ? KM_EUCLIDEAN : (p == (T)1.0) ? KM_MANHATTAN : KM_MINKOWSKI;
    case da_euclidean_gemm:
        return KM_EUCLIDEAN_GEMM;
    case da_sqeuclidean_gemm:
        return KM_SQEUCLIDEAN_GEMM;
    case da_inner_product:
        return KM_INNER;
    default:
        return -1;
    }
}

// The reference forces a zero diagonal for cosine (cosine_distance.cpp:57-61) and for the
// GEMM form of X against itself (euclidean_gemm_distance.cpp:196-226); for the direct forms
// the diagonal is exactly zero already.
inline bool pairwise_zero_diag(int km, bool x_is_y) {
    return x_is_y && (km == KM_COSINE || km == KM_EUCLIDEAN_GEMM || km == KM_SQEUCLIDEAN_GEMM);
}

// Row norms: squared norms for the GEMM-form tile (every metric that is a function of x.y),
// turned into cosine norms for cosine.
inline bool pairwise_needs_norms(int km) { return pairwise_mfma_metric(km); }

// The distance tile of a metric: the fp64 matrix-core GEMM form where the metric is a function
// of x.y, direct differences (manhattan, minkowski) otherwise.
template <typename T>
inline hipError_t pairwise_tile(int km, const T *A, const T *B, int m, int n, int k, T *D,
                                int64_t ldd, const double *na, const double *nb, double p,
                                int zero_diag, int diag_offset, hipStream_t s) {
    if (pairwise_mfma_metric(km))
        return launch_pairwise_mfma<T, T>(km, A, B, m, n, k, D, ldd, na, nb, zero_diag,
                                          diag_offset, s);
    return launch_pairwise_direct<T, T>(km, A, B, m, n, k, D, ldd, na, nb, p, zero_diag,
                                        diag_offset, s);
}

} // namespace dagpu
