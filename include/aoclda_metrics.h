/*
 * Pairwise distances. Replaces source/include/aoclda_metrics.h:44-106 of the reference
 * (same enum values, same signature, same argument checks as
 * source/core/metrics/pairwise_distances.cpp:79-114). X, Y and D are host arrays; the
 * distances are computed on the GPU.
 */
#ifndef AOCLDA_METRICS
#define AOCLDA_METRICS

#include "aoclda_error.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum da_metric_ {
    da_euclidean,
    da_l2 = da_euclidean,
    da_sqeuclidean,
    da_minkowski,
    da_manhattan,
    da_l1 = da_manhattan,
    da_cityblock = da_manhattan,
    da_cosine,
    da_euclidean_gemm,
    da_sqeuclidean_gemm,
    da_inner_product
};
typedef enum da_metric_ da_metric;

enum da_data_types_ { da_all_finite, da_allow_infinite, da_allow_NaN };
typedef enum da_data_types_ da_data_types;

/* aoclda_metrics.h:99-102 */
da_status da_pairwise_distances_d(da_order order, da_int m, da_int n, da_int k,
                                  const double *X, da_int ldx, const double *Y,
                                  da_int ldy, double *D, da_int ldd, double p,
                                  da_metric metric);
/* aoclda_metrics.h:104-106 */
da_status da_pairwise_distances_s(da_order order, da_int m, da_int n, da_int k,
                                  const float *X, da_int ldx, const float *Y, da_int ldy,
                                  float *D, da_int ldd, float p, da_metric metric);

#ifdef __cplusplus
}
#endif

#endif
