// Neighbour votes for k-NN classification and regression on gfx950: one thread per query
// turns its neighbour list (k fixed, or a CSR row from a radius search) into class
// probabilities and a predicted label, or a predicted target.
//
// The arithmetic follows nearest_neighbors.cpp:1166-1280 (k-NN probabilities),
// :1280-1480 (radius probabilities), :1489-1600 (labels) and :1610-1910 (targets) of the
// reference, in the handle's type T and in the same order: per class, the weights of the
// neighbours in list order; the denominator as the sum of the class totals in class order;
// weights 1 (uniform) or 1/d with d <= epsilon(T) counted as 1 (get_weights, :1021-1033).
#include "launch.hpp"
#include <hip/hip_runtime.h>
#include <limits>

namespace dagpu {

namespace {

template <typename T> __device__ inline T nn_weight(const T *dist, int64_t t, int weights) {
    if (!weights)
        return (T)1;
    const T d = dist[t];
    return d <= std::numeric_limits<T>::epsilon() ? (T)1 : (T)1 / d;
}

template <typename T>
__global__ void nn_classify_kernel(int nq, const int *__restrict__ off, int k,
                                   const int *__restrict__ ind, const T *__restrict__ dist,
                                   int weights, const int *__restrict__ cls,
                                   const int *__restrict__ classes, int ncls, int outlier_mode,
                                   int fill_idx, int fallback_label, T *__restrict__ proba,
                                   int *__restrict__ pred, int qbase,
                                   int *__restrict__ first_empty) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const int64_t b = off ? off[q] : (int64_t)q * k;
    const int cnt = off ? off[q + 1] - off[q] : k;
    T *row = proba + (int64_t)q * ncls;
    for (int c = 0; c < ncls; c++)
        row[c] = (T)0;
    if (cnt == 0) {
        // No neighbour within the radius: the host learns the first such query ("none" makes
        // it an error, "manual" may warn); "manual" and "most frequent" set the probability
        // of one class (fill_idx, -1 for none of them).
        atomicMin(first_empty, qbase + q);
        if (outlier_mode != 0 && fill_idx >= 0)
            row[fill_idx] = (T)1;
    } else {
        for (int t = 0; t < cnt; t++)
            row[cls[ind[b + t]]] += nn_weight(dist, b + t, weights);
        T den = (T)0;
        for (int c = 0; c < ncls; c++)
            den += row[c];
        for (int c = 0; c < ncls; c++)
            row[c] = row[c] / den;
    }
    if (pred) {
        // First class of largest probability (cblas_iamax); a zero largest probability (an
        // outlier without a class) takes the fallback label.
        int best = 0;
        for (int c = 1; c < ncls; c++)
            if (row[c] > row[best])
                best = c;
        pred[q] = row[best] == (T)0 ? fallback_label : classes[best];
    }
}

template <typename T>
__global__ void nn_regress_kernel(int nq, const int *__restrict__ off, int k,
                                  const int *__restrict__ ind, const T *__restrict__ dist,
                                  int weights, const T *__restrict__ tgt, int outlier_mode,
                                  T fallback, T *__restrict__ out, int qbase,
                                  int *__restrict__ first_empty) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const int64_t b = off ? off[q] : (int64_t)q * k;
    const int cnt = off ? off[q + 1] - off[q] : k;
    if (cnt == 0) {
        atomicMin(first_empty, qbase + q);
        out[q] = outlier_mode == 0 ? (T)0 : fallback;
        return;
    }
    T num = (T)0, den = (T)0;
    for (int t = 0; t < cnt; t++) {
        const T w = nn_weight(dist, b + t, weights);
        num += tgt[ind[b + t]] * w;
        den += w;
    }
    out[q] = weights ? num / den : num / (T)cnt;
}

} // namespace

template <typename T>
hipError_t launch_nn_classify(int nq, const int *off, int k, const int *ind, const T *dist,
                              int weights, const int *cls, const int *classes, int ncls,
                              int outlier_mode, int fill_idx, int fallback_label, T *proba,
                              int *pred, int qbase, int *first_empty, hipStream_t s) {
    if (nq <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(nn_classify_kernel<T>, dim3((nq + 255) / 256), dim3(256), 0, s, nq, off,
                       k, ind, dist, weights, cls, classes, ncls, outlier_mode, fill_idx,
                       fallback_label, proba, pred, qbase, first_empty);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_nn_regress(int nq, const int *off, int k, const int *ind, const T *dist,
                             int weights, const T *tgt, int outlier_mode, T fallback, T *out,
                             int qbase, int *first_empty, hipStream_t s) {
    if (nq <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(nn_regress_kernel<T>, dim3((nq + 255) / 256), dim3(256), 0, s, nq, off, k,
                       ind, dist, weights, tgt, outlier_mode, fallback, out, qbase, first_empty);
    return hipGetLastError();
}

#define DA_VOTE_INST(T)                                                                          \
    template hipError_t launch_nn_classify<T>(int, const int *, int, const int *, const T *,     \
                                              int, const int *, const int *, int, int, int, int, \
                                              T *, int *, int, int *, hipStream_t);              \
    template hipError_t launch_nn_regress<T>(int, const int *, int, const int *, const T *, int, \
                                             const T *, int, T, T *, int, int *, hipStream_t);
DA_VOTE_INST(float)
DA_VOTE_INST(double)
#undef DA_VOTE_INST

} // namespace dagpu
