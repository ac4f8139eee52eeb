// da_pairwise_distances_{d,s}: argument checks of
// source/core/metrics/pairwise_distances.cpp:79-114 and the metric switch of :36-75 of
// the reference, with the distance tiles computed on the GPU.
#include "../core/gpu.hpp"
#include "../kernels/launch.hpp"
#include "aoclda_metrics.h"
#include "pairwise_common.hpp"
#include <algorithm>
#include <cfloat>
#include <limits>
#include <mutex>

using namespace dagpu;

namespace {

// Process-wide state of the stateless distance entry points: one stream and cached
// device buffers, serialised by a mutex.
struct PairwiseState {
    std::mutex mu;
    GpuContext gpu;
    ErrorStack err;
    DeviceBuffer stage, xbuf, ybuf, dbuf, nx, ny;
};
PairwiseState &state() {
    static PairwiseState *s = new PairwiseState;
    return *s;
}

// Copies an m x k array (given order and leading dimension) to device row-major compact.
template <typename T>
da_status upload_rows(ErrorStack *err, hipStream_t st, da_order order, da_int m, da_int k,
                      const T *A, da_int lda, DeviceBuffer &dst, DeviceBuffer &stage) {
    const size_t bytes = (size_t)m * k * sizeof(T);
    DA_HIP(err, dst.reserve(bytes));
    if (order == row_major) {
        DA_HIP(err, copy_2d(dst.ptr(), k * sizeof(T), A, (size_t)lda * sizeof(T),
                                     k * sizeof(T), m, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, stage.reserve(bytes));
        DA_HIP(err, copy_2d(stage.ptr(), m * sizeof(T), A, (size_t)lda * sizeof(T),
                                     m * sizeof(T), k, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(stage.as<T>(), m, m, k, dst.as<T>(), st));
    }
    return da_status_success;
}

template <typename T>
da_status pairwise(da_order order, da_int m, da_int n, da_int k, const T *X, da_int ldx,
                   const T *Y, da_int ldy, T *D, da_int ldd, T p, da_metric metric) {
    da_status chk = pairwise_check<T>(order, m, n, k, X, ldx, Y, ldy, D, ldd, p, metric);
    if (chk != da_status_success)
        return chk;
    const int km = pairwise_kmetric<T>(metric, p);
    if (km < 0)
        return da_status_not_implemented;

    PairwiseState &S = state();
    std::lock_guard<std::mutex> lock(S.mu);
    ErrorStack *err = &S.err;
    err->clear();
    da_status st = S.gpu.init(err);
    if (st != da_status_success)
        return st;
    hipStream_t s = S.gpu.stream();

    const bool x_is_y = (Y == nullptr);
    if (x_is_y) {
        n = m;
        ldy = ldx;
    }
    st = upload_rows<T>(err, s, order, m, k, X, ldx, S.xbuf, S.stage);
    if (st != da_status_success)
        return st;
    const T *dX = S.xbuf.as<T>();
    const T *dY = dX;
    if (!x_is_y) {
        st = upload_rows<T>(err, s, order, n, k, Y, ldy, S.ybuf, S.stage);
        if (st != da_status_success)
            return st;
        dY = S.ybuf.as<T>();
    }

    double *dnx = nullptr, *dny = nullptr;
    if (pairwise_needs_norms(km)) {
        DA_HIP(err, S.nx.reserve((size_t)m * sizeof(double)));
        DA_HIP(err, launch_row_sqnorms<T>(dX, k, m, k, S.nx.as<double>(), s));
        dnx = S.nx.as<double>();
        if (x_is_y) {
            dny = dnx;
        } else {
            DA_HIP(err, S.ny.reserve((size_t)n * sizeof(double)));
            DA_HIP(err, launch_row_sqnorms<T>(dY, k, n, k, S.ny.as<double>(), s));
            dny = S.ny.as<double>();
        }
        if (km == KM_COSINE) {
            const double eps = std::numeric_limits<T>::epsilon();
            DA_HIP(err, launch_cosine_norms(dnx, m, eps, s));
            if (!x_is_y)
                DA_HIP(err, launch_cosine_norms(dny, n, eps, s));
        }
    }
    const int zero_diag = pairwise_zero_diag(km, x_is_y);

    // Row-major output is computed as rows of X against Y. Column-major output is the
    // transposed problem (rows of Y against X, written row by row), which keeps the device
    // stores contiguous; every metric here is symmetric in its two arguments.
    const bool colmaj = (order == column_major);
    const T *A = colmaj ? dY : dX;
    const T *B = colmaj ? dX : dY;
    const double *na = colmaj ? dny : dnx;
    const double *nb = colmaj ? dnx : dny;
    const da_int ma = colmaj ? n : m;
    const da_int nb_cols = colmaj ? m : n;

    // Output is produced in row chunks that fit a bounded device buffer.
    const size_t row_bytes = (size_t)nb_cols * sizeof(T);
    const size_t max_bytes = (size_t)1 << 30;
    da_int chunk = (da_int)std::max<size_t>(1, std::min<size_t>(ma, max_bytes / row_bytes));
    DA_HIP(err, S.dbuf.reserve((size_t)chunk * row_bytes));
    for (da_int r0 = 0; r0 < ma; r0 += chunk) {
        const da_int rows = std::min(chunk, ma - r0);
        {
            KernelScope ks(S.gpu, pairwise_mfma_metric(km) ? "pairwise_mfma" : "pairwise_direct");
            DA_HIP(err, pairwise_tile<T>(km, A + (size_t)r0 * k, B, rows, nb_cols, k,
                                         S.dbuf.as<T>(), nb_cols, na ? na + r0 : nullptr, nb,
                                         (double)p, zero_diag, (int)r0, s));
        }
        DA_HIP(err, copy_2d(D + (size_t)r0 * ldd, (size_t)ldd * sizeof(T),
                                     S.dbuf.ptr(), row_bytes, row_bytes, rows,
                                     hipMemcpyDeviceToHost, s));
    }
    DA_HIP(err, hipStreamSynchronize(s));
    return da_status_success;
}

} // namespace

extern "C" {

// Per-kernel timing of the da_pairwise_distances_? calls (their own stream), as
// da_amd_profile_* does for a handle: kernels "pairwise_mfma" and "pairwise_direct".
da_status da_amd_pairwise_profile_enable(da_int on) {
    PairwiseState &S = state();
    std::lock_guard<std::mutex> lock(S.mu);
    S.gpu.enable_profile(on != 0);
    return da_status_success;
}
da_status da_amd_pairwise_profile_reset(void) {
    PairwiseState &S = state();
    std::lock_guard<std::mutex> lock(S.mu);
    S.gpu.reset_profile();
    return da_status_success;
}
da_status da_amd_pairwise_profile_get(const char *kernel, double *total_ms, da_int *count) {
    if (!kernel || !total_ms || !count)
        return da_status_invalid_pointer;
    PairwiseState &S = state();
    std::lock_guard<std::mutex> lock(S.mu);
    double ms = 0;
    int c = 0;
    S.gpu.profile_get(kernel, ms, c);
    *total_ms = ms;
    *count = c;
    return da_status_success;
}

da_status da_pairwise_distances_d(da_order order, da_int m, da_int n, da_int k,
                                  const double *X, da_int ldx, const double *Y, da_int ldy,
                                  double *D, da_int ldd, double p, da_metric metric) {
    return pairwise<double>(order, m, n, k, X, ldx, Y, ldy, D, ldd, p, metric);
}

da_status da_pairwise_distances_s(da_order order, da_int m, da_int n, da_int k, const float *X,
                                  da_int ldx, const float *Y, da_int ldy, float *D, da_int ldd,
                                  float p, da_metric metric) {
    return pairwise<float>(order, m, n, k, X, ldx, Y, ldy, D, ldd, p, metric);
}

} // extern "C"
