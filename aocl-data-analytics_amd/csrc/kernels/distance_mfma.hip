// GEMM-form distance tiles on the fp64 matrix cores of gfx950 (v_mfma_f64_16x16x4f64), for
// the metrics whose value is a function of the dot product x.y: squared euclidean and
// euclidean (|x|^2 + |y|^2 - 2 x.y), their GEMM-form variants, cosine (1 - x.y/(|x||y|)) and
// the inner product. Both input precisions are widened to double when staged, so the cross
// term carries fp64 accuracy whatever the data type.
//
// The reference computes euclidean and squared euclidean with direct differences
// (source/core/metrics/metrics_kernels.hpp:413-1070), which are accurate to a few ulps of the
// distance itself; the GEMM form is accurate to a few ulps of |x|^2 + |y|^2 instead. Each
// element therefore carries the rounding bound of its GEMM form,
//     E = 2.01 gamma_{k+2} (|x|^2 + |y|^2),   gamma_n = n u / (1 - n u),  u = 2^-53,
// and an element whose value does not exceed E / rtol (rtol = 5e-13 for double output,
// 5e-7 for float output) is recomputed with direct differences in double: the distances keep
// the direct form's relative accuracy, the diagonal of X against itself is exactly zero.
// The GEMM-form variants, cosine and the inner product are GEMM products in the reference
// too (euclidean_gemm_distance.cpp:59-230, cosine_distance.cpp:66-161,
// inner_product_distance.cpp) and are stored as computed.
//
// Work decomposition: a 256-thread workgroup owns a 64 x 64 output tile, stages 16-deep k
// slices of X and Y in LDS as double (transposed, as the direct tile does), and each wave
// accumulates a 32 x 32 quarter as 2 x 2 blocks of 16 x 16 on the matrix cores. Operand
// layout of v_mfma_f64_16x16x4f64 for lane l: A[l % 16][l / 16], B[l / 16][l % 16]; the
// accumulator's register r holds C[4 r + l / 16][l % 16].
#include "launch.hpp"
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

constexpr int TB = 64; // output tile edge
constexpr int KC = 16; // k slice staged in LDS

typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T, typename TY>
__device__ double direct_sqdist(const T *__restrict__ x, const TY *__restrict__ y, int k) {
    double acc = 0.0;
    for (int l = 0; l < k; l++) {
        const double d = (double)x[l] - (double)y[l];
        acc = fma(d, d, acc);
    }
    return acc;
}

template <typename T, typename TY, typename U, int METRIC>
__global__ __launch_bounds__(256) void pairwise_mfma_kernel(
    const T *__restrict__ X, const TY *__restrict__ Y, int m, int n, int k, U *__restrict__ D,
    int64_t ldd, const double *__restrict__ nx, const double *__restrict__ ny, int zero_diag,
    int diag_offset, double gam, double inv_rtol) {
    __shared__ double Xs[KC][TB + 1];
    __shared__ double Ys[KC][TB + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const int i0 = blockIdx.y * TB, j0 = blockIdx.x * TB;

    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
            acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

    // Loader mapping: 16 consecutive k per row, 16 rows per pass, 4 passes.
    const int lk = tid & 15, lr = tid >> 4;
    const int kq = lane >> 4, col = lane & 15;
    for (int k0 = 0; k0 < k; k0 += KC) {
        const int kk = k0 + lk;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int row = lr + 16 * s;
            const int gi = i0 + row, gj = j0 + row;
            double xv = 0.0, yv = 0.0;
            if (kk < k) {
                if (gi < m)
                    xv = (double)X[(int64_t)gi * k + kk];
                if (gj < n)
                    yv = (double)Y[(int64_t)gj * k + kk];
            }
            Xs[lk][row] = xv;
            Ys[lk][row] = yv;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KC; q += 4) {
            double a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                a[t] = Xs[q + kq][wr * 32 + t * 16 + col];
                b[t] = Ys[q + kq][wc * 32 + t * 16 + col];
            }
#pragma unroll
            for (int ra = 0; ra < 2; ra++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
                    acc[ra][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ra], b[cb], acc[ra][cb],
                                                                        0, 0, 0);
        }
        __syncthreads();
    }

#pragma unroll
    for (int ra = 0; ra < 2; ra++) {
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wr * 32 + ra * 16 + 4 * r + kq;
                const int j = j0 + wc * 32 + cb * 16 + col;
                if (i >= m || j >= n)
                    continue;
                const double g = acc[ra][cb][r];
                double v;
                if constexpr (METRIC == KM_SQEUCLIDEAN || METRIC == KM_EUCLIDEAN) {
                    const double s = nx[i] + ny[j];
                    v = s - 2.0 * g;
                    if (v <= 2.01 * gam * s * inv_rtol)
                        v = direct_sqdist(X + (int64_t)i * k, Y + (int64_t)j * k, k);
                    if constexpr (METRIC == KM_EUCLIDEAN)
                        v = sqrt(v);
                } else if constexpr (METRIC == KM_SQEUCLIDEAN_GEMM ||
                                     METRIC == KM_EUCLIDEAN_GEMM) {
                    v = nx[i] + ny[j] - 2.0 * g;
                    v = v < 0.0 ? 0.0 : v;
                    if constexpr (METRIC == KM_EUCLIDEAN_GEMM)
                        v = sqrt(v);
                } else if constexpr (METRIC == KM_COSINE) {
                    v = 1.0 - g / (nx[i] * ny[j]);
                } else {
                    v = g;
                }
                if (zero_diag && i + diag_offset == j)
                    v = 0.0;
                D[(int64_t)i * ldd + j] = (U)v;
            }
        }
    }
}

} // namespace

bool pairwise_mfma_metric(int metric) {
    return metric == KM_SQEUCLIDEAN || metric == KM_EUCLIDEAN || metric == KM_COSINE ||
           metric == KM_INNER || metric == KM_SQEUCLIDEAN_GEMM || metric == KM_EUCLIDEAN_GEMM;
}

template <typename T, typename U>
hipError_t launch_pairwise_mfma(int metric, const T *X, const T *Y, int m, int n, int k, U *D,
                                int64_t ldd, const double *nx, const double *ny, int zero_diag,
                                int diag_offset, hipStream_t s) {
    if (!pairwise_mfma_metric(metric) || !nx || !ny)
        return hipErrorInvalidValue;
    if (m <= 0 || n <= 0)
        return hipSuccess;
    dim3 grid((n + TB - 1) / TB, (m + TB - 1) / TB);
    const double u = 1.0 / 9007199254740992.0; // 2^-53
    const double nk = (double)(k + 2);
    const double gam = nk * u / (1.0 - nk * u);
    const double inv_rtol = sizeof(U) == 8 ? 2e12 : 2e6;
#define DA_PWM_CASE(M)                                                                       \
    case M:                                                                                  \
        hipLaunchKernelGGL((pairwise_mfma_kernel<T, T, U, M>), grid, dim3(256), 0, s, X, Y, m, n, \
                           k, D, ldd, nx, ny, zero_diag, diag_offset, gam, inv_rtol);         \
        break;
    switch (metric) {
        DA_PWM_CASE(KM_SQEUCLIDEAN)
        DA_PWM_CASE(KM_EUCLIDEAN)
        DA_PWM_CASE(KM_COSINE)
        DA_PWM_CASE(KM_INNER)
        DA_PWM_CASE(KM_SQEUCLIDEAN_GEMM)
        DA_PWM_CASE(KM_EUCLIDEAN_GEMM)
    default:
        return hipErrorInvalidValue;
    }
#undef DA_PWM_CASE
    return hipGetLastError();
}

template <typename TX, typename TY>
hipError_t launch_inner_mfma(const TX *X, const TY *Y, int m, int n, int k, float *D,
                             int64_t ldd, hipStream_t s) {
    if (m <= 0 || n <= 0)
        return hipSuccess;
    // At most 65,535 row tiles per launch (the grid's y extent).
    const int chunk = 65535 * TB;
    for (int i0 = 0; i0 < m; i0 += chunk) {
        const int mm = m - i0 < chunk ? m - i0 : chunk;
        dim3 grid((n + TB - 1) / TB, (mm + TB - 1) / TB);
        hipLaunchKernelGGL((pairwise_mfma_kernel<TX, TY, float, KM_INNER>), grid, dim3(256), 0,
                           s, X + (int64_t)i0 * k, Y, mm, n, k, D + (int64_t)i0 * ldd, ldd,
                           nullptr, nullptr, 0, 0, 0.0, 0.0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}
template hipError_t launch_inner_mfma<double, float>(const double *, const float *, int, int,
                                                     int, float *, int64_t, hipStream_t);
template hipError_t launch_inner_mfma<double, double>(const double *, const double *, int, int,
                                                      int, float *, int64_t, hipStream_t);

template hipError_t launch_pairwise_mfma<float, float>(int, const float *, const float *, int,
                                                       int, int, float *, int64_t,
                                                       const double *, const double *, int, int,
                                                       hipStream_t);
template hipError_t launch_pairwise_mfma<double, double>(int, const double *, const double *,
                                                         int, int, int, double *, int64_t,
                                                         const double *, const double *, int,
                                                         int, hipStream_t);

} // namespace dagpu
