// Direct-form distance tiles and layout helpers for gfx950.
//
// The tile kernel follows the arithmetic of the reference's packed micro-kernels
// (source/core/metrics/metrics_kernels.hpp:413-1070: (x-y)^2, |x-y|, |x-y|^p and x*y
// accumulated per output element) and its post-operations (sqrt for euclidean,
// euclidean_distance.cpp:74-83; 1 - dot/(|x||y|) for cosine, cosine_distance.cpp:66-161;
// pow(., 1/p) for minkowski, minkowski_distance.cpp). The accumulation is done in double
// for both input precisions: on CDNA4 a double FMA issues at the same rate as an unpacked
// float FMA, so the fp32 path gets fp64 accuracy at no cost.
//
// Work decomposition: a 256-thread workgroup owns a 64x64 output tile, stages 64x16
// slices of X and Y in LDS (as double, transposed so that a thread's operands are a
// conflict-free broadcast), and each thread accumulates a 4x4 sub-tile with a stride of 16
// rows/columns so that the stores of a wave are contiguous along the output row.
#include "launch.hpp"
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

constexpr int TB = 64;  // output tile edge
constexpr int KC = 16;  // k slice staged in LDS
constexpr int TR = 4;   // per-thread rows/cols

template <int METRIC> __device__ __forceinline__ double accum(double acc, double x, double y, double p) {
    if constexpr (METRIC == KM_SQEUCLIDEAN || METRIC == KM_EUCLIDEAN) {
        double d = x - y;
        return fma(d, d, acc);
    } else if constexpr (METRIC == KM_MANHATTAN) {
        return acc + fabs(x - y);
    } else if constexpr (METRIC == KM_MINKOWSKI || METRIC == KM_MINKOWSKI_KEY) {
        return acc + pow(fabs(x - y), p);
    } else {
        return fma(x, y, acc);
    }
}

template <int METRIC>
__device__ __forceinline__ double finish(double acc, int i, int j, const double *nx,
                                         const double *ny, double p) {
    if constexpr (METRIC == KM_EUCLIDEAN) {
        return sqrt(acc);
    } else if constexpr (METRIC == KM_MINKOWSKI) {
        return pow(acc, 1.0 / p);
    } else if constexpr (METRIC == KM_COSINE) {
        return 1.0 - acc / (nx[i] * ny[j]);
    } else if constexpr (METRIC == KM_SQEUCLIDEAN_GEMM || METRIC == KM_EUCLIDEAN_GEMM) {
        double v = nx[i] + ny[j] - 2.0 * acc;
        v = v < 0.0 ? 0.0 : v;
        return METRIC == KM_EUCLIDEAN_GEMM ? sqrt(v) : v;
    } else {
        return acc;
    }
}

// With COUNT, the tile also counts per row the stored keys <= thr (skipping column
// self0 + row when self0 >= 0) and writes the count to tilecnt[row * gridDim.x + blockIdx.x],
// which saves a radius search one read of the key tile.
template <typename T, typename U, int METRIC, bool COUNT>
__global__ __launch_bounds__(256) void pairwise_direct_kernel(
    const T *__restrict__ X, const T *__restrict__ Y, int m, int n, int k, U *__restrict__ D,
    int64_t ldd, const double *__restrict__ nx, const double *__restrict__ ny, double p,
    int zero_diag, int diag_offset, U thr, int self0, int *__restrict__ tilecnt) {
    __shared__ double Xs[KC][TB + 1];
    __shared__ double Ys[KC][TB + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * TB, j0 = blockIdx.x * TB;

    double acc[TR][TR];
#pragma unroll
    for (int r = 0; r < TR; r++)
#pragma unroll
        for (int c = 0; c < TR; c++)
            acc[r][c] = 0.0;

    // Loader mapping: 16 consecutive k per row, 16 rows per pass, 4 passes.
    const int lk = tid & 15, lr = tid >> 4;
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
        for (int q = 0; q < KC; q++) {
            double a[TR], b[TR];
#pragma unroll
            for (int r = 0; r < TR; r++)
                a[r] = Xs[q][ty + 16 * r];
#pragma unroll
            for (int c = 0; c < TR; c++)
                b[c] = Ys[q][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < TR; r++)
#pragma unroll
                for (int c = 0; c < TR; c++)
                    acc[r][c] = accum<METRIC>(acc[r][c], a[r], b[c], p);
        }
        __syncthreads();
    }

    int hits[TR] = {};
#pragma unroll
    for (int r = 0; r < TR; r++) {
        const int i = i0 + ty + 16 * r;
        if (i >= m)
            continue;
#pragma unroll
        for (int c = 0; c < TR; c++) {
            const int j = j0 + tx + 16 * c;
            if (j >= n)
                continue;
            double v = finish<METRIC>(acc[r][c], i, j, nx, ny, p);
            if (zero_diag && i + diag_offset == j)
                v = 0.0;
            const U key = (U)v;
            D[(int64_t)i * ldd + j] = key;
            if constexpr (COUNT)
                hits[r] += key <= thr && !(self0 >= 0 && j == self0 + i);
        }
    }
    if constexpr (COUNT) {
        // The 16 lanes of equal ty hold one row's 64 columns between them.
#pragma unroll
        for (int r = 0; r < TR; r++) {
            int h = hits[r];
#pragma unroll
            for (int off = 8; off > 0; off >>= 1)
                h += __shfl_xor(h, off, 16);
            const int i = i0 + ty + 16 * r;
            if (tx == 0 && i < m)
                tilecnt[(int64_t)i * gridDim.x + blockIdx.x] = h;
        }
    }
}

template <typename T>
__global__ void colmajor_to_rowmajor_kernel(const T *__restrict__ src, int64_t ld, int rows,
                                            int cols, T *__restrict__ dst) {
    __shared__ T tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    for (int s = ty; s < 32; s += 8) {
        const int r = r0 + tx, c = c0 + s;
        if (r < rows && c < cols)
            tile[s][tx] = src[(int64_t)c * ld + r];
    }
    __syncthreads();
    for (int s = ty; s < 32; s += 8) {
        const int r = r0 + s, c = c0 + tx;
        if (r < rows && c < cols)
            dst[(int64_t)r * cols + c] = tile[tx][s];
    }
}

template <typename T, typename U>
__global__ void pack_rows_kernel(const T *__restrict__ src, int64_t lds, int rows, int cols,
                                 U *__restrict__ dst, int64_t ldd, int rows_pad) {
    const int64_t total = (int64_t)rows_pad * ldd;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / ldd;
        const int64_t c = e - r * ldd;
        U v = (U)0.0f;
        if (r < rows && c < cols)
            v = (U)(float)src[r * lds + c];
        if constexpr (sizeof(U) == 8)
            if (r < rows && c < cols)
                v = (U)src[r * lds + c];
        dst[e] = v;
    }
}

// Split operands: each value x is stored as hi = U(x) and lo = U(x - hi), rows laid out
// as [hi kp][lo kp].
template <typename T, typename U>
__global__ void pack_rows_split_kernel(const T *__restrict__ src, int64_t lds, int rows, int cols,
                                       U *__restrict__ dst, int64_t kp, int rows_pad) {
    const int64_t total = (int64_t)rows_pad * kp;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / kp;
        const int64_t c = e - r * kp;
        const double x = (r < rows && c < cols) ? (double)src[r * lds + c] : 0.0;
        const U hi = (U)(float)x;
        const U lo = (U)(float)(x - (double)(float)hi);
        dst[r * 2 * kp + c] = hi;
        dst[r * 2 * kp + kp + c] = lo;
    }
}

template <typename T>
__global__ void row_sqnorms_kernel(const T *__restrict__ X, int64_t ld, int rows, int cols,
                                   double *__restrict__ out) {
    // One wave per row.
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (row >= rows)
        return;
    double s = 0.0;
    for (int c = lane; c < cols; c += 64) {
        double v = (double)X[(int64_t)row * ld + c];
        s = fma(v, v, s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, 64);
    if (lane == 0)
        out[row] = s;
}

__global__ void cosine_norms_kernel(double *sq, int rows, double eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) {
        double v = sqrt(sq[i]);
        sq[i] = (v <= eps) ? 1.0 : v;
    }
}

} // namespace

template <typename T>
hipError_t launch_colmajor_to_rowmajor(const T *src, int64_t ld, int rows, int cols, T *dst,
                                       hipStream_t s) {
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(colmajor_to_rowmajor_kernel<T>, grid, dim3(256), 0, s, src, ld, rows,
                       cols, dst);
    return hipGetLastError();
}

template <typename T, typename U>
hipError_t launch_pack_rows(const T *src, int64_t lds, int rows, int cols, U *dst, int64_t ldd,
                            int rows_pad, hipStream_t s) {
    const int64_t total = (int64_t)rows_pad * ldd;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    if (blocks < 1)
        blocks = 1;
    hipLaunchKernelGGL((pack_rows_kernel<T, U>), dim3((unsigned)blocks), dim3(256), 0, s, src,
                       lds, rows, cols, dst, ldd, rows_pad);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_row_sqnorms(const T *X, int64_t ld, int rows, int cols, double *out,
                              hipStream_t s) {
    const int waves = 4;
    dim3 grid((rows + waves - 1) / waves);
    hipLaunchKernelGGL(row_sqnorms_kernel<T>, grid, dim3(64 * waves), 0, s, X, ld, rows, cols,
                       out);
    return hipGetLastError();
}

hipError_t launch_cosine_norms(double *sq, int rows, double eps, hipStream_t s) {
    hipLaunchKernelGGL(cosine_norms_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, sq, rows,
                       eps);
    return hipGetLastError();
}

template <typename T, typename U>
hipError_t launch_pairwise_direct(int metric, const T *X, const T *Y, int m, int n, int k,
                                  U *D, int64_t ldd, const double *nx, const double *ny,
                                  double p, int zero_diag, int diag_offset, hipStream_t s) {
    dim3 grid((n + TB - 1) / TB, (m + TB - 1) / TB);
    dim3 block(256);
#define DA_PW_CASE(M)                                                                        \
    case M:                                                                                  \
        hipLaunchKernelGGL((pairwise_direct_kernel<T, U, M, false>), grid, block, 0, s, X, Y, \
                           m, n, k, D, ldd, nx, ny, p, zero_diag, diag_offset, (U)0, -1,      \
                           nullptr);                                                         \
        break;
    switch (metric) {
        DA_PW_CASE(KM_SQEUCLIDEAN)
        DA_PW_CASE(KM_EUCLIDEAN)
        DA_PW_CASE(KM_MANHATTAN)
        DA_PW_CASE(KM_MINKOWSKI)
        DA_PW_CASE(KM_COSINE)
        DA_PW_CASE(KM_INNER)
        DA_PW_CASE(KM_SQEUCLIDEAN_GEMM)
        DA_PW_CASE(KM_EUCLIDEAN_GEMM)
        DA_PW_CASE(KM_MINKOWSKI_KEY)
    default:
        return hipErrorInvalidValue;
    }
#undef DA_PW_CASE
    return hipGetLastError();
}

template <typename T, typename U>
hipError_t launch_pairwise_direct_count(int metric, const T *X, const T *Y, int m, int n, int k,
                                        U *D, int64_t ldd, const double *nx, const double *ny,
                                        double p, U thr, int self0, int *tilecnt,
                                        hipStream_t s) {
    dim3 grid((n + TB - 1) / TB, (m + TB - 1) / TB);
    dim3 block(256);
#define DA_PWC_CASE(M)                                                                       \
    case M:                                                                                  \
        hipLaunchKernelGGL((pairwise_direct_kernel<T, U, M, true>), grid, block, 0, s, X, Y,  \
                           m, n, k, D, ldd, nx, ny, p, 0, 0, thr, self0, tilecnt);           \
        break;
    switch (metric) {
        DA_PWC_CASE(KM_SQEUCLIDEAN)
        DA_PWC_CASE(KM_EUCLIDEAN)
        DA_PWC_CASE(KM_MANHATTAN)
        DA_PWC_CASE(KM_MINKOWSKI)
        DA_PWC_CASE(KM_COSINE)
        DA_PWC_CASE(KM_SQEUCLIDEAN_GEMM)
    default:
        return hipErrorInvalidValue;
    }
#undef DA_PWC_CASE
    return hipGetLastError();
}
template hipError_t launch_pairwise_direct_count<float, float>(int, const float *, const float *,
                                                               int, int, int, float *, int64_t,
                                                               const double *, const double *,
                                                               double, float, int, int *,
                                                               hipStream_t);
template hipError_t launch_pairwise_direct_count<double, double>(
    int, const double *, const double *, int, int, int, double *, int64_t, const double *,
    const double *, double, double, int, int *, hipStream_t);

template <typename T>
hipError_t launch_pack_rows_prec(const T *src, int64_t lds, int rows, int cols, void *dst,
                                 int64_t ldd, int rows_pad, int prec, hipStream_t s) {
    switch (prec) {
    case 1: // KP_F64
        return launch_pack_rows<T, double>(src, lds, rows, cols, static_cast<double *>(dst), ldd,
                                           rows_pad, s);
    case 0: // KP_F32
        return launch_pack_rows<T, float>(src, lds, rows, cols, static_cast<float *>(dst), ldd,
                                          rows_pad, s);
    case 2: // KP_F16
    case 6: // KP_F16C (the queries are plain fp16)
        return launch_pack_rows<T, _Float16>(src, lds, rows, cols, static_cast<_Float16 *>(dst),
                                             ldd, rows_pad, s);
    case 3: // KP_BF16
        return launch_pack_rows<T, __bf16>(src, lds, rows, cols, static_cast<__bf16 *>(dst), ldd,
                                           rows_pad, s);
    case 4: // KP_F16X3
    case 5: { // KP_BF16X3
        const int64_t total = (int64_t)rows_pad * ldd;
        int64_t blocks = (total + 255) / 256;
        blocks = blocks > 65536 ? 65536 : (blocks < 1 ? 1 : blocks);
        if (prec == 4)
            hipLaunchKernelGGL((pack_rows_split_kernel<T, _Float16>), dim3((unsigned)blocks),
                               dim3(256), 0, s, src, lds, rows, cols, static_cast<_Float16 *>(dst),
                               ldd, rows_pad);
        else
            hipLaunchKernelGGL((pack_rows_split_kernel<T, __bf16>), dim3((unsigned)blocks),
                               dim3(256), 0, s, src, lds, rows, cols, static_cast<__bf16 *>(dst),
                               ldd, rows_pad);
        return hipGetLastError();
    }
    default:
        return hipErrorInvalidValue;
    }
}
template hipError_t launch_pack_rows_prec<float>(const float *, int64_t, int, int, void *,
                                                 int64_t, int, int, hipStream_t);
template hipError_t launch_pack_rows_prec<double>(const double *, int64_t, int, int, void *,
                                                  int64_t, int, int, hipStream_t);

template hipError_t launch_colmajor_to_rowmajor<float>(const float *, int64_t, int, int, float *,
                                                       hipStream_t);
template hipError_t launch_colmajor_to_rowmajor<double>(const double *, int64_t, int, int,
                                                        double *, hipStream_t);
template hipError_t launch_pack_rows<float, float>(const float *, int64_t, int, int, float *,
                                                   int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<double, double>(const double *, int64_t, int, int, double *,
                                                     int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<float, _Float16>(const float *, int64_t, int, int, _Float16 *,
                                                      int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<float, __bf16>(const float *, int64_t, int, int, __bf16 *,
                                                    int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<double, _Float16>(const double *, int64_t, int, int,
                                                       _Float16 *, int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<double, __bf16>(const double *, int64_t, int, int, __bf16 *,
                                                     int64_t, int, hipStream_t);
template hipError_t launch_pack_rows<double, float>(const double *, int64_t, int, int, float *,
                                                    int64_t, int, hipStream_t);
template hipError_t launch_row_sqnorms<float>(const float *, int64_t, int, int, double *,
                                              hipStream_t);
template hipError_t launch_row_sqnorms<double>(const double *, int64_t, int, int, double *,
                                               hipStream_t);
template hipError_t launch_pairwise_direct<float, float>(int, const float *, const float *, int,
                                                         int, int, float *, int64_t,
                                                         const double *, const double *, double,
                                                         int, int, hipStream_t);
template hipError_t launch_pairwise_direct<double, double>(int, const double *, const double *,
                                                           int, int, int, double *, int64_t,
                                                           const double *, const double *,
                                                           double, int, int, hipStream_t);
template hipError_t launch_pairwise_direct<float, double>(int, const float *, const float *, int,
                                                          int, int, double *, int64_t,
                                                          const double *, const double *, double,
                                                          int, int, hipStream_t);

} // namespace dagpu
