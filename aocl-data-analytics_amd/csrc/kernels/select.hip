// Per-row k-smallest selection over a materialised key matrix.
//
// The reference selects with a "replace the current maximum" scan per query followed by
// a stable sort (source/core/nearest_neighbors/nearest_neighbors_utils.hpp:44-101). On the
// GPU one workgroup owns one row: each element is compared with a running threshold (the
// k-th best (key, index) pair kept so far) and, when it beats it, appended to an LDS
// candidate buffer with an LDS atomic. When the buffer is close to full the workgroup
// bitonic-sorts it, keeps the k best and tightens the threshold. Elements are ordered by
// (key, column index), so the result is deterministic, including at equal keys.
#include "launch.hpp"
#include <climits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

template <typename K> __device__ __forceinline__ bool pair_less(K ka, int ia, K kb, int ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// Sorts bkey/bidx[0..CAP) ascending by (key, index). All threads of the block must call.
template <typename K, int CAP>
__device__ void block_bitonic_sort(K *bkey, int *bidx) {
    for (int size = 2; size <= CAP; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < CAP / 2; t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                K ki = bkey[i], kj = bkey[j];
                int ii = bidx[i], ij = bidx[j];
                const bool gt = pair_less(kj, ij, ki, ii);
                if (gt == asc) {
                    bkey[i] = kj;
                    bkey[j] = ki;
                    bidx[i] = ij;
                    bidx[j] = ii;
                }
            }
            __syncthreads();
        }
    }
}

template <typename K, int CAP>
__global__ __launch_bounds__(256) void row_topk_kernel(const K *__restrict__ keys, int64_t ld,
                                                       int rows, int n, int k, int largest,
                                                       K *__restrict__ out_key,
                                                       int *__restrict__ out_idx) {
    __shared__ K bkey[CAP];
    __shared__ int bidx[CAP];
    __shared__ int count;
    __shared__ K thr_k;
    __shared__ int thr_i;
    const int row = blockIdx.x;
    if (row >= rows)
        return;
    const K inf = (K)__builtin_huge_val();
    if (threadIdx.x == 0) {
        count = 0;
        thr_k = inf;
        thr_i = INT_MAX;
    }
    __syncthreads();
    const K *rk = keys + (int64_t)row * ld;
    for (int base = 0; base < n; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n) {
            K v = rk[j];
            if (largest)
                v = -v;
            if (v != v)
                v = inf;
            if (pair_less(v, j, thr_k, thr_i)) {
                const int pos = atomicAdd(&count, 1);
                bkey[pos] = v;
                bidx[pos] = j;
            }
        }
        __syncthreads();
        const int c = count;
        // Every thread must read the count before any thread appends again.
        __syncthreads();
        if (c > CAP - 256) {
            for (int t = c + threadIdx.x; t < CAP; t += blockDim.x) {
                bkey[t] = inf;
                bidx[t] = INT_MAX;
            }
            __syncthreads();
            block_bitonic_sort<K, CAP>(bkey, bidx);
            if (threadIdx.x == 0) {
                count = c < k ? c : k;
                if (c >= k) {
                    thr_k = bkey[k - 1];
                    thr_i = bidx[k - 1];
                }
            }
            __syncthreads();
        }
    }
    const int c = count;
    for (int t = c + threadIdx.x; t < CAP; t += blockDim.x) {
        bkey[t] = inf;
        bidx[t] = INT_MAX;
    }
    __syncthreads();
    block_bitonic_sort<K, CAP>(bkey, bidx);
    for (int t = threadIdx.x; t < k; t += blockDim.x) {
        K v = bkey[t];
        out_key[(int64_t)row * k + t] = largest ? -v : v;
        out_idx[(int64_t)row * k + t] = bidx[t];
    }
}

} // namespace

template <typename K>
hipError_t launch_row_topk(const K *keys, int64_t ld, int rows, int n, int k, int largest,
                           K *out_key, int *out_idx, hipStream_t s) {
    if (k < 1 || k > 3840 || rows < 1)
        return hipErrorInvalidValue;
    dim3 grid(rows), block(256);
    if (k + 256 <= 512)
        hipLaunchKernelGGL((row_topk_kernel<K, 512>), grid, block, 0, s, keys, ld, rows, n, k,
                           largest, out_key, out_idx);
    else if (k + 256 <= 1024)
        hipLaunchKernelGGL((row_topk_kernel<K, 1024>), grid, block, 0, s, keys, ld, rows, n, k,
                           largest, out_key, out_idx);
    else if (k + 256 <= 2048)
        hipLaunchKernelGGL((row_topk_kernel<K, 2048>), grid, block, 0, s, keys, ld, rows, n, k,
                           largest, out_key, out_idx);
    else
        hipLaunchKernelGGL((row_topk_kernel<K, 4096>), grid, block, 0, s, keys, ld, rows, n, k,
                           largest, out_key, out_idx);
    return hipGetLastError();
}

template hipError_t launch_row_topk<float>(const float *, int64_t, int, int, int, int, float *,
                                           int *, hipStream_t);
template hipError_t launch_row_topk<double>(const double *, int64_t, int, int, int, int,
                                            double *, int *, hipStream_t);

} // namespace dagpu

namespace dagpu {
namespace {
template <typename T>
__global__ void knn_finalize_kernel(const double *__restrict__ keys, const int *__restrict__ idx,
                                    int rows, int k, int post, double p, T *__restrict__ out_dist,
                                    int *__restrict__ out_idx, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * k)
        return;
    const int64_t q = e / k, j = e - q * k;
    double v = keys[e];
    if (post == 1)
        v = sqrt(v < 0.0 ? 0.0 : v);
    else if (post == 2)
        v = pow(v, 1.0 / p);
    if (out_dist)
        out_dist[q * ldo + j] = (T)v;
    out_idx[q * ldo + j] = idx[e];
}
} // namespace

template <typename T>
hipError_t launch_knn_finalize(const double *keys, const int *idx, int rows, int k, int post,
                               double p, T *out_dist, int *out_idx, int64_t ldo,
                               hipStream_t s) {
    const int64_t total = (int64_t)rows * k;
    hipLaunchKernelGGL(knn_finalize_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, keys, idx, rows, k, post, p, out_dist, out_idx, ldo);
    return hipGetLastError();
}
template hipError_t launch_knn_finalize<float>(const double *, const int *, int, int, int,
                                               double, float *, int *, int64_t, hipStream_t);
template hipError_t launch_knn_finalize<double>(const double *, const int *, int, int, int,
                                                double, double *, int *, int64_t, hipStream_t);
} // namespace dagpu

namespace dagpu {
namespace {
template <typename E>
__global__ void gather_scatter_kernel(const E *__restrict__ src, E *__restrict__ dst,
                                      const int *__restrict__ idx, int m, int cols, int scatter) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * cols)
        return;
    const int64_t i = e / cols, c = e - i * cols;
    const int64_t row = idx[i];
    if (scatter)
        dst[row * cols + c] = src[i * cols + c];
    else
        dst[i * cols + c] = src[row * cols + c];
}
} // namespace
hipError_t launch_gather_scatter(const void *src, void *dst, const int *idx, int m, int cols,
                                 int elem, int scatter, hipStream_t s) {
    const int64_t total = (int64_t)m * cols;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (elem == 8)
        hipLaunchKernelGGL(gather_scatter_kernel<long long>, grid, block, 0, s,
                           static_cast<const long long *>(src), static_cast<long long *>(dst), idx,
                           m, cols, scatter);
    else
        hipLaunchKernelGGL(gather_scatter_kernel<int>, grid, block, 0, s,
                           static_cast<const int *>(src), static_cast<int *>(dst), idx, m, cols,
                           scatter);
    return hipGetLastError();
}
} // namespace dagpu
