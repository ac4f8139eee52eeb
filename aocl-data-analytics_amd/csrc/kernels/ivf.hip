// IVF-Flat search kernels for gfx950: the probed inverted lists of each query are scanned
// with exact distances into one padded candidate row per query, which the row top-k
// selection then orders by (key, column).
//
// Layout: the indexed vectors are held once, in the order they were added (row-major,
// d columns); an inverted list is a range [list_off[c], list_off[c+1]) of perm, the global
// ids of its vectors in list order. Candidate row q of a query block holds the keys of list
// probes[q][p] at columns seg[q][p] + r; the host places the probed lists in ascending list
// number, so that column order is (list, position in list).
#include "launch.hpp"
#include <climits>
#include <type_traits>
#include <hip/hip_runtime.h>
#include <limits>

namespace dagpu {

namespace {

// One wave per (query, probe) pair. Sixteen lanes share a vector: lane s of a group sums
// features s, s + 16, ... (consecutive lanes read consecutive features, so each group's
// loads are contiguous), then the group adds its partial sums with lane shuffles. A wave
// therefore takes four vectors of the list at a time; distances accumulate in double.
template <typename T>
__global__ __launch_bounds__(256) void ivf_scan_kernel(
    const T *__restrict__ Q, int nq, int n_probe, int d, const T *__restrict__ X,
    const int *__restrict__ perm, const int *__restrict__ list_off,
    const int *__restrict__ probes, const int *__restrict__ seg, int64_t ld,
    T *__restrict__ keys, int *__restrict__ idx) {
    constexpr int G = 16;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (G - 1), slot = lane / G;
    const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= (int64_t)nq * n_probe)
        return;
    const int q = (int)(pair / n_probe);
    const int c = probes[pair];
    const int b = list_off[c], e = list_off[c + 1];
    const T *qv = Q + (int64_t)q * d;
    T *kr = keys + (int64_t)q * ld + seg[pair];
    int *ir = idx + (int64_t)q * ld + seg[pair];
    for (int r0 = b; r0 < e; r0 += 64 / G) {
        const int r = r0 + slot;
        const bool live = r < e;
        const int g = live ? perm[r] : 0;
        const T *xv = X + (int64_t)g * d;
        double acc = 0.0;
        if (live)
            for (int l = sub; l < d; l += G) {
                const double t = (double)qv[l] - (double)xv[l];
                acc += t * t;
            }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (live && sub == 0) {
            kr[r - b] = (T)acc;
            ir[r - b] = g;
        }
    }
}

// The same scan for fp32 rows of d = 4 * d4 <= 128 features: eight lanes share a vector and
// read it as 16-byte chunks (lane s takes chunks s, s + 8, ...), so a wave takes eight
// vectors per iteration with a quarter of the load instructions; the query's chunks stay
// in registers for the whole list. Distances accumulate in double, chunk by chunk.
template <int NJ>
__global__ __launch_bounds__(256) void ivf_scan4_kernel(
    const float *__restrict__ Q, int nq, int n_probe, int d4, const float *__restrict__ X,
    const int *__restrict__ perm, const int *__restrict__ list_off,
    const int *__restrict__ probes, const int *__restrict__ seg, int64_t ld,
    float *__restrict__ keys, int *__restrict__ idx) {
    constexpr int G = 8;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (G - 1), slot = lane / G;
    const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= (int64_t)nq * n_probe)
        return;
    const int q = (int)(pair / n_probe);
    const int c = probes[pair];
    const int b = list_off[c], e = list_off[c + 1];
    const float4 *qv = reinterpret_cast<const float4 *>(Q) + (int64_t)q * d4;
    float4 qr[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int ch = sub + G * j;
        qr[j] = ch < d4 ? qv[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float *kr = keys + (int64_t)q * ld + seg[pair];
    int *ir = idx + (int64_t)q * ld + seg[pair];
    for (int r0 = b; r0 < e; r0 += 64 / G) {
        const int r = r0 + slot;
        const bool live = r < e;
        const int g = live ? perm[r] : 0;
        const float4 *xv = reinterpret_cast<const float4 *>(X) + (int64_t)g * d4;
        double acc = 0.0;
        if (live) {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int ch = sub + G * j;
                if (ch < d4) {
                    const float4 x = xv[ch];
                    const double t0 = (double)qr[j].x - (double)x.x;
                    const double t1 = (double)qr[j].y - (double)x.y;
                    const double t2 = (double)qr[j].z - (double)x.z;
                    const double t3 = (double)qr[j].w - (double)x.w;
                    acc += t0 * t0;
                    acc += t1 * t1;
                    acc += t2 * t2;
                    acc += t3 * t3;
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (live && sub == 0) {
            kr[r - b] = (float)acc;
            ir[r - b] = g;
        }
    }
}

template <typename T>
__global__ void ivf_pad_kernel(T *__restrict__ keys, int *__restrict__ idx, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        keys[t] = std::numeric_limits<T>::infinity();
        idx[t] = INT_MAX;
    }
}

// Selected candidates (key, column of the candidate row) -> reported neighbours: the global
// id stored at that column, index -1 and distance +inf where a query's probed lists held
// fewer than k vectors; the square root for the euclidean metric.
template <typename T>
__global__ void ivf_finish_kernel(const T *__restrict__ key, const int *__restrict__ sel,
                                  const int *__restrict__ cand_idx, int64_t ld, int k, int64_t n,
                                  int take_sqrt, T *__restrict__ dist, int *__restrict__ ind) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int g = cand_idx[(t / k) * ld + sel[t]];
        T v = key[t];
        if (g == INT_MAX) {
            ind[t] = -1;
            v = std::numeric_limits<T>::infinity();
        } else {
            ind[t] = g;
            if (take_sqrt && v >= (T)0)
                v = sqrt(v);
        }
        if (dist)
            dist[t] = v;
    }
}

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

} // namespace

template <typename T>
hipError_t launch_ivf_scan(const T *Q, int nq, int n_probe, int d, const T *X, const int *perm,
                           const int *list_off, const int *probes, const int *seg, int64_t ld,
                           T *keys, int *idx, hipStream_t s) {
    const int64_t pairs = (int64_t)nq * n_probe;
    if (pairs <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((pairs + 3) / 4));
    if constexpr (std::is_same_v<T, float>) {
        // 16-byte chunks need rows of whole chunks (d % 4 == 0; the buffers are allocated
        // 256-byte aligned), at most 32 chunks.
        const int d4 = d / 4;
        if (d % 4 == 0 && d4 <= 32) {
            auto run = [&](auto kern) {
                hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, Q, nq, n_probe, d4, X, perm,
                                   list_off, probes, seg, ld, keys, idx);
                return hipGetLastError();
            };
            if (d4 <= 8)
                return run(ivf_scan4_kernel<1>);
            if (d4 <= 16)
                return run(ivf_scan4_kernel<2>);
            if (d4 <= 24)
                return run(ivf_scan4_kernel<3>);
            return run(ivf_scan4_kernel<4>);
        }
    }
    hipLaunchKernelGGL(ivf_scan_kernel<T>, grid, dim3(256), 0, s, Q, nq, n_probe, d, X, perm,
                       list_off, probes, seg, ld, keys, idx);
    return hipGetLastError();
}

template <typename T> hipError_t launch_ivf_pad(T *keys, int *idx, int64_t n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(ivf_pad_kernel<T>, dim3(blocks_for(n)), dim3(256), 0, s, keys, idx, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ivf_finish(const T *key, const int *sel, const int *cand_idx, int64_t ld, int k,
                             int64_t n, int take_sqrt, T *dist, int *ind, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(ivf_finish_kernel<T>, dim3(blocks_for(n)), dim3(256), 0, s, key, sel,
                       cand_idx, ld, k, n, take_sqrt, dist, ind);
    return hipGetLastError();
}

#define DA_IVF_INST(T)                                                                           \
    template hipError_t launch_ivf_scan<T>(const T *, int, int, int, const T *, const int *,     \
                                           const int *, const int *, const int *, int64_t, T *,  \
                                           int *, hipStream_t);                                  \
    template hipError_t launch_ivf_pad<T>(T *, int *, int64_t, hipStream_t);                     \
    template hipError_t launch_ivf_finish<T>(const T *, const int *, const int *, int64_t, int,  \
                                             int64_t, int, T *, int *, hipStream_t);
DA_IVF_INST(float)
DA_IVF_INST(double)
#undef DA_IVF_INST

} // namespace dagpu
