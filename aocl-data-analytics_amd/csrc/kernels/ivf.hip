// IVF-Flat search kernels for gfx950: the probed inverted lists of each query are scanned
// with exact distances into one padded candidate row per query, which the row top-k
// selection then orders by (key, column).
//
// Layout: the indexed vectors are held once, in the order they were added (row-major,
// d columns); an inverted list is a range [list_off[c], list_off[c+1]) of perm, the global
// ids of its vectors in list order. Candidate row q of a query block holds the keys of list
// probes[q][p] at columns seg[q][p] + r; the host places the probed lists in ascending list
// number, so that column order is (list, position in list).
//
// Keys: the squared euclidean distance, or for the inner-product and cosine metrics (IP)
// the negated dot product -q.x (the reference's GEMM with alpha = -1, approximate_neighbors.cpp
// :1145-1151), so that the smallest key is the nearest vector for every metric.
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
template <typename T, bool IP>
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
                if constexpr (IP) {
                    acc += (double)qv[l] * (double)xv[l];
                } else {
                    const double t = (double)qv[l] - (double)xv[l];
                    acc += t * t;
                }
            }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (live && sub == 0) {
            kr[r - b] = (T)(IP ? -acc : acc);
            ir[r - b] = g;
        }
    }
}

// 16-byte chunks of a row: four floats or two doubles. acc adds the squared differences of
// a chunk's elements in element order, dot their products.
template <typename T> struct Chunk16;
template <> struct Chunk16<float> {
    using type = float4;
    __device__ static type zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
    __device__ static void acc(double &a, const type &q, const type &x) {
        const double t0 = (double)q.x - (double)x.x, t1 = (double)q.y - (double)x.y;
        const double t2 = (double)q.z - (double)x.z, t3 = (double)q.w - (double)x.w;
        a += t0 * t0;
        a += t1 * t1;
        a += t2 * t2;
        a += t3 * t3;
    }
    __device__ static void dot(double &a, const type &q, const type &x) {
        a += (double)q.x * (double)x.x;
        a += (double)q.y * (double)x.y;
        a += (double)q.z * (double)x.z;
        a += (double)q.w * (double)x.w;
    }
};
template <> struct Chunk16<double> {
    using type = double2;
    __device__ static type zero() { return make_double2(0.0, 0.0); }
    __device__ static void acc(double &a, const type &q, const type &x) {
        const double t0 = q.x - x.x, t1 = q.y - x.y;
        a += t0 * t0;
        a += t1 * t1;
    }
    __device__ static void dot(double &a, const type &q, const type &x) {
        a += q.x * x.x;
        a += q.y * x.y;
    }
};

// The same scan for rows of whole 16-byte chunks (dc chunks, at most 8 * NJ): eight lanes
// share a vector and read it chunk by chunk (lane s takes chunks s, s + 8, ...), so a wave
// takes eight vectors per iteration with a quarter (fp32) or half (fp64) of the load
// instructions; the query's chunks stay in registers for the whole list. Distances
// accumulate in double, chunk by chunk.
template <typename T, int NJ, bool IP>
__global__ __launch_bounds__(256) void ivf_scan_chunk_kernel(
    const T *__restrict__ Q, int nq, int n_probe, int dc, const T *__restrict__ X,
    const int *__restrict__ perm, const int *__restrict__ list_off,
    const int *__restrict__ probes, const int *__restrict__ seg, int64_t ld,
    T *__restrict__ keys, int *__restrict__ idx) {
    using CT = Chunk16<T>;
    using ch_t = typename CT::type;
    constexpr int G = 8;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (G - 1), slot = lane / G;
    const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= (int64_t)nq * n_probe)
        return;
    const int q = (int)(pair / n_probe);
    const int c = probes[pair];
    const int b = list_off[c], e = list_off[c + 1];
    const ch_t *qv = reinterpret_cast<const ch_t *>(Q) + (int64_t)q * dc;
    ch_t qr[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int ch = sub + G * j;
        qr[j] = ch < dc ? qv[ch] : CT::zero();
    }
    T *kr = keys + (int64_t)q * ld + seg[pair];
    int *ir = idx + (int64_t)q * ld + seg[pair];
    for (int r0 = b; r0 < e; r0 += 64 / G) {
        const int r = r0 + slot;
        const bool live = r < e;
        const int g = live ? perm[r] : 0;
        const ch_t *xv = reinterpret_cast<const ch_t *>(X) + (int64_t)g * dc;
        double acc = 0.0;
        if (live) {
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int ch = sub + G * j;
                if (ch < dc) {
                    if constexpr (IP)
                        CT::dot(acc, qr[j], xv[ch]);
                    else
                        CT::acc(acc, qr[j], xv[ch]);
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (live && sub == 0) {
            kr[r - b] = (T)(IP ? -acc : acc);
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
// id stored at that column, index -1 where a query's probed lists held fewer than k vectors.
// Distances (approximate_neighbors.cpp:1287-1312): post 0 the key, 1 its square root
// (euclidean), 2 1 + key (cosine), 3 -key (inner product); a missing neighbour has key +inf,
// so its distance is +inf, or -inf for the inner product.
template <typename T>
__global__ void ivf_finish_kernel(const T *__restrict__ key, const int *__restrict__ sel,
                                  const int *__restrict__ cand_idx, int64_t ld, int k, int64_t n,
                                  int post, T *__restrict__ dist, int *__restrict__ ind) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int g = cand_idx[(t / k) * ld + sel[t]];
        T v = key[t];
        if (g == INT_MAX) {
            ind[t] = -1;
            v = std::numeric_limits<T>::infinity();
        } else {
            ind[t] = g;
        }
        if (post == 1 && v >= (T)0)
            v = sqrt(v);
        else if (post == 2)
            v = (T)1 + v;
        else if (post == 3)
            v = -v;
        if (dist)
            dist[t] = v;
    }
}

// Rows scaled to unit length in place (normalize_rows, da_utils.cpp:857-905): 1/|x| in
// double, the product rounded to T; a zero row is left as it is. One thread per row.
template <typename T>
__global__ void ivf_normalize_kernel(T *__restrict__ X, int64_t n, int d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T *x = X + i * d;
    double s = 0.0;
    for (int l = 0; l < d; l++)
        s += (double)x[l] * (double)x[l];
    const double nrm = sqrt(s);
    if (nrm > 0.0) {
        const double inv = 1.0 / nrm;
        for (int l = 0; l < d; l++)
            x[l] = (T)((double)x[l] * inv);
    }
}

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

} // namespace

template <typename T, bool IP>
static hipError_t ivf_scan_dispatch(const T *Q, int nq, int n_probe, int d, const T *X,
                                    const int *perm, const int *list_off, const int *probes,
                                    const int *seg, int64_t ld, T *keys, int *idx,
                                    hipStream_t s) {
    const int64_t pairs = (int64_t)nq * n_probe;
    if (pairs <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((pairs + 3) / 4));
    // Rows of whole 16-byte chunks (d a multiple of 4 for fp32, of 2 for fp64; the buffers
    // are allocated 256-byte aligned), at most 128 elements, take the chunked scan.
    constexpr int per = 16 / (int)sizeof(T);
    const int dc = d / per;
    if (d % per == 0 && d <= 128) {
        auto run = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, Q, nq, n_probe, dc, X, perm,
                               list_off, probes, seg, ld, keys, idx);
            return hipGetLastError();
        };
        switch ((dc + 7) / 8) {
        case 1: return run(ivf_scan_chunk_kernel<T, 1, IP>);
        case 2: return run(ivf_scan_chunk_kernel<T, 2, IP>);
        case 3: return run(ivf_scan_chunk_kernel<T, 3, IP>);
        case 4: return run(ivf_scan_chunk_kernel<T, 4, IP>);
        }
        if constexpr (per == 2) {
            switch ((dc + 7) / 8) {
            case 5: return run(ivf_scan_chunk_kernel<T, 5, IP>);
            case 6: return run(ivf_scan_chunk_kernel<T, 6, IP>);
            case 7: return run(ivf_scan_chunk_kernel<T, 7, IP>);
            case 8: return run(ivf_scan_chunk_kernel<T, 8, IP>);
            }
        }
    }
    hipLaunchKernelGGL((ivf_scan_kernel<T, IP>), grid, dim3(256), 0, s, Q, nq, n_probe, d, X,
                       perm, list_off, probes, seg, ld, keys, idx);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ivf_scan(const T *Q, int nq, int n_probe, int d, const T *X, const int *perm,
                           const int *list_off, const int *probes, const int *seg, int64_t ld,
                           T *keys, int *idx, hipStream_t s, int ip) {
    return ip ? ivf_scan_dispatch<T, true>(Q, nq, n_probe, d, X, perm, list_off, probes, seg,
                                           ld, keys, idx, s)
              : ivf_scan_dispatch<T, false>(Q, nq, n_probe, d, X, perm, list_off, probes, seg,
                                            ld, keys, idx, s);
}

template <typename T> hipError_t launch_ivf_normalize(T *X, int64_t n, int d, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(ivf_normalize_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, X, n, d);
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
                             int64_t n, int post, T *dist, int *ind, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(ivf_finish_kernel<T>, dim3(blocks_for(n)), dim3(256), 0, s, key, sel,
                       cand_idx, ld, k, n, post, dist, ind);
    return hipGetLastError();
}

#define DA_IVF_INST(T)                                                                           \
    template hipError_t launch_ivf_scan<T>(const T *, int, int, int, const T *, const int *,     \
                                           const int *, const int *, const int *, int64_t, T *,  \
                                           int *, hipStream_t, int);                             \
    template hipError_t launch_ivf_normalize<T>(T *, int64_t, int, hipStream_t);                 \
    template hipError_t launch_ivf_pad<T>(T *, int *, int64_t, hipStream_t);                     \
    template hipError_t launch_ivf_finish<T>(const T *, const int *, const int *, int64_t, int,  \
                                             int64_t, int, T *, int *, hipStream_t);
DA_IVF_INST(float)
DA_IVF_INST(double)
#undef DA_IVF_INST

} // namespace dagpu
