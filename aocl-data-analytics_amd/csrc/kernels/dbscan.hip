// DBSCAN cluster formation on gfx950 over the eps-neighbour graph of the samples, held in
// CSR form (offsets, neighbour indices; a point is not its own neighbour).
//
// The reference forms clusters by a breadth-first search from each unvisited core point in
// index order (dbscan.cpp:395-445, the serial method). Its result is fixed by three facts,
// which the kernels below compute in parallel:
//  * a cluster is a connected set of core points (neighbouring core points are joined);
//  * clusters are numbered in the order of their smallest core index (the point that
//    starts each search);
//  * a non-core point joins the first numbered cluster that reaches it, i.e. the smallest
//    cluster number among its core neighbours, and is noise (-1) without one.
// The connected sets come from a lock-free union-find: every tree is hooked under the
// smaller root with a compare-and-swap, so each final root is its set's smallest index and
// an exclusive scan of the root flags numbers the clusters in that order.
#include "launch.hpp"
#include <climits>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace dagpu {

namespace {

// Parent links are read and written by many waves at once: device-scope atomic loads and
// stores keep every wave's view coherent (no stale per-CU cache lines).
__device__ inline int ld_link(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_link(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving. Every link points to a smaller index (parent[x] < x unless x
// is a root), so the walk strictly decreases and ends; a halving store replaces a link by
// an ancestor, which keeps that property under concurrent updates.
__device__ int uf_find(int *parent, int x) {
    int p = ld_link(parent + x);
    while (p != x) {
        const int g = ld_link(parent + p);
        if (g == p)
            return p;
        st_link(parent + x, g);
        x = g;
        p = ld_link(parent + x);
    }
    return x;
}

// Root of x without writes.
__device__ int uf_root(const int *parent, int x) {
    int p = ld_link(parent + x);
    while (p != x) {
        x = p;
        p = ld_link(parent + x);
    }
    return x;
}

// Joins the trees of a and b: the larger root is hooked under the smaller one. A failed
// compare-and-swap means the larger root was hooked elsewhere in the meantime, so its root
// has become strictly smaller; roots only decrease, hence the loop ends.
__device__ void uf_unite(int *parent, int a, int b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b)
            return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        int expected = b;
        if (__hip_atomic_compare_exchange_strong(parent + b, &expected, a, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

__global__ void dbscan_init_kernel(int *__restrict__ parent, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        parent[i] = i;
}

// One wave per core point; its lanes take the neighbours 64 at a time and join each core
// neighbour of larger index (every edge once).
__global__ __launch_bounds__(256) void dbscan_union_kernel(const long long *__restrict__ offs,
                                                           const int *__restrict__ ind,
                                                           const unsigned char *__restrict__ core,
                                                           int *parent, int n) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (row >= n || !core[row])
        return;
    const long long e = offs[row + 1];
    for (long long t = offs[row] + lane; t < e; t += 64) {
        const int j = ind[t];
        if (j > row && core[j])
            uf_unite(parent, row, j);
    }
}

// After the union launch has completed the links no longer change: each core point's link
// is replaced by its root (only thread i writes parent[i], and the roots read by others are
// unaffected by it).
__global__ void dbscan_flatten_kernel(const unsigned char *__restrict__ core, int *parent,
                                      int *__restrict__ is_root, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!core[i]) {
            is_root[i] = 0;
            continue;
        }
        const int r = uf_root(parent, i);
        is_root[i] = r == i;
        st_link(parent + i, r);
    }
}

__device__ inline int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(256) void dbscan_label_kernel(const long long *__restrict__ offs,
                                                           const int *__restrict__ ind,
                                                           const unsigned char *__restrict__ core,
                                                           const int *__restrict__ parent,
                                                           const int *__restrict__ rid,
                                                           int *__restrict__ labels, int n) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (row >= n)
        return;
    if (core[row]) {
        if (lane == 0)
            labels[row] = rid[parent[row]];
        return;
    }
    int best = INT_MAX;
    const long long e = offs[row + 1];
    for (long long t = offs[row] + lane; t < e; t += 64) {
        const int j = ind[t];
        if (core[j])
            best = min(best, rid[parent[j]]);
    }
    best = wave_min(best);
    if (lane == 0)
        labels[row] = best == INT_MAX ? -1 : best;
}

inline unsigned grid_for(int n, int per_block) {
    return (unsigned)std::max(1, std::min((n + per_block - 1) / per_block, 1 << 20));
}

} // namespace

hipError_t launch_dbscan_init(int *parent, int n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(dbscan_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, parent, n);
    return hipGetLastError();
}

hipError_t launch_dbscan_union(const long long *offs, const int *ind, const unsigned char *core,
                               int *parent, int n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(dbscan_union_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, offs,
                       ind, core, parent, n);
    return hipGetLastError();
}

hipError_t launch_dbscan_flatten(const unsigned char *core, int *parent, int *is_root, int n,
                                 hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(dbscan_flatten_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, core,
                       parent, is_root, n);
    return hipGetLastError();
}

hipError_t dbscan_scan_bytes(int n, size_t *bytes) {
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (const int *)nullptr,
                                            (int *)nullptr, n);
}

hipError_t launch_dbscan_scan(void *temp, size_t bytes, const int *in, int *out, int n,
                              hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    return hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, n, s);
}

hipError_t launch_dbscan_label(const long long *offs, const int *ind, const unsigned char *core,
                               const int *parent, const int *rid, int *labels, int n,
                               hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(dbscan_label_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, offs,
                       ind, core, parent, rid, labels, n);
    return hipGetLastError();
}

} // namespace dagpu
