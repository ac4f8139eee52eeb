// Radius search over materialised distance-key tiles for gfx950: per query, the training
// rows whose key is at most a threshold (the reference's brute-force radius search,
// nearest_neighbors.cpp:2080-2215: D <= radius, or <= radius^2 on squared euclidean keys).
//
// A row of keys is cut into segments of radius_segment_len(n) columns, each swept by one
// wave 64 entries at a time: a ballot of the compares gives the segment's count, and in the
// fill pass the position of each hit (the segment's offset, the running count and the hits
// of lower lanes), so every query's neighbours come out in ascending training index without
// atomics. Optional ordering by distance is a stable
// segmented radix sort of (key, index) pairs, which keeps ties in index order. With
// self0 >= 0 the entry of row r at column self0 + r is skipped (a point is not its own
// neighbour when a set is compared with itself).
#include "launch.hpp"
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace dagpu {

namespace {

// Per-row counts of a row segment from the per-tile counts of the key-tile kernel: the
// segment's tiles_per_seg 64-column tiles (the last segment may hold fewer).
__global__ void radius_segsum_kernel(const int *__restrict__ tilecnt, int rows, int ntiles,
                                     int tiles_per_seg, int nseg, int *__restrict__ segcnt) {
    const int64_t total = (int64_t)rows * nseg;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / nseg;
        const int g = (int)(t - r * nseg);
        const int b = g * tiles_per_seg, e = min(ntiles, b + tiles_per_seg);
        int c = 0;
        for (int j = b; j < e; j++)
            c += tilecnt[r * ntiles + j];
        segcnt[t] = c;
    }
}

// A wave sweeps one segment of one row, four 64-entry steps per iteration so that four
// loads are in flight; the grid is (row groups of 4) x (segments).
template <typename U>
__global__ __launch_bounds__(256) void radius_fill_kernel(const U *__restrict__ D, int64_t ldd,
                                                          int rows, int n, U thr, int self0,
                                                          int seg, int nseg,
                                                          const int *__restrict__ segoff,
                                                          int *__restrict__ ind,
                                                          U *__restrict__ key) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (row >= rows)
        return;
    const int c0 = (int)blockIdx.y * seg, c1 = min(n, c0 + seg);
    const U *d = D + (int64_t)row * ldd;
    const unsigned long long below = (1ull << lane) - 1;
    const int self = self0 >= 0 ? self0 + row : -1;
    int pos = segoff[(int64_t)row * nseg + blockIdx.y];
    for (int j0 = c0; j0 < c1; j0 += 256) {
        U v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int j = j0 + u * 64 + lane;
            v[u] = j < c1 ? d[j] : thr;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int j = j0 + u * 64 + lane;
            const bool hit = j < c1 && v[u] <= thr && j != self;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(hit);
            if (hit) {
                const int p = pos + __popcll(m & below);
                ind[p] = j;
                if (key)
                    key[p] = v[u];
            }
            pos += __popcll(m);
        }
    }
}

template <typename U>
__global__ void radius_sqrt_kernel(U *__restrict__ v, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        v[t] = sqrt(v[t]);
}

} // namespace

int radius_segment_len(int n) {
    // 2,048 columns a segment, longer when the segment count would pass the grid's y limit.
    int seg = 2048;
    while ((int64_t)(n + seg - 1) / seg > 65535)
        seg *= 2;
    return seg;
}

hipError_t launch_radius_segsum(const int *tilecnt, int rows, int n, int *segcnt,
                                hipStream_t s) {
    if (rows <= 0 || n <= 0)
        return hipSuccess;
    const int seg = radius_segment_len(n), nseg = (n + seg - 1) / seg;
    const int ntiles = (n + 63) / 64;
    const int64_t total = (int64_t)rows * nseg;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256,
                                                                              4096));
    hipLaunchKernelGGL(radius_segsum_kernel, dim3(blocks), dim3(256), 0, s, tilecnt, rows, ntiles,
                       seg / 64, nseg, segcnt);
    return hipGetLastError();
}

template <typename U>
hipError_t launch_radius_fill(const U *D, int64_t ldd, int rows, int n, U thr, int self0,
                              const int *segoff, int *ind, U *key, hipStream_t s) {
    if (rows <= 0 || n <= 0)
        return hipSuccess;
    const int seg = radius_segment_len(n), nseg = (n + seg - 1) / seg;
    hipLaunchKernelGGL(radius_fill_kernel<U>, dim3((rows + 3) / 4, nseg), dim3(256), 0, s, D,
                       ldd, rows, n, thr, self0, seg, nseg, segoff, ind, key);
    return hipGetLastError();
}

template <typename U>
hipError_t radius_sort_bytes(int total, int rows, size_t *bytes) {
    return hipcub::DeviceSegmentedRadixSort::SortPairs(
        nullptr, *bytes, (const U *)nullptr, (U *)nullptr, (const int *)nullptr, (int *)nullptr,
        total, rows, (const int *)nullptr, (const int *)nullptr, 0, (int)sizeof(U) * 8);
}

template <typename U>
hipError_t launch_radius_sort(void *temp, size_t bytes, const U *key_in, U *key_out,
                              const int *ind_in, int *ind_out, int total, int rows,
                              const int *offs, hipStream_t s) {
    if (total <= 0)
        return hipSuccess;
    return hipcub::DeviceSegmentedRadixSort::SortPairs(temp, bytes, key_in, key_out, ind_in,
                                                       ind_out, total, rows, offs, offs + 1, 0,
                                                       (int)sizeof(U) * 8, s);
}

template <typename U> hipError_t launch_radius_sqrt(U *v, int64_t n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(radius_sqrt_kernel<U>, dim3((unsigned)blocks), dim3(256), 0, s, v, n);
    return hipGetLastError();
}

#define DA_RADIUS_INST(U)                                                                       \
    template hipError_t launch_radius_fill<U>(const U *, int64_t, int, int, U, int,             \
                                              const int *, int *, U *, hipStream_t);            \
    template hipError_t radius_sort_bytes<U>(int, int, size_t *);                               \
    template hipError_t launch_radius_sort<U>(void *, size_t, const U *, U *, const int *,      \
                                              int *, int, int, const int *, hipStream_t);      \
    template hipError_t launch_radius_sqrt<U>(U *, int64_t, hipStream_t);
DA_RADIUS_INST(float)
DA_RADIUS_INST(double)
#undef DA_RADIUS_INST

} // namespace dagpu
