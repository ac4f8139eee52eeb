// k-means kernels for gfx950: exact assignment, deterministic centroid update, inertia,
// transform and the k-means++ sampling helpers.
//
// Assignment follows lloyd_iteration_kernel (source/core/clustering/kmeans/
// kmeans_kernels.cpp:209-680 of the reference): argmin over centres with ties to the
// lowest centre index (the scalar kernel's rule), but on exact double-precision squared
// distances instead of the GEMM form.
//
// Centroid sums (kmeans_lloyd.hpp:274-361) are accumulated in 64-bit fixed point: every
// value of feature l is scaled by 2^e[l] and split into an integer part and a remainder
// scaled by a further 2^lshift. Integer addition is associative, so the sums are
// identical whatever the order in which points are visited or workgroups are scheduled,
// and the totals are exact to about 2^-100 of the data range. The exponents are chosen on
// the host from the column maxima so that no partial or total can overflow.
//
// Spherical k-means (the "distance" option "cosine" of the reference) shares these kernels:
// the assignment takes the argmax of x.c (the reference's GEMM form with zero centre norms,
// kmeans_lloyd.hpp:131-146), the sums are of x or of x/|x| (kmeans_lloyd.hpp:279-300) and
// the centres are normalised rather than divided by the counts (:363-405).
#include "launch.hpp"
#include <type_traits>
#include <algorithm>
#include <climits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

constexpr int CS_ELEMS = 4096; // centres staged in LDS per chunk (doubles)

// One term of the assignment key: (x - c)^2 for the Euclidean k-means, -x c for the
// spherical one (the key -x.c is minimised, so the largest x.c wins with the same ties).
template <bool SPH>
__device__ __forceinline__ double km_term(double x, double c, double acc) {
    if constexpr (SPH) {
        return fma(-x, c, acc);
    } else {
        const double t = x - c;
        return fma(t, t, acc);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void col_absmax_kernel(const T *__restrict__ X, int n, int d,
                                                         double *out) {
    // Thread t takes feature t % d of every (256/d)-th sample of the workgroup's stride
    // (d <= 256), or features t, t + 256, ... of every sample (d > 256); maxima meet in LDS
    // and each workgroup issues one atomic per feature. Non-negative doubles order like their
    // bit patterns.
    __shared__ double m[256];
    auto fold = [](double mx, T x) {
        const double v = fabs((double)x);
        return (v == v && v > mx) ? v : mx;
    };
    if (d <= 256) {
        const int spp = 256 / d, sub = (int)threadIdx.x / d, l = (int)threadIdx.x - sub * d;
        double mx = 0.0;
        if (sub < spp)
            for (int64_t i = (int64_t)blockIdx.x * spp + sub; i < n;
                 i += (int64_t)gridDim.x * spp)
                mx = fold(mx, X[i * d + l]);
        m[threadIdx.x] = mx;
        __syncthreads();
        if ((int)threadIdx.x < d) {
            double r = m[threadIdx.x];
            for (int t = 1; t < spp; t++)
                r = fmax(r, m[threadIdx.x + t * d]);
            if (r > 0.0)
                atomicMax(reinterpret_cast<unsigned long long *>(out + threadIdx.x),
                          (unsigned long long)__double_as_longlong(r));
        }
    } else {
        for (int l = threadIdx.x; l < d; l += 256) {
            double mx = 0.0;
            for (int64_t i = blockIdx.x; i < n; i += gridDim.x)
                mx = fold(mx, X[i * d + l]);
            if (mx > 0.0)
                atomicMax(reinterpret_cast<unsigned long long *>(out + l),
                          (unsigned long long)__double_as_longlong(mx));
        }
    }
}

template <typename T, int DMAX, bool SPH>
__global__ __launch_bounds__(256) void kmeans_assign_kernel(
    const T *__restrict__ X, int n, int d, const double *__restrict__ C, int k,
    int *__restrict__ labels, const int *__restrict__ prev, int *__restrict__ changed,
    double *__restrict__ mind) {
    __shared__ double Cs[CS_ELEMS];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    // The sample stays in registers in its own type (widened exactly to double when used),
    // which halves the registers of a float sample.
    T x[DMAX > 0 ? DMAX : 1];
    if constexpr (DMAX > 0) {
#pragma unroll
        for (int l = 0; l < DMAX; l++)
            x[l] = (valid && l < d) ? X[(int64_t)i * d + l] : (T)0;
    }
    double best = __builtin_huge_val();
    int bl = 0;
    if (DMAX == 0 && d > CS_ELEMS) {
        // A centre does not fit the LDS stage: read the centres from global memory (L2).
        if (valid) {
            for (int c = 0; c < k; c++) {
                const double *p = C + (int64_t)c * d;
                double acc = 0.0;
                for (int l = 0; l < d; l++)
                    acc = km_term<SPH>((double)X[(int64_t)i * d + l], p[l], acc);
                if (acc < best) {
                    best = acc;
                    bl = c;
                }
            }
            labels[i] = bl;
            if (mind)
                mind[i] = best;
            if (prev && prev[i] != bl)
                atomicAdd(changed, 1);
        }
        return;
    }
    const int cc = CS_ELEMS / d;
    for (int c0 = 0; c0 < k; c0 += cc) {
        const int nc = (k - c0) < cc ? (k - c0) : cc;
        __syncthreads();
        for (int t = threadIdx.x; t < nc * d; t += 256)
            Cs[t] = C[(int64_t)c0 * d + t];
        __syncthreads();
        if (!valid)
            continue;
        int c = 0;
        if constexpr (DMAX > 0) {
            // Four centres at a time: each sample element is widened once for four
            // distances. Each distance keeps its own accumulation order and the centres are
            // compared in ascending order, so the result is that of one centre at a time.
            for (; c + 4 <= nc; c += 4) {
                const double *p = Cs + c * d;
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int l = 0; l < DMAX; l++) {
                    if (l < d) {
                        const double xl = (double)x[l];
                        a0 = km_term<SPH>(xl, p[l], a0);
                        a1 = km_term<SPH>(xl, p[d + l], a1);
                        a2 = km_term<SPH>(xl, p[2 * d + l], a2);
                        a3 = km_term<SPH>(xl, p[3 * d + l], a3);
                    }
                }
                if (a0 < best) {
                    best = a0;
                    bl = c0 + c;
                }
                if (a1 < best) {
                    best = a1;
                    bl = c0 + c + 1;
                }
                if (a2 < best) {
                    best = a2;
                    bl = c0 + c + 2;
                }
                if (a3 < best) {
                    best = a3;
                    bl = c0 + c + 3;
                }
            }
        }
        for (; c < nc; c++) {
            double acc = 0.0;
            if constexpr (DMAX > 0) {
#pragma unroll
                for (int l = 0; l < DMAX; l++) {
                    if (l < d)
                        acc = km_term<SPH>((double)x[l], Cs[c * d + l], acc);
                }
            } else {
                for (int l = 0; l < d; l++)
                    acc = km_term<SPH>((double)X[(int64_t)i * d + l], Cs[c * d + l], acc);
            }
            if (acc < best) {
                best = acc;
                bl = c0 + c;
            }
        }
    }
    if (valid) {
        labels[i] = bl;
        if (mind)
            mind[i] = best;
        if (prev && prev[i] != bl)
            atomicAdd(changed, 1);
    }
}

// Exclusive prefix sum of v[0..m) in LDS by a 1024-thread workgroup: each thread scans a
// contiguous chunk, the chunk totals are scanned across the waves, then added back.
__device__ void block_exclusive_scan(int *v, int m, int *wsum) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (m + 1023) / 1024;
    const int lo = t * per, hi = min(m, lo + per);
    int s = 0;
    for (int j = lo; j < hi; j++)
        s += v[j];
    // Inclusive scan of the chunk totals within the wave.
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int j = 0; j < 16; j++) {
            const int c = wsum[j];
            wsum[j] = run;
            run += c;
        }
    }
    __syncthreads();
    int run = wsum[w] + x - s;
    for (int j = lo; j < hi; j++) {
        const int c = v[j];
        v[j] = run;
        run += c;
    }
    __syncthreads();
}

// Fixed-point totals per (cluster, feature) pair of one pass: the clusters [c0, c0+kc) and the
// features [l0, l0+dl) (kc * dl <= 1024 * P pairs). The grid's workgroups take the chunks of
// ppb points in turn (chunk b, b + gridDim.x, ...): a chunk's points are counting-sorted by
// label in LDS (two 32-bit LDS atomics per point), then thread t adds the values of its pairs
// p = t + 1024 u (u < P) in registers, across all of its workgroup's chunks. Each workgroup
// finally writes its totals, wpart[(b * kc * dl + p) * 2 + w] (w = 0: the value scaled by
// 2^e[l] and rounded to int64, w = 1: the remainder scaled by 2^(e[l] + lshift)), and its
// cluster sizes wcnt[b * kc + c]: plain stores, no device-wide atomics. Integer sums are
// exact, so neither the order of the points nor that of the workgroups changes the result.
// With winv, point i contributes x_i * winv[i] (the normalised samples of spherical k-means).
template <typename T, int P>
__global__ __launch_bounds__(1024) void kmeans_partial_kernel(
    const T *__restrict__ X, int n, int d, const int *__restrict__ labels, int c0, int kc,
    int l0, int dl, const int *__restrict__ escale, int ppb, long long *__restrict__ wpart,
    int *__restrict__ wcnt, const double *__restrict__ winv) {
    extern __shared__ int sh[];
    int *off = sh;                 // kc + 1 offsets
    int *fill = off + kc + 1;      // kc fill counters
    int *list = fill + kc;         // ppb point offsets, sorted by label
    __shared__ int wsum[16];
    const int pairs = kc * dl;
    const int lshift = escale[d];
    long long H[P], L[P];
    int cnt[P];
#pragma unroll
    for (int u = 0; u < P; u++) {
        H[u] = 0;
        L[u] = 0;
        cnt[u] = 0;
    }
    const int nchunks = (int)(((int64_t)n + ppb - 1) / ppb);
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int64_t start = (int64_t)ch * ppb;
        const int m = (int)min((int64_t)ppb, (int64_t)n - start);
        for (int t = threadIdx.x; t <= kc; t += 1024)
            off[t] = 0;
        for (int t = threadIdx.x; t < kc; t += 1024)
            fill[t] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += 1024) {
            const int lab = labels[start + i] - c0;
            if (lab >= 0 && lab < kc)
                atomicAdd(&off[lab], 1);
        }
        __syncthreads();
        block_exclusive_scan(off, kc + 1, wsum);
        for (int i = threadIdx.x; i < m; i += 1024) {
            const int lab = labels[start + i] - c0;
            if (lab >= 0 && lab < kc)
                list[off[lab] + atomicAdd(&fill[lab], 1)] = i;
        }
        __syncthreads();
        const double *wv = winv ? winv + start : nullptr;
#pragma unroll
        for (int u = 0; u < P; u++) {
            const int p = (int)threadIdx.x + 1024 * u;
            if (p >= pairs)
                continue;
            const int c = p / dl, l = l0 + (p - c * dl);
            const int j0 = off[c], j1 = off[c + 1];
            cnt[u] += j1 - j0;
            const int el = escale[l];
            long long h = 0, lo = 0;
            auto add = [&](T x, int j) {
                const double v = wv ? (double)x * wv[list[j]] : (double)x;
                if (v == v && v - v == 0.0) {
                    const double sv = ldexp(v, el);
                    const double hv = rint(sv);
                    h += (long long)hv;
                    lo += (long long)rint(ldexp(sv - hv, lshift));
                }
            };
            const T *xs = X + start * d + l;
            int j = j0;
            // Eight independent gathers in flight per thread.
            for (; j + 8 <= j1; j += 8) {
                T v[8];
#pragma unroll
                for (int e = 0; e < 8; e++)
                    v[e] = xs[(int64_t)list[j + e] * d];
#pragma unroll
                for (int e = 0; e < 8; e++)
                    add(v[e], j + e);
            }
            for (; j < j1; j++)
                add(xs[(int64_t)list[j] * d], j);
            H[u] += h;
            L[u] += lo;
        }
        __syncthreads(); // the next chunk reuses the LDS lists
    }
    long long *wp = wpart + (int64_t)blockIdx.x * pairs * 2;
#pragma unroll
    for (int u = 0; u < P; u++) {
        const int p = (int)threadIdx.x + 1024 * u;
        if (p >= pairs)
            continue;
        wp[2 * (int64_t)p] = H[u];
        wp[2 * (int64_t)p + 1] = L[u];
        if (p % dl == 0)
            wcnt[(int64_t)blockIdx.x * kc + p / dl] = cnt[u];
    }
}

// One workgroup per cluster of the pass: the totals of the nb workgroups' partials (summed as
// integers, so in any order), then the centre: the sum divided by the cluster size (an empty
// cluster keeps count 1 and a zero sum, i.e. the origin: scale_current_cluster_centres,
// kmeans_lloyd.hpp:363-373), or for spherical k-means (sph) the sum itself, normalised
// afterwards.
__global__ __launch_bounds__(256) void kmeans_reduce_kernel(
    const long long *__restrict__ wpart, const int *__restrict__ wcnt, int nb, int c0, int kc,
    int l0, int dl, int d, const int *__restrict__ escale, double *__restrict__ C,
    int *__restrict__ counts, int sph) {
    __shared__ long long sh_h[256], sh_l[256];
    __shared__ int sh_cnt;
    const int c = blockIdx.x, t = threadIdx.x;
    const int pairs = kc * dl;
    if (t < 64) {
        int s = 0;
        for (int b = t; b < nb; b += 64)
            s += wcnt[(int64_t)b * kc + c];
        for (int o = 32; o > 0; o >>= 1)
            s += __shfl_xor(s, o, 64);
        if (t == 0)
            sh_cnt = s;
    }
    // Threads are (feature, group) pairs: the groups split the workgroups' partials.
    const int width = dl < 256 ? dl : 256;
    const int groups = 256 / width;
    const int lw = t % width, grp = t / width;
    const int lshift = escale[d];
    for (int lb = 0; lb < dl; lb += width) {
        const int li = lb + lw;
        long long h = 0, lo = 0;
        if (grp < groups && li < dl) {
            const int64_t p = (int64_t)c * dl + li;
            for (int b = grp; b < nb; b += groups) {
                h += wpart[((int64_t)b * pairs + p) * 2];
                lo += wpart[((int64_t)b * pairs + p) * 2 + 1];
            }
        }
        sh_h[t] = h;
        sh_l[t] = lo;
        __syncthreads();
        if (grp == 0 && li < dl) {
            for (int g = 1; g < groups; g++) {
                h += sh_h[t + g * width];
                lo += sh_l[t + g * width];
            }
            const int l = l0 + li;
            const int ex = escale[l];
            const double sum = ldexp((double)h, -ex) + ldexp((double)lo, -(ex + lshift));
            const int count = sh_cnt;
            C[(int64_t)(c0 + c) * d + l] = sph ? sum : sum / (double)(count > 0 ? count : 1);
        }
        __syncthreads();
    }
    if (t == 0)
        counts[c0 + c] = sh_cnt;
}

__global__ void zero_words_kernel(long long *__restrict__ p, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = 0;
}

__global__ __launch_bounds__(1024) void kmeans_shift_kernel(const double *__restrict__ a,
                                                            const double *__restrict__ b,
                                                            int len, double *out) {
    __shared__ double s0[1024], s1[1024];
    double x0 = 0.0, x1 = 0.0;
    for (int t = threadIdx.x; t < len; t += 1024) {
        const double va = a[t];
        const double dv = va - b[t];
        x0 = fma(va, va, x0);
        x1 = fma(dv, dv, x1);
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

// Block sums over a fixed partition: block b sums elements b, b+G, b+2G... per thread.
__global__ __launch_bounds__(256) void sum_partial_kernel(const double *__restrict__ v, int n,
                                                          double *tmp) {
    __shared__ double s[256];
    double x = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        x += v[t];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        tmp[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void sum_final_kernel(const double *__restrict__ tmp, int nb,
                                                        double *out) {
    __shared__ double s[256];
    double x = 0.0;
    for (int t = threadIdx.x; t < nb; t += 256)
        x += tmp[t];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = s[0];
}

template <typename T>
__global__ void convert_kernel(const T *__restrict__ src, int64_t n, double *__restrict__ dst) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = (double)src[e];
}
template <typename T>
__global__ void convert_back_kernel(const double *__restrict__ src, int64_t n, T *__restrict__ dst) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = (T)src[e];
}
template <typename T>
__global__ void gather_rows_kernel(const T *__restrict__ X, int d, const int *__restrict__ idx,
                                   int m, double *__restrict__ C) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * d)
        return;
    const int64_t r = e / d, l = e - r * d;
    C[e] = (double)X[(int64_t)idx[r] * d + l];
}

template <typename T>
__device__ __forceinline__ double sqdist_rows(const T *__restrict__ X, int d, int64_t i,
                                              const double *__restrict__ cs) {
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double t = (double)X[i * d + l] - cs[l];
        acc = fma(t, t, acc);
    }
    return acc;
}

// Squared distance of rows i and j of X, widened to double (the row j read from global
// memory: used where a row does not fit the LDS stage).
template <typename T>
__device__ __forceinline__ double sqdist_xx(const T *__restrict__ X, int d, int64_t i, int64_t j) {
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double t = (double)X[i * d + l] - (double)X[j * d + l];
        acc = fma(t, t, acc);
    }
    return acc;
}

// Dot products of the same operands (spherical k-means).
template <typename T>
__device__ __forceinline__ double dot_rows(const T *__restrict__ X, int d, int64_t i,
                                           const double *__restrict__ cs) {
    double acc = 0.0;
    for (int l = 0; l < d; l++)
        acc = fma((double)X[i * d + l], cs[l], acc);
    return acc;
}
template <typename T>
__device__ __forceinline__ double dot_xx(const T *__restrict__ X, int d, int64_t i, int64_t j) {
    double acc = 0.0;
    for (int l = 0; l < d; l++)
        acc = fma((double)X[i * d + l], (double)X[j * d + l], acc);
    return acc;
}

// The distance of samples i and j used by the initialisations: the squared Euclidean
// distance, or for spherical k-means (sph) max(0, 1 - x_i.x_j inv_i inv_j) with the inverse
// norms inv (compute_distances_to_point, kmeans.cpp:1488-1522; without inv the samples are
// taken as unit rows). The _lds form reads row j staged in LDS (cs), the _global form from X.
struct KmGeom {
    int sph;
    const double *inv;
};
__device__ __forceinline__ double cosine_dist(double dot, KmGeom g, int64_t i, int64_t j) {
    if (g.inv)
        dot = dot * g.inv[i] * g.inv[j];
    const double v = 1.0 - dot;
    return v > 0.0 ? v : 0.0;
}
template <typename T>
__device__ __forceinline__ double sample_dist_lds(const T *__restrict__ X, int d, int64_t i,
                                                  int64_t j, const double *cs, KmGeom g) {
    return g.sph ? cosine_dist(dot_rows(X, d, i, cs), g, i, j) : sqdist_rows(X, d, i, cs);
}
template <typename T>
__device__ __forceinline__ double sample_dist_global(const T *__restrict__ X, int d, int64_t i,
                                                     int64_t j, KmGeom g) {
    return g.sph ? cosine_dist(dot_xx(X, d, i, j), g, i, j) : sqdist_xx(X, d, i, j);
}

// Rows of at most ROW_STAGE_MAX features are staged in LDS by the distance passes of the
// initialisations; longer rows are read from global memory.
constexpr int ROW_STAGE_MAX = 8192;

template <typename T>
__global__ __launch_bounds__(256) void dist_to_row_kernel(const T *__restrict__ X, int n, int d,
                                                          int ci, double *__restrict__ w,
                                                          int mode, KmGeom g) {
    extern __shared__ double cs[];
    const bool staged = d <= ROW_STAGE_MAX;
    if (staged) {
        for (int t = threadIdx.x; t < d; t += 256)
            cs[t] = (double)X[(int64_t)ci * d + t];
        __syncthreads();
    }
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    double v = staged ? sample_dist_lds(X, d, i, ci, cs, g) : sample_dist_global(X, d, i, ci, g);
    if (mode == 1) {
        // min update: max(0, min(w, v))
        const double o = w[i];
        v = o < v ? o : v;
        v = v > 0.0 ? v : 0.0;
    } else {
        if (!(v - v == 0.0) || v < 0.0)
            v = 0.0;
    }
    if (i == ci)
        v = 0.0;
    w[i] = v;
}

constexpr int SCAN_BLOCKS = 1024;

// Phase 1: per-block sums of w/total over a contiguous chunk.
__global__ __launch_bounds__(256) void scan_blocksum_kernel(const double *__restrict__ w, int n,
                                                            const double *__restrict__ total,
                                                            double *__restrict__ bsum, int chunk) {
    __shared__ double s[256];
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const double tot = total[0];
    double x = 0.0;
    // Thread t handles the contiguous sub-chunk [b0 + t*sub, b0 + (t+1)*sub).
    const int sub = (chunk + 255) / 256;
    for (int j = 0; j < sub; j++) {
        const int64_t e = b0 + (int64_t)threadIdx.x * sub + j;
        if (e < n && e < b0 + chunk)
            x += w[e] / tot;
    }
    s[threadIdx.x] = x;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m)
            s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bsum[blockIdx.x] = s[0];
}

// Phase 2: exclusive scan of the nb <= 1024 block sums by one workgroup (Hillis-Steele in LDS,
// a fixed order of additions).
__global__ __launch_bounds__(1024) void scan_blockoffsets_kernel(double *bsum, int nb) {
    __shared__ double s[1024];
    const int t = threadIdx.x;
    s[t] = t < nb ? bsum[t] : 0.0;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const double v = t >= off ? s[t - off] : 0.0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    if (t < nb)
        bsum[t] = t > 0 ? s[t - 1] : 0.0;
}

// Phase 3: each block scans its chunk with its offset.
__global__ __launch_bounds__(256) void scan_final_kernel(const double *__restrict__ w, int n,
                                                         const double *__restrict__ total,
                                                         const double *__restrict__ boff,
                                                         double *__restrict__ cp, int chunk) {
    __shared__ double s[256];
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const double tot = total[0];
    const int sub = (chunk + 255) / 256;
    double x = 0.0;
    for (int j = 0; j < sub; j++) {
        const int64_t e = b0 + (int64_t)threadIdx.x * sub + j;
        if (e < n && e < b0 + chunk)
            x += w[e] / tot;
    }
    s[threadIdx.x] = x;
    __syncthreads();
    // Exclusive scan of the thread sums (Hillis-Steele on a copy, fixed order).
    for (int off = 1; off < 256; off <<= 1) {
        double v = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0.0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    double run = boff[blockIdx.x] + ((int)threadIdx.x > 0 ? s[threadIdx.x - 1] : 0.0);
    for (int j = 0; j < sub; j++) {
        const int64_t e = b0 + (int64_t)threadIdx.x * sub + j;
        if (e < n && e < b0 + chunk) {
            run += w[e] / tot;
            cp[e] = (e == n - 1) ? 1.0 : run;
        }
    }
}

__global__ void lower_bound_kernel(const double *__restrict__ cp, int n, const double *__restrict__ u,
                                   int nt, int *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt)
        return;
    const double p = u[t];
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (cp[mid] < p)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[t] = lo < n ? lo : n - 1;
}

constexpr int NT_MAX = 32;

template <typename T>
__global__ __launch_bounds__(256) void candidate_cost_kernel(const T *__restrict__ X, int n, int d,
                                                             const double *__restrict__ w,
                                                             const int *__restrict__ cand, int nt,
                                                             double *__restrict__ part, int nb,
                                                             KmGeom g) {
    extern __shared__ double cs[]; // nt x d when staged
    __shared__ double red[256];
    const bool staged = (int64_t)nt * d <= ROW_STAGE_MAX;
    if (staged) {
        for (int t = threadIdx.x; t < nt * d; t += 256) {
            const int c = t / d, l = t - c * d;
            cs[t] = (double)X[(int64_t)cand[c] * d + l];
        }
        __syncthreads();
    }
    double acc[NT_MAX];
#pragma unroll
    for (int c = 0; c < NT_MAX; c++)
        acc[c] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double wi = w[i];
#pragma unroll
        for (int c = 0; c < NT_MAX; c++) {
            if (c < nt) {
                double v = staged ? sample_dist_lds(X, d, i, (int64_t)cand[c], cs + c * d, g)
                                  : sample_dist_global(X, d, i, (int64_t)cand[c], g);
                v = wi < v ? wi : v;
                acc[c] += v > 0.0 ? v : 0.0;
            }
        }
    }
    for (int c = 0; c < nt; c++) {
        red[threadIdx.x] = acc[c];
        __syncthreads();
        for (int m = 128; m > 0; m >>= 1) {
            if ((int)threadIdx.x < m)
                red[threadIdx.x] += red[threadIdx.x + m];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            part[(int64_t)c * nb + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// The same costs for rows of exactly D features (D * sizeof(T) a multiple of 16 bytes): each
// sample's row is read once into registers with 16-byte loads and every candidate's distance
// is formed from there, with the operations and their order of candidate_cost_kernel, so the
// costs are bitwise the same. The candidates are staged in LDS (nt * D <= ROW_STAGE_MAX).
template <typename T, int D>
__global__ __launch_bounds__(256) void candidate_cost_reg_kernel(
    const T *__restrict__ X, int n, const double *__restrict__ w, const int *__restrict__ cand,
    int nt, double *__restrict__ part, int nb, KmGeom g) {
    extern __shared__ double cs[]; // nt x D
    __shared__ double red[256];
    for (int t = threadIdx.x; t < nt * D; t += 256) {
        const int c = t / D, l = t - c * D;
        cs[t] = (double)X[(int64_t)cand[c] * D + l];
    }
    __syncthreads();
    constexpr int VE = 16 / (int)sizeof(T); // elements per 16-byte load
    static_assert(D % VE == 0, "whole 16-byte loads per row");
    double acc[NT_MAX];
#pragma unroll
    for (int c = 0; c < NT_MAX; c++)
        acc[c] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double xr[D];
        const T *row = X + i * D;
#pragma unroll
        for (int l = 0; l < D; l += VE) {
            if constexpr (sizeof(T) == 4) {
                const float4 v = *reinterpret_cast<const float4 *>(row + l);
                xr[l] = v.x;
                xr[l + 1] = v.y;
                xr[l + 2] = v.z;
                xr[l + 3] = v.w;
            } else {
                const double2 v = *reinterpret_cast<const double2 *>(row + l);
                xr[l] = v.x;
                xr[l + 1] = v.y;
            }
        }
        const double wi = w[i];
#pragma unroll
        for (int c = 0; c < NT_MAX; c++) {
            if (c < nt) {
                const double *cc = cs + c * D;
                double v = 0.0;
                if (g.sph) {
#pragma unroll
                    for (int l = 0; l < D; l++)
                        v = fma(xr[l], cc[l], v);
                    v = cosine_dist(v, g, i, (int64_t)cand[c]);
                } else {
#pragma unroll
                    for (int l = 0; l < D; l++) {
                        const double t = xr[l] - cc[l];
                        v = fma(t, t, v);
                    }
                }
                v = wi < v ? wi : v;
                acc[c] += v > 0.0 ? v : 0.0;
            }
        }
    }
    for (int c = 0; c < nt; c++) {
        red[threadIdx.x] = acc[c];
        __syncthreads();
        for (int m = 128; m > 0; m >>= 1) {
            if ((int)threadIdx.x < m)
                red[threadIdx.x] += red[threadIdx.x + m];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            part[(int64_t)c * nb + blockIdx.x] = red[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void reduce_rows_kernel(const double *__restrict__ part, int nt,
                                                          int nb, double *__restrict__ out) {
    __shared__ double s[256];
    const int t = blockIdx.x;
    double x = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256)
        x += part[(int64_t)t * nb + b];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m)
            s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[t] = s[0];
}

// Distances of rows of X to the centres: Euclidean, or for spherical k-means (sph) the cosine
// distance 1 - clamp(x.c inv_x, -1, 1) (kmeans.cpp:78-105, 821-863), inv_x = 1 without inv.
template <typename T>
__global__ __launch_bounds__(256) void kmeans_transform_kernel(const T *__restrict__ X, int n, int d,
                                                               const double *__restrict__ C, int k,
                                                               T *__restrict__ D, int64_t ldd,
                                                               KmGeom g) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * k)
        return;
    const int64_t i = e / k;
    const int c = (int)(e - i * k);
    if (g.sph) {
        double cosv = dot_rows(X, d, i, C + (int64_t)c * d);
        if (g.inv)
            cosv *= g.inv[i];
        cosv = cosv < -1.0 ? -1.0 : (cosv > 1.0 ? 1.0 : cosv);
        D[i * ldd + c] = (T)(1.0 - cosv);
        return;
    }
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double t = (double)X[i * d + l] - C[(int64_t)c * d + l];
        acc = fma(t, t, acc);
    }
    D[i * ldd + c] = (T)sqrt(acc);
}

// inv[i] = 1/|X_i| (0 for a zero row), the norm in double.
template <typename T>
__global__ __launch_bounds__(256) void row_inv_norms_kernel(const T *__restrict__ X, int n, int d,
                                                            double *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    double s = 0.0;
    for (int l = 0; l < d; l++) {
        const double v = (double)X[i * d + l];
        s = fma(v, v, s);
    }
    const double nrm = sqrt(s);
    inv[i] = nrm > 0.0 ? 1.0 / nrm : 0.0;
}

__host__ inline unsigned grid_for(int64_t n, int block, int64_t cap = 65536) {
    int64_t g = (n + block - 1) / block;
    if (g > cap)
        g = cap;
    if (g < 1)
        g = 1;
    return (unsigned)g;
}

} // namespace

template <typename T>
hipError_t launch_col_absmax(const T *X, int n, int d, double *out, hipStream_t s) {
    const int64_t spp = d <= 256 ? 256 / d : 1;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + spp - 1) / spp, 1024));
    hipLaunchKernelGGL(col_absmax_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, X, n, d,
                       out);
    return hipGetLastError();
}

template <typename T, bool SPH>
static void assign_dispatch(dim3 grid, dim3 block, hipStream_t s, const T *X, int n, int d,
                            const double *C, int k, int *labels, const int *prev, int *changed,
                            double *mind) {
#define DA_KM_ASSIGN(DM)                                                                       \
    hipLaunchKernelGGL((kmeans_assign_kernel<T, DM, SPH>), grid, block, 0, s, X, n, d, C, k,  \
                       labels, prev, changed, mind)
    if (d <= 8)
        DA_KM_ASSIGN(8);
    else if (d <= 16)
        DA_KM_ASSIGN(16);
    else if (d <= 32)
        DA_KM_ASSIGN(32);
    else if (d <= 64)
        DA_KM_ASSIGN(64);
    else if (std::is_same_v<T, float> && d <= 96)
        DA_KM_ASSIGN(96);
    else if (std::is_same_v<T, float> && d <= 128)
        DA_KM_ASSIGN(128);
    else
        DA_KM_ASSIGN(0);
#undef DA_KM_ASSIGN
}

template <typename T>
hipError_t launch_kmeans_assign(const T *X, int n, int d, const double *C, int k, int *labels,
                                const int *prev, int *changed, double *mind, hipStream_t s,
                                int sph) {
    dim3 grid((n + 255) / 256), block(256);
    if (sph)
        assign_dispatch<T, true>(grid, block, s, X, n, d, C, k, labels, prev, changed, mind);
    else
        assign_dispatch<T, false>(grid, block, s, X, n, d, C, k, labels, prev, changed, mind);
    return hipGetLastError();
}

template <typename T, int P>
static hipError_t partial_launch(const T *X, int n, int d, const int *labels, int c0, int kc,
                                 int l0, int dl, const int *escale, int ppb, int grid,
                                 long long *wpart, int *wcnt, hipStream_t s, const double *winv,
                                 size_t lds) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)kmeans_partial_kernel<T, P>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((kmeans_partial_kernel<T, P>), dim3(grid), dim3(1024), lds, s, X, n, d,
                       labels, c0, kc, l0, dl, escale, ppb, wpart, wcnt, winv);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_kmeans_partial_sums(const T *X, int n, int d, const int *labels, int c0,
                                      int kc, int l0, int dl, const int *escale, int ppb,
                                      int grid, long long *wpart, int *wcnt, hipStream_t s,
                                      const double *winv) {
    const size_t lds = kmeans_partial_lds(kc, ppb);
    const int64_t pairs = (int64_t)kc * dl;
    if (lds > 160 * 1024 || pairs > KMEANS_PASS_PAIRS || pairs < 1 || grid < 1)
        return hipErrorInvalidValue;
    const int per = (int)((pairs + 1023) / 1024);
    if (per <= 1)
        return partial_launch<T, 1>(X, n, d, labels, c0, kc, l0, dl, escale, ppb, grid, wpart,
                                    wcnt, s, winv, lds);
    if (per <= 2)
        return partial_launch<T, 2>(X, n, d, labels, c0, kc, l0, dl, escale, ppb, grid, wpart,
                                    wcnt, s, winv, lds);
    if (per <= 4)
        return partial_launch<T, 4>(X, n, d, labels, c0, kc, l0, dl, escale, ppb, grid, wpart,
                                    wcnt, s, winv, lds);
    if (per <= 8)
        return partial_launch<T, 8>(X, n, d, labels, c0, kc, l0, dl, escale, ppb, grid, wpart,
                                    wcnt, s, winv, lds);
    return partial_launch<T, 16>(X, n, d, labels, c0, kc, l0, dl, escale, ppb, grid, wpart,
                                 wcnt, s, winv, lds);
}

size_t kmeans_partial_lds(int kc, int ppb) {
    return (size_t)(2 * kc + 1 + ppb) * sizeof(int);
}

hipError_t launch_kmeans_reduce(const long long *wpart, const int *wcnt, int nb, int c0, int kc,
                                int l0, int dl, int d, const int *escale, double *C,
                                int *counts, hipStream_t s, int sph) {
    hipLaunchKernelGGL(kmeans_reduce_kernel, dim3(kc), dim3(256), 0, s, wpart, wcnt, nb, c0, kc,
                       l0, dl, d, escale, C, counts, sph);
    return hipGetLastError();
}

hipError_t launch_zero_words(long long *p, int64_t n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_kmeans_shift(const double *Cprev, const double *Cnew, int len, double *out,
                               hipStream_t s) {
    hipLaunchKernelGGL(kmeans_shift_kernel, dim3(1), dim3(1024), 0, s, Cprev, Cnew, len, out);
    return hipGetLastError();
}

hipError_t launch_sum_double(const double *v, int n, double *tmp, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_partial_kernel, dim3(256), dim3(256), 0, s, v, n, tmp);
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(256), 0, s, tmp, 256, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_convert(const T *src, int64_t n, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(convert_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n, dst);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_convert_back(const double *src, int64_t n, T *dst, hipStream_t s) {
    hipLaunchKernelGGL(convert_back_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n,
                       dst);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_gather_rows(const T *X, int d, const int *idx, int m, double *C,
                              hipStream_t s) {
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(grid_for((int64_t)m * d, 256, 1 << 30)),
                       dim3(256), 0, s, X, d, idx, m, C);
    return hipGetLastError();
}

// Dynamic LDS of a pass that stages `elems` doubles: at most ROW_STAGE_MAX (64 KB, within
// the default limit), none when the rows are read from global memory instead.
static size_t row_stage_bytes(int64_t elems) {
    return elems <= ROW_STAGE_MAX ? (size_t)elems * sizeof(double) : 0;
}

template <typename T>
hipError_t launch_dist_to_row(const T *X, int n, int d, int ci, double *w, hipStream_t s, int sph,
                              const double *inv) {
    hipLaunchKernelGGL(dist_to_row_kernel<T>, dim3((n + 255) / 256), dim3(256),
                       row_stage_bytes(d), s, X, n, d, ci, w, 0, KmGeom{sph, inv});
    return hipGetLastError();
}
template <typename T>
hipError_t launch_min_update(const T *X, int n, int d, int ci, double *w, hipStream_t s, int sph,
                             const double *inv) {
    hipLaunchKernelGGL(dist_to_row_kernel<T>, dim3((n + 255) / 256), dim3(256),
                       row_stage_bytes(d), s, X, n, d, ci, w, 1, KmGeom{sph, inv});
    return hipGetLastError();
}

hipError_t launch_normalised_scan(const double *w, int n, const double *total, double *cp,
                                  double *tmp, hipStream_t s) {
    const int nb = SCAN_BLOCKS;
    const int chunk = (n + nb - 1) / nb;
    hipLaunchKernelGGL(scan_blocksum_kernel, dim3(nb), dim3(256), 0, s, w, n, total, tmp, chunk);
    static_assert(SCAN_BLOCKS <= 1024, "one workgroup scans the block sums");
    hipLaunchKernelGGL(scan_blockoffsets_kernel, dim3(1), dim3(1024), 0, s, tmp, nb);
    hipLaunchKernelGGL(scan_final_kernel, dim3(nb), dim3(256), 0, s, w, n, total, tmp, cp, chunk);
    return hipGetLastError();
}

hipError_t launch_lower_bound(const double *cp, int n, const double *u, int nt, int *out,
                              hipStream_t s) {
    hipLaunchKernelGGL(lower_bound_kernel, dim3((nt + 63) / 64), dim3(64), 0, s, cp, n, u, nt, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_candidate_cost(const T *X, int n, int d, const double *w, const int *cand,
                                 int nt, double *part, int nb, hipStream_t s, int sph,
                                 const double *inv) {
    if (nt > NT_MAX)
        return hipErrorInvalidValue;
    const KmGeom geo{sph, inv};
    const size_t lds = row_stage_bytes((int64_t)nt * d);
    if ((int64_t)nt * d <= ROW_STAGE_MAX) {
#define DA_CC(DV)                                                                                 \
    if (d == DV) {                                                                                \
        hipLaunchKernelGGL((candidate_cost_reg_kernel<T, DV>), dim3(nb), dim3(256), lds, s, X, n, \
                           w, cand, nt, part, nb, geo);                                           \
        return hipGetLastError();                                                                 \
    }
        DA_CC(8)
        DA_CC(16)
        DA_CC(32)
        DA_CC(64)
#undef DA_CC
    }
    hipLaunchKernelGGL(candidate_cost_kernel<T>, dim3(nb), dim3(256), lds, s, X, n, d, w, cand,
                       nt, part, nb, geo);
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double *part, int nt, int nb, double *out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(nt), dim3(256), 0, s, part, nt, nb, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_kmeans_transform(const T *X, int n, int d, const double *C, int k, T *D,
                                   int64_t ldd, hipStream_t s, int sph, const double *inv) {
    const int64_t total = (int64_t)n * k;
    hipLaunchKernelGGL(kmeans_transform_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, X, n, d, C, k, D, ldd, KmGeom{sph, inv});
    return hipGetLastError();
}

// One workgroup per row of C (k x d doubles): the row scaled to unit length, a zero row left
// as it is (normalize_rows_inplace, da_utils.cpp:807-855).
__global__ __launch_bounds__(256) void normalize_rows_kernel(double *__restrict__ C, int d) {
    __shared__ double red[256];
    double *row = C + (int64_t)blockIdx.x * d;
    double ss = 0.0;
    for (int l = threadIdx.x; l < d; l += 256)
        ss = fma(row[l], row[l], ss);
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m)
            red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    const double nrm = sqrt(red[0]);
    const double inv = nrm == 0.0 ? 1.0 : 1.0 / nrm;
    for (int l = threadIdx.x; l < d; l += 256)
        row[l] *= inv;
}

hipError_t launch_normalize_rows(double *C, int k, int d, hipStream_t s) {
    if (k <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(normalize_rows_kernel, dim3(k), dim3(256), 0, s, C, d);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_row_inv_norms(const T *X, int n, int d, double *inv, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(row_inv_norms_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, X, n, d,
                       inv);
    return hipGetLastError();
}

#define DA_KM_INST(T)                                                                           \
    template hipError_t launch_col_absmax<T>(const T *, int, int, double *, hipStream_t);       \
    template hipError_t launch_kmeans_assign<T>(const T *, int, int, const double *, int, int *, \
                                                const int *, int *, double *, hipStream_t, int); \
    template hipError_t launch_kmeans_partial_sums<T>(const T *, int, int, const int *, int, int, \
                                                      int, int, const int *, int, int,          \
                                                      long long *, int *, hipStream_t,          \
                                                      const double *);                           \
    template hipError_t launch_convert<T>(const T *, int64_t, double *, hipStream_t);           \
    template hipError_t launch_convert_back<T>(const double *, int64_t, T *, hipStream_t);      \
    template hipError_t launch_gather_rows<T>(const T *, int, const int *, int, double *,       \
                                              hipStream_t);                                      \
    template hipError_t launch_dist_to_row<T>(const T *, int, int, int, double *, hipStream_t,  \
                                              int, const double *);                              \
    template hipError_t launch_min_update<T>(const T *, int, int, int, double *, hipStream_t,   \
                                             int, const double *);                               \
    template hipError_t launch_candidate_cost<T>(const T *, int, int, const double *,           \
                                                 const int *, int, double *, int, hipStream_t,  \
                                                 int, const double *);                           \
    template hipError_t launch_kmeans_transform<T>(const T *, int, int, const double *, int, T *, \
                                                   int64_t, hipStream_t, int, const double *);  \
    template hipError_t launch_row_inv_norms<T>(const T *, int, int, double *, hipStream_t);
DA_KM_INST(float)
DA_KM_INST(double)
#undef DA_KM_INST

} // namespace dagpu

namespace dagpu {
namespace {
// The inertia term of each sample: squared distance to its centre, or for spherical k-means
// (sph) 1 - x.c inv_x (compute_current_inertia, kmeans.cpp:1060-1124; inv_x = 1 without inv).
template <typename T>
__global__ void label_sqdist_kernel(const T *__restrict__ X, int n, int d,
                                    const double *__restrict__ C, const int *__restrict__ labels,
                                    double *__restrict__ out, int sph,
                                    const double *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *c = C + (int64_t)labels[i] * d;
    double acc = 0.0;
    if (sph) {
        for (int l = 0; l < d; l++)
            acc = fma((double)X[i * d + l], c[l], acc);
        out[i] = 1.0 - (inv ? acc * inv[i] : acc);
        return;
    }
    for (int l = 0; l < d; l++) {
        const double t = (double)X[i * d + l] - c[l];
        acc = fma(t, t, acc);
    }
    out[i] = acc;
}
} // namespace
template <typename T>
hipError_t launch_label_sqdist(const T *X, int n, int d, const double *C, const int *labels,
                               double *out, hipStream_t s, int sph, const double *inv) {
    hipLaunchKernelGGL(label_sqdist_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, X, n, d, C,
                       labels, out, sph, inv);
    return hipGetLastError();
}
template hipError_t launch_label_sqdist<float>(const float *, int, int, const double *,
                                               const int *, double *, hipStream_t, int,
                                               const double *);
template hipError_t launch_label_sqdist<double>(const double *, int, int, const double *,
                                                const int *, double *, hipStream_t, int,
                                                const double *);
} // namespace dagpu

namespace dagpu {
namespace {
__global__ void round_through_kernel(double *v, int64_t n, int mode) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double x = v[e];
        v[e] = mode == 2 ? (double)(float)(_Float16)(float)x : (double)(float)x;
    }
}
template <typename T>
__global__ void round_data_kernel(const T *__restrict__ src, int64_t n, int mode, T *__restrict__ dst) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const T x = src[e];
        dst[e] = mode == 2 ? (T)(float)(_Float16)(float)x : (T)(float)x;
    }
}
} // namespace
hipError_t launch_round_through(double *v, int64_t n, int mode, hipStream_t s) {
    hipLaunchKernelGGL(round_through_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, v, n, mode);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_round_data(const T *src, int64_t n, int mode, T *dst, hipStream_t s) {
    hipLaunchKernelGGL(round_data_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n, mode,
                       dst);
    return hipGetLastError();
}
template hipError_t launch_round_data<float>(const float *, int64_t, int, float *, hipStream_t);
template hipError_t launch_round_data<double>(const double *, int64_t, int, double *, hipStream_t);
} // namespace dagpu
