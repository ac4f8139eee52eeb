// Elkan's k-means bounds and the AFK-MC2 seeding chain on the GPU.
//
// Elkan (source/core/clustering/kmeans/kmeans_elkan.hpp:143-566 and
// kmeans_kernels.cpp:66-207 of the reference): every sample keeps an upper bound u on its
// distance to its own centre and k lower bounds l_j on its distance to centre j; the half
// distances between centres cc (k x k) and each centre's half distance to its nearest other
// centre nxt prune the exact distance evaluations. One thread owns one sample: the bound
// tests are data dependent and sparse, so the work is a per-sample loop over the centres
// (read from global memory through L2) with exact double distances, which keeps the
// decisions -- and therefore the labels, also at ties -- those of the sequential method.
// The bounds live in HBM: u (n doubles) and l (n x k doubles).
//
// AFK-MC2 (kmeans.cpp:1667-1811): the Markov chain that picks one more centre runs in one
// thread over draws the host made with the reference's engine. Each draw is one
// generate_canonical value; a chain consumes one per proposal and one per acceptance test
// that needs a uniform, and reports how many it used so that the host advances its engine by
// exactly that many.
#include "launch.hpp"
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

template <typename T>
__device__ __forceinline__ double elkan_dist(const T *__restrict__ x, const double *__restrict__ c,
                                             int d) {
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double t = (double)x[l] - c[l];
        acc = fma(t, t, acc);
    }
    return sqrt(acc);
}

// cc[i*k + j] = |C_i - C_j| / 2 (zero diagonal), one thread per pair.
__global__ __launch_bounds__(256) void elkan_half_kernel(const double *__restrict__ C, int k, int d,
                                                         double *__restrict__ cc) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)k * k)
        return;
    const int i = (int)(t / k), j = (int)(t - (int64_t)i * k);
    if (i == j) {
        cc[t] = 0.0;
        return;
    }
    // The pair is measured as (lower index, higher index) so that cc is exactly symmetric.
    const int a = i < j ? i : j, b = i < j ? j : i;
    const double *ca = C + (int64_t)a * d, *cb = C + (int64_t)b * d;
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double v = ca[l] - cb[l];
        acc = fma(v, v, acc);
    }
    cc[t] = 0.5 * sqrt(acc);
}

// nxt[i] = min_{j != i} cc[i*k + j] (+inf for a single centre), one wave per centre.
__global__ __launch_bounds__(64) void elkan_next_kernel(const double *__restrict__ cc, int k,
                                                        double *__restrict__ nxt) {
    const int i = blockIdx.x;
    double m = __builtin_huge_val();
    for (int j = threadIdx.x; j < k; j += 64)
        if (j != i)
            m = fmin(m, cc[(int64_t)i * k + j]);
    for (int o = 32; o > 0; o >>= 1)
        m = fmin(m, __shfl_xor(m, o, 64));
    if (threadIdx.x == 0)
        nxt[i] = m;
}

// init_elkan (kmeans_elkan.hpp:143-233): centre 0 first; centre j is measured only when the
// best distance so far exceeds the half distance between the best centre and j; a measured
// distance replaces the best only when strictly smaller.
template <typename T>
__global__ __launch_bounds__(256) void elkan_init_kernel(const T *__restrict__ X, int n, int d,
                                                         const double *__restrict__ C, int k,
                                                         const double *__restrict__ cc,
                                                         int *__restrict__ lab,
                                                         double *__restrict__ u,
                                                         double *__restrict__ lb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const T *x = X + i * d;
    double *l = lb + i * k;
    int label = 0;
    double sd = elkan_dist(x, C, d);
    l[0] = sd;
    for (int j = 1; j < k; j++) {
        double v = 0.0;
        if (sd > cc[(int64_t)label * k + j]) {
            v = elkan_dist(x, C + (int64_t)j * d, d);
            if (v < sd) {
                label = j;
                sd = v;
            }
        }
        l[j] = v;
    }
    lab[i] = label;
    u[i] = sd;
}

// elkan_iteration_assign_block (kmeans_elkan.hpp:427-513): a sample whose upper bound exceeds
// its centre's nxt tests every other centre j against l_j and cc; the first test that passes
// tightens u to the exact distance to the own centre; a centre that still passes is measured
// and taken when strictly nearer.
template <typename T>
__global__ __launch_bounds__(256) void elkan_assign_kernel(
    const T *__restrict__ X, int n, int d, const double *__restrict__ C, int k,
    const double *__restrict__ cc, const double *__restrict__ nxt, const int *__restrict__ prev,
    int *__restrict__ lab, double *__restrict__ u, double *__restrict__ lb,
    int *__restrict__ changed) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const T *x = X + i * d;
    double *l = lb + i * k;
    const int label0 = prev[i];
    int label = label0;
    double ub = u[i];
    if (ub > nxt[label]) {
        bool tight = false;
        for (int j = 0; j < k; j++) {
            const double lj = l[j];
            const double h = cc[(int64_t)label * k + j];
            if (j != label && ub > lj && ub > h) {
                if (!tight) {
                    ub = elkan_dist(x, C + (int64_t)label * d, d);
                    l[label] = ub;
                    tight = true;
                }
                if (ub > lj || ub > h) {
                    const double v = elkan_dist(x, C + (int64_t)j * d, d);
                    l[j] = v;
                    if (v < ub) {
                        ub = v;
                        label = j;
                    }
                }
            }
        }
    }
    u[i] = ub;
    lab[i] = label;
    if (label != label0)
        atomicAdd(changed, 1);
}

// shift[c] = |Cprev_c - Cnew_c|, one thread per centre.
__global__ __launch_bounds__(256) void elkan_shift_kernel(const double *__restrict__ Cprev,
                                                          const double *__restrict__ Cnew, int k,
                                                          int d, double *__restrict__ shift) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= k)
        return;
    double acc = 0.0;
    for (int l = 0; l < d; l++) {
        const double v = Cprev[(int64_t)c * d + l] - Cnew[(int64_t)c * d + l];
        acc = fma(v, v, acc);
    }
    shift[c] = sqrt(acc);
}

// The bound update after the centres moved (kmeans_elkan.hpp:372-418,
// kmeans_kernels.cpp:116-131): u += shift of the own centre, l_j = max(l_j - shift_j, 0).
// One workgroup row of threads per sample block; the lower bounds are updated with
// coalesced accesses (thread t of the grid owns consecutive entries of l).
__global__ __launch_bounds__(256) void elkan_bounds_kernel(int n, int k,
                                                           const double *__restrict__ shift,
                                                           const int *__restrict__ lab,
                                                           double *__restrict__ u,
                                                           double *__restrict__ lb) {
    const int64_t total = (int64_t)n * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * 256) {
        const int j = (int)(t % k);
        const double v = lb[t] - shift[j];
        lb[t] = v < 0.0 ? 0.0 : v;
        if (j == 0) {
            const int64_t i = t / k;
            u[i] += shift[lab[i]];
        }
    }
}

// Bounds invalidated after an empty-cluster split (kmeans.cpp:1312-1317): u = DBL_MAX, l = 0.
__global__ __launch_bounds__(256) void elkan_reset_kernel(int n, int k, double *__restrict__ u,
                                                          double *__restrict__ lb) {
    const int64_t total = (int64_t)n * k;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * 256) {
        lb[t] = 0.0;
        if (t < n)
            u[t] = 1.7976931348623157e308;
    }
}

__device__ int lower_bound_cp(const double *__restrict__ cp, int n, double p) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (cp[mid] < p)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// q[i] = w[i] / (2 sum) + 1/(2n), or 1/n when the sum is zero or not finite.
__global__ __launch_bounds__(256) void afk_proposal_kernel(const double *__restrict__ w, int n,
                                                           const double *__restrict__ sum,
                                                           double *__restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const double s = sum[0];
    const bool ok = (s - s == 0.0) && s > 0.0;
    q[i] = ok ? 0.5 * w[i] / s + 1.0 / (2.0 * n) : 1.0 / n;
}

// One chain of m steps (afk_mcmc_init, kmeans.cpp:1667-1811). u holds the canonical draws in
// order; dmin[i] is the squared distance of sample i to its nearest chosen centre. out[0]:
// the chosen sample, out[1]: the draws consumed.
__global__ void afk_chain_kernel(const double *__restrict__ cp, const double *__restrict__ q,
                                 const double *__restrict__ dmin, int n, int m,
                                 const double *__restrict__ u, int *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    int used = 0;
    auto draw = [&]() {
        const int r = lower_bound_cp(cp, n, u[used++]);
        return r < n ? r : n - 1;
    };
    int x = draw();
    double dx = dmin[x];
    for (int s = 0; s < m; s++) {
        const int y = draw();
        const double dy = dmin[y];
        const double dxqy = dx * q[y];
        bool accept = dxqy == 0.0;
        if (!accept)
            accept = (dy * q[x]) / dxqy > u[used++];
        if (accept) {
            x = y;
            dx = dy;
        }
    }
    out[0] = x;
    out[1] = used;
}

unsigned blocks_for(int64_t n, int64_t cap = 65536) {
    int64_t g = (n + 255) / 256;
    if (g > cap)
        g = cap;
    return (unsigned)(g < 1 ? 1 : g);
}

} // namespace

hipError_t launch_elkan_half(const double *C, int k, int d, double *cc, double *nxt,
                             hipStream_t s) {
    hipLaunchKernelGGL(elkan_half_kernel, dim3(blocks_for((int64_t)k * k, 1 << 30)), dim3(256), 0,
                       s, C, k, d, cc);
    hipLaunchKernelGGL(elkan_next_kernel, dim3(k), dim3(64), 0, s, cc, k, nxt);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_elkan_init(const T *X, int n, int d, const double *C, int k, const double *cc,
                             int *lab, double *u, double *lb, hipStream_t s) {
    hipLaunchKernelGGL(elkan_init_kernel<T>, dim3(blocks_for(n, 1 << 30)), dim3(256), 0, s, X, n, d,
                       C, k, cc, lab, u, lb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_elkan_assign(const T *X, int n, int d, const double *C, int k, const double *cc,
                               const double *nxt, const int *prev, int *lab, double *u,
                               double *lb, int *changed, hipStream_t s) {
    hipLaunchKernelGGL(elkan_assign_kernel<T>, dim3(blocks_for(n, 1 << 30)), dim3(256), 0, s, X, n,
                       d, C, k, cc, nxt, prev, lab, u, lb, changed);
    return hipGetLastError();
}

hipError_t launch_elkan_bounds(const double *Cprev, const double *Cnew, int n, int k, int d,
                               const int *lab, double *shift, double *u, double *lb,
                               hipStream_t s) {
    hipLaunchKernelGGL(elkan_shift_kernel, dim3(blocks_for(k)), dim3(256), 0, s, Cprev, Cnew, k, d,
                       shift);
    hipLaunchKernelGGL(elkan_bounds_kernel, dim3(blocks_for((int64_t)n * k)), dim3(256), 0, s, n, k,
                       shift, lab, u, lb);
    return hipGetLastError();
}

hipError_t launch_elkan_reset(int n, int k, double *u, double *lb, hipStream_t s) {
    hipLaunchKernelGGL(elkan_reset_kernel, dim3(blocks_for((int64_t)n * k)), dim3(256), 0, s, n, k,
                       u, lb);
    return hipGetLastError();
}

hipError_t launch_afk_proposal(const double *w, int n, const double *sum, double *q,
                               hipStream_t s) {
    hipLaunchKernelGGL(afk_proposal_kernel, dim3(blocks_for(n, 1 << 30)), dim3(256), 0, s, w, n,
                       sum, q);
    return hipGetLastError();
}

hipError_t launch_afk_chain(const double *cp, const double *q, const double *dmin, int n, int m,
                            const double *u, int *out, hipStream_t s) {
    hipLaunchKernelGGL(afk_chain_kernel, dim3(1), dim3(64), 0, s, cp, q, dmin, n, m, u, out);
    return hipGetLastError();
}

template hipError_t launch_elkan_init<float>(const float *, int, int, const double *, int,
                                             const double *, int *, double *, double *,
                                             hipStream_t);
template hipError_t launch_elkan_init<double>(const double *, int, int, const double *, int,
                                              const double *, int *, double *, double *,
                                              hipStream_t);
template hipError_t launch_elkan_assign<float>(const float *, int, int, const double *, int,
                                               const double *, const double *, const int *, int *,
                                               double *, double *, int *, hipStream_t);
template hipError_t launch_elkan_assign<double>(const double *, int, int, const double *, int,
                                                const double *, const double *, const int *, int *,
                                                double *, double *, int *, hipStream_t);

} // namespace dagpu
