// MFMA assignment step of Lloyd's algorithm for gfx950 (fp32 or fp64 data, and a bf16 form).
//
// The reference forms -2 C X^T with a BLAS gemm per block of samples and takes, per sample,
// the argmin of ||c||^2 - 2 x.c (kmeans_lloyd.hpp:185-197, kmeans_kernels.cpp:209-680).
// Here the product runs on v_mfma_f32_16x16x4f32 (v_mfma_f64_16x16x4f64 for fp64 data, and
// v_mfma_f32_16x16x32_bf16 for the bf16 cross term) with the centres as the A operand
// (staged in LDS) and 64 samples per wave held in registers as the B operand.
// The accumulator of centre c starts at -||c||^2/2, so it ends at x.c - ||c||^2/2 and the
// nearest centre is its argmax. Each lane tracks the best and second-best value of its
// sample; a sample whose two best values are closer than the error bound of the MFMA
// arithmetic is listed for an exact double-precision re-check, so the labels equal the exact
// argmin (ties to the lowest centre index).
#include "launch.hpp"
#include <algorithm>
#include <climits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int PT = 4;             // 16-sample tiles per wave
constexpr int WAVES = 4;
constexpr int SB = WAVES * PT * 16; // samples per workgroup (bf16 form)
// LDS bytes a workgroup of the fp32 / fp64 assignment stages at most: centres beyond it are
// taken in chunks, each staged in turn (two workgroups per CU).
constexpr size_t ASSIGN_LDS_MAX = 80 * 1024;

typedef double v4d __attribute__((ext_vector_type(4)));

// MFMA of the assignment per data type: v_mfma_f32_16x16x4f32 returns rows 4(l/16) + i of
// lane l's column, v_mfma_f64_16x16x4f64 rows l/16 + 4i. unit: the rounding unit.
template <typename E> struct AssignOp;
template <> struct AssignOp<float> {
    using acc_t = v4f;
    static constexpr bool strided = false;
    static constexpr double unit = 1.0 / 16777216.0;
    __device__ static acc_t mfma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct AssignOp<double> {
    using acc_t = v4d;
    static constexpr bool strided = true;
    static constexpr double unit = 1.0 / 9007199254740992.0;
    __device__ static acc_t mfma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// KCH consecutive elements (16-byte aligned) into registers as 16-byte loads.
template <typename E, int KCH> __device__ __forceinline__ void load_row(const E *p, E (&v)[KCH]) {
    constexpr int PER = 16 / (int)sizeof(E);
#pragma unroll
    for (int f = 0; f < KCH; f += PER) {
        if constexpr (sizeof(E) == 4) {
            const v4f x = *reinterpret_cast<const v4f *>(p + f);
            v[f] = x[0];
            v[f + 1] = x[1];
            v[f + 2] = x[2];
            v[f + 3] = x[3];
        } else {
            typedef double v2d __attribute__((ext_vector_type(2)));
            const v2d x = *reinterpret_cast<const v2d *>(p + f);
            v[f] = x[0];
            v[f + 1] = x[1];
        }
    }
}

// Centres per staged chunk: all of them when they fit ASSIGN_LDS_MAX, else the most
// multiples of 16 that do.
template <typename E> int assign_chunk(int k, int Kp) {
    const int kpad = (k + 15) / 16 * 16;
    const int fit = (int)(ASSIGN_LDS_MAX / ((size_t)(Kp + 1) * sizeof(E))) / 16 * 16;
    return std::max(16, std::min(kpad, fit));
}

// The assignment in the data's own precision (fp32 or fp64 MFMA). Each wave holds PTT
// tiles of 16 samples in registers as B operands; the centres, A operands, are staged in LDS
// kch at a time (one chunk: staged once per workgroup; more: each chunk staged in turn for
// every tile, with the best and second-best scores carried across chunks).
template <typename E, int KCH, int PTT>
__global__ __launch_bounds__(256) void kmeans_assign_mfma_kernel(
    const E *__restrict__ X, int n, int Kp, const E *__restrict__ CP,
    const E *__restrict__ chalf, int k, int kch, double cmax, int *__restrict__ labels,
    const int *__restrict__ prev, int *__restrict__ changed, int *__restrict__ list_count,
    int *__restrict__ list) {
    using Op = AssignOp<E>;
    using acc_t = typename Op::acc_t;
    extern __shared__ double smem_d[];
    E *Cs = reinterpret_cast<E *>(smem_d); // kch x Kp centres
    E *Hs = Cs + (size_t)kch * Kp;         // -||c||^2/2, padded with -inf
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int kpad = (k + 15) / 16 * 16;
    const int nchunk = (kpad + kch - 1) / kch;
    auto stage = [&](int cb) {
        const int kc = min(kch, kpad - cb);
        for (int t = threadIdx.x; t < kc * Kp; t += 256)
            Cs[t] = cb * Kp + t < k * Kp ? CP[(int64_t)cb * Kp + t] : (E)0;
        for (int t = threadIdx.x; t < kc; t += 256)
            Hs[t] = cb + t < k ? chalf[cb + t] : (E)-__builtin_huge_val();
    };
    if (nchunk == 1) {
        stage(0);
        __syncthreads();
    }
    constexpr int SBW = WAVES * PTT * 16;
    const int ntiles = (n + SBW - 1) / SBW;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int s0 = tile * SBW + w * PTT * 16;
        E b[PTT][KCH];
        E xn[PTT];
#pragma unroll
        for (int pt = 0; pt < PTT; pt++) {
            const int i = s0 + pt * 16 + r;
            E sq = 0;
            load_row<E, KCH>(X + (int64_t)(i < n ? i : 0) * Kp + g * KCH, b[pt]);
#pragma unroll
            for (int f = 0; f < KCH; f++) {
                if (i >= n)
                    b[pt][f] = (E)0;
                sq = fma(b[pt][f], b[pt][f], sq);
            }
            xn[pt] = sq;
        }
        E best1[PTT], best2[PTT];
        int bidx[PTT];
#pragma unroll
        for (int pt = 0; pt < PTT; pt++) {
            best1[pt] = (E)-__builtin_huge_val();
            best2[pt] = (E)-__builtin_huge_val();
            bidx[pt] = INT_MAX;
        }
        for (int ci = 0; ci < nchunk; ci++) {
            const int cb = ci * kch;
            if (nchunk > 1) {
                __syncthreads();
                stage(cb);
                __syncthreads();
            }
            const int kc = min(kch, kpad - cb);
            for (int c0 = 0; c0 < kc; c0 += 16) {
                E a[KCH];
                load_row<E, KCH>(Cs + (c0 + r) * Kp + g * KCH, a);
                acc_t h;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    h[i] = Hs[c0 + (Op::strided ? g + 4 * i : 4 * g + i)];
#pragma unroll
                for (int pt = 0; pt < PTT; pt++) {
                    acc_t acc = h;
#pragma unroll
                    for (int f = 0; f < KCH; f++)
                        acc = Op::mfma(a[f], b[pt][f], acc);
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const E v = acc[i];
                        const int c = cb + c0 + (Op::strided ? g + 4 * i : 4 * g + i);
                        if (v > best1[pt]) {
                            best2[pt] = best1[pt];
                            best1[pt] = v;
                            bidx[pt] = c;
                        } else if (v > best2[pt]) {
                            best2[pt] = v;
                        }
                    }
                }
            }
        }
        // Combine the four lane groups that hold the same sample (lanes r, r+16, r+32, r+48).
#pragma unroll
        for (int pt = 0; pt < PTT; pt++) {
            E xs = xn[pt];
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                const E o1 = __shfl_xor(best1[pt], off, 64);
                const E o2 = __shfl_xor(best2[pt], off, 64);
                const int oi = __shfl_xor(bidx[pt], off, 64);
                xs += __shfl_xor(xs, off, 64);
                const bool other = (o1 > best1[pt]) || (o1 == best1[pt] && oi < bidx[pt]);
                const E lo = other ? best1[pt] : o1; // the loser's best
                if (other) {
                    best1[pt] = o1;
                    bidx[pt] = oi;
                }
                best2[pt] = fmax(fmax(best2[pt], o2), lo);
            }
            const int i = s0 + pt * 16 + r;
            const bool valid = g == 0 && i < n;
            bool listed = false, moved = false;
            if (valid) {
                // |acc - exact| <= gamma_{d+2} (|x| cmax + cmax^2/2) + unit (|x| cmax + cmax^2/2)
                const double unit = Op::unit;
                const double xnorm = sqrt((double)xs) * (1.0 + 4.0 * unit);
                const double gam = (Kp + 2) * unit / (1.0 - (Kp + 2) * unit);
                const double e = (gam + 2.0 * unit) * (xnorm * cmax + 0.5 * cmax * cmax);
                const bool near_tie = (double)best1[pt] - (double)best2[pt] <= 2.0 * e;
                const int lab = bidx[pt] < k ? bidx[pt] : 0;
                labels[i] = lab;
                listed = near_tie || !(best1[pt] - best1[pt] == (E)0);
                moved = !listed && prev && prev[i] != lab;
            }
            // One atomic per wave for the re-check list and one for the change count.
            const unsigned long long lm = __ballot(listed);
            const unsigned long long cm = __ballot(moved);
            if (lm) {
                const int first = __ffsll((long long)lm) - 1;
                int base = 0;
                if (lane == first)
                    base = atomicAdd(list_count, __popcll(lm));
                base = __builtin_amdgcn_readlane(base, first);
                if (listed)
                    list[base + __popcll(lm & ((1ull << lane) - 1))] = i;
            }
            if (cm && lane == 0)
                atomicAdd(changed, __popcll(cm));
        }
    }
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// The assignment with a bf16 cross term (v_mfma_f32_16x16x32_bf16, fp32 accumulation). The
// data are a bf16 copy padded to Kp (a multiple of 32) features; lane group g holds features
// g*KCH .. g*KCH + KCH - 1 of its sample (the k-permutation, so that a lane's operands are
// contiguous), KCH / 8 MFMA steps per 16 x 16 tile. The near-tie bound adds the bf16
// rounding of both operands to the fp32 bound.
template <int KCH>
__global__ __launch_bounds__(256) void kmeans_assign_bf16_kernel(
    const __bf16 *__restrict__ X, int n, int Kp, const __bf16 *__restrict__ C16,
    const float *__restrict__ chalf, int k, double cmax, int *__restrict__ labels,
    const int *__restrict__ prev, int *__restrict__ changed, int *__restrict__ list_count,
    int *__restrict__ list) {
    constexpr int NF = KCH / 8;
    extern __shared__ float smem[];
    const int kpad = (k + 15) / 16 * 16;
    __bf16 *Cs = reinterpret_cast<__bf16 *>(smem); // kpad x Kp centres
    float *Hs = smem + (size_t)kpad * Kp / 2;       // -||c||^2/2, padded with -inf
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(C16);
        uint4 *dst = reinterpret_cast<uint4 *>(Cs);
        const int n16 = kpad * Kp / 8;
        for (int t = threadIdx.x; t < n16; t += 256)
            dst[t] = src[t];
    }
    for (int t = threadIdx.x; t < kpad; t += 256)
        Hs[t] = t < k ? chalf[t] : -__builtin_huge_valf();
    __syncthreads();

    // The centres are staged once per workgroup, which then takes every gridDim.x-th tile of
    // SB samples.
    const int ntiles = (n + SB - 1) / SB;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int s0 = tile * SB + w * PT * 16;
        v8bf b[PT][NF];
        float xn[PT];
#pragma unroll
        for (int pt = 0; pt < PT; pt++) {
            const int i = s0 + pt * 16 + r;
            float s = 0.f;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                v8bf v = {};
                if (i < n)
                    v = *reinterpret_cast<const v8bf *>(X + (int64_t)i * Kp + g * KCH + f * 8);
                b[pt][f] = v;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const float x = (float)v[e];
                    s = fmaf(x, x, s);
                }
            }
            xn[pt] = s;
        }
        float best1[PT], best2[PT];
        int bidx[PT];
#pragma unroll
        for (int pt = 0; pt < PT; pt++) {
            best1[pt] = -__builtin_huge_valf();
            best2[pt] = -__builtin_huge_valf();
            bidx[pt] = INT_MAX;
        }
        for (int c0 = 0; c0 < kpad; c0 += 16) {
            v8bf a[NF];
            const __bf16 *src = Cs + (c0 + r) * Kp + g * KCH;
#pragma unroll
            for (int f = 0; f < NF; f++)
                a[f] = *reinterpret_cast<const v8bf *>(src + f * 8);
            v4f h;
#pragma unroll
            for (int i = 0; i < 4; i++)
                h[i] = Hs[c0 + 4 * g + i];
#pragma unroll
            for (int pt = 0; pt < PT; pt++) {
                v4f acc = h;
#pragma unroll
                for (int f = 0; f < NF; f++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[f], b[pt][f], acc, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float v = acc[i];
                    const int c = c0 + 4 * g + i;
                    if (v > best1[pt]) {
                        best2[pt] = best1[pt];
                        best1[pt] = v;
                        bidx[pt] = c;
                    } else if (v > best2[pt]) {
                        best2[pt] = v;
                    }
                }
            }
        }
        const double u = 1.0 / 256.0, u32 = 1.0 / 16777216.0;
#pragma unroll
        for (int pt = 0; pt < PT; pt++) {
            float xs = xn[pt];
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                const float o1 = __shfl_xor(best1[pt], off, 64);
                const float o2 = __shfl_xor(best2[pt], off, 64);
                const int oi = __shfl_xor(bidx[pt], off, 64);
                xs += __shfl_xor(xs, off, 64);
                const bool other = (o1 > best1[pt]) || (o1 == best1[pt] && oi < bidx[pt]);
                const float lo = other ? best1[pt] : o1;
                if (other) {
                    best1[pt] = o1;
                    bidx[pt] = oi;
                }
                best2[pt] = fmaxf(fmaxf(best2[pt], o2), lo);
            }
            const int i = s0 + pt * 16 + r;
            const bool valid = g == 0 && i < n;
            bool listed = false, moved = false;
            if (valid) {
                // |x| of the data from the norm of its bf16 copy (relative error <= u), with a
                // margin for the fp32 sum. Score error: operand rounding (2u + u^2)|x| cmax, the
                // fp32 accumulation gamma_{Kp+2} over |x~| |c~| + cmax^2/2, the rounded -|c|^2/2.
                const double xnorm = sqrt((double)xs) / (1.0 - u) * (1.0 + 1e-6);
                const double gam = (Kp + 2) * u32 / (1.0 - (Kp + 2) * u32);
                const double S = xnorm * cmax;
                const double e = (2.0 * u + u * u) * S + gam * (S * (1.0 + u) * (1.0 + u) + 0.5 * cmax * cmax) +
                                 u32 * 0.5 * cmax * cmax;
                const bool near_tie = (double)best1[pt] - (double)best2[pt] <= 2.0 * e;
                const int lab = bidx[pt] < k ? bidx[pt] : 0;
                labels[i] = lab;
                listed = near_tie || !(best1[pt] - best1[pt] == 0.0f);
                moved = !listed && prev && prev[i] != lab;
            }
            // One atomic per wave for the re-check list and one for the change count.
            const unsigned long long lm = __ballot(listed);
            const unsigned long long cm = __ballot(moved);
            if (lm) {
                const int first = __ffsll((long long)lm) - 1;
                int base = 0;
                if (lane == first)
                    base = atomicAdd(list_count, __popcll(lm));
                base = __builtin_amdgcn_readlane(base, first);
                if (listed)
                    list[base + __popcll(lm & ((1ull << lane) - 1))] = i;
            }
            if (cm && lane == 0)
                atomicAdd(changed, __popcll(cm));
        }
    }
}

// Exact re-check of listed samples: argmin of double squared distances, ties to the lowest
// centre. One wave per listed sample (fixed grid, the count is read on the device): lane c
// takes centres c, c + 64, ..., then a wave argmin by (distance, index).
template <typename T, bool SPH>
__global__ __launch_bounds__(256) void kmeans_recheck_kernel(
    const T *__restrict__ X, int ldx, int d, const double *__restrict__ CT, int k,
    const int *__restrict__ list_count, const int *__restrict__ list, int *__restrict__ labels,
    const int *__restrict__ prev, int *__restrict__ changed) {
    const int cnt = *list_count;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int t = wave; t < cnt; t += nwaves) {
        const int i = list[t];
        const T *x = X + (int64_t)i * ldx;
        double best = __builtin_huge_val();
        int bl = INT_MAX;
        for (int c = lane; c < k; c += 64) {
            // Centres transposed (feature-major): the 64 lanes read 64 consecutive centres.
            double acc = 0.0;
            for (int l = 0; l < d; l++) {
                const double cv = CT[(int64_t)l * k + c];
                if constexpr (SPH) {
                    acc = fma(-(double)x[l], cv, acc); // key -x.c
                } else {
                    const double v = (double)x[l] - cv;
                    acc = fma(v, v, acc);
                }
            }
            if (acc < best) {
                best = acc;
                bl = c;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bl, off, 64);
            if (ob < best || (ob == best && oi < bl)) {
                best = ob;
                bl = oi;
            }
        }
        if (lane == 0) {
            labels[i] = bl;
            if (prev && prev[i] != bl)
                atomicAdd(changed, 1);
        }
    }
}

template <typename E>
__global__ void centre_prep_kernel(const double *__restrict__ C, int k, int d, int Kp,
                                   E *__restrict__ CP, E *__restrict__ chalf,
                                   double *__restrict__ cmax2, int sph) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k)
        return;
    double s = 0.0;
    for (int l = 0; l < Kp; l++) {
        const double v = l < d ? C[(int64_t)c * d + l] : 0.0;
        CP[(int64_t)c * Kp + l] = (E)v;
        s = fma(v, v, s);
    }
    chalf[c] = sph ? (E)0 : (E)(-0.5 * s);
    atomicMax(reinterpret_cast<unsigned long long *>(cmax2), (unsigned long long)__double_as_longlong(s));
}

// bf16 centres padded to Kp features (rows up to kpad are zero), -||c||^2/2 in fp32 and
// the largest ||c||^2 (atomic max on the bit pattern of a non-negative double).
__global__ void centre_prep_bf16_kernel(const double *__restrict__ C, int k, int kpad, int d,
                                        int Kp, __bf16 *__restrict__ C16,
                                        float *__restrict__ chalf, double *__restrict__ cmax2,
                                        int sph) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kpad)
        return;
    double s = 0.0;
    for (int l = 0; l < Kp; l++) {
        const double v = (c < k && l < d) ? C[(int64_t)c * d + l] : 0.0;
        C16[(int64_t)c * Kp + l] = (__bf16)(float)v;
        s = fma(v, v, s);
    }
    if (c < k) {
        chalf[c] = sph ? 0.0f : (float)(-0.5 * s);
        atomicMax(reinterpret_cast<unsigned long long *>(cmax2),
                  (unsigned long long)__double_as_longlong(s));
    }
}

__global__ void to_bf16_rows_kernel(const float *__restrict__ X, int n, int d, int ldx,
                                    __bf16 *__restrict__ Y, int Kp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * Kp)
        return;
    const int64_t i = t / Kp;
    const int l = (int)(t - i * Kp);
    Y[t] = l < d ? (__bf16)X[i * ldx + l] : (__bf16)0.0f;
}

// Workgroups of the assignments: one per tile of SB samples (wg_per_cu = 0), or at most
// wg_per_cu per CU, each taking every gridDim.x-th tile.
unsigned assign_grid(int n, int wg_per_cu, int sb = SB) {
    const int ntiles = (n + sb - 1) / sb;
    if (wg_per_cu <= 0)
        return (unsigned)std::max(1, ntiles);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 256;
    return (unsigned)std::max(1, std::min(ntiles, wg_per_cu * cus));
}

} // namespace

hipError_t launch_centre_prep_bf16(const double *C, int k, int d, int Kp, void *C16,
                                   float *chalf, double *cmax2, hipStream_t s, int sph) {
    const int kpad = (k + 15) / 16 * 16;
    hipLaunchKernelGGL(centre_prep_bf16_kernel, dim3((kpad + 255) / 256), dim3(256), 0, s, C, k,
                       kpad, d, Kp, static_cast<__bf16 *>(C16), chalf, cmax2, sph);
    return hipGetLastError();
}

hipError_t launch_rows_to_bf16(const float *X, int n, int d, int ldx, void *Y, int Kp,
                               hipStream_t s) {
    const int64_t total = (int64_t)n * Kp;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(to_bf16_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       s, X, n, d, ldx, static_cast<__bf16 *>(Y), Kp);
    return hipGetLastError();
}

size_t kmeans_assign_bf16_lds(int k, int Kp) {
    const int kpad = (k + 15) / 16 * 16;
    return (size_t)kpad * Kp * 2 + (size_t)kpad * sizeof(float);
}

hipError_t launch_kmeans_assign_bf16(const void *X, int n, int Kp, const void *C16,
                                     const float *chalf, int k, double cmax, int *labels,
                                     const int *prev, int *changed, int *list_count, int *list,
                                     hipStream_t s, int wg_per_cu) {
    const size_t lds = kmeans_assign_bf16_lds(k, Kp);
    if (lds > 160 * 1024 || Kp % 32 != 0)
        return hipErrorInvalidValue;
    dim3 grid(assign_grid(n, wg_per_cu)), block(256);
    const __bf16 *x = static_cast<const __bf16 *>(X);
    const __bf16 *c = static_cast<const __bf16 *>(C16);
#define DA_ASSIGN(KCHV)                                                                          \
    {                                                                                            \
        auto kern = kmeans_assign_bf16_kernel<KCHV>;                                             \
        if (lds > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)kern,                                        \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);     \
        hipLaunchKernelGGL(kern, grid, block, lds, s, x, n, Kp, c, chalf, k, cmax, labels, prev, \
                           changed, list_count, list);                                           \
    }
    switch (Kp / 4) {
    case 8: DA_ASSIGN(8) break;
    case 16: DA_ASSIGN(16) break;
    case 24: DA_ASSIGN(24) break;
    case 32: DA_ASSIGN(32) break;
    default:
        return hipErrorInvalidValue;
    }
#undef DA_ASSIGN
    return hipGetLastError();
}

template <typename E>
hipError_t launch_centre_prep(const double *C, int k, int d, int Kp, E *CP, E *chalf,
                              double *cmax2, hipStream_t s, int sph) {
    hipLaunchKernelGGL(centre_prep_kernel<E>, dim3((k + 255) / 256), dim3(256), 0, s, C, k, d,
                       Kp, CP, chalf, cmax2, sph);
    return hipGetLastError();
}
template hipError_t launch_centre_prep<float>(const double *, int, int, int, float *, float *,
                                              double *, hipStream_t, int);
template hipError_t launch_centre_prep<double>(const double *, int, int, int, double *,
                                               double *, double *, hipStream_t, int);

template <typename E>
hipError_t launch_kmeans_assign_mfma(const E *X, int n, int Kp, const E *CP, const E *chalf,
                                     int k, double cmax, int *labels, const int *prev,
                                     int *changed, int *list_count, int *list, hipStream_t s,
                                     int wg_per_cu) {
    if (Kp % 16 != 0 || Kp > 128 || k < 1)
        return hipErrorInvalidValue;
    const int kch = assign_chunk<E>(k, Kp);
    const size_t lds = (size_t)kch * (Kp + 1) * sizeof(E);
    // fp64 samples of more than 64 features: two tiles per wave (registers).
    const bool wide = sizeof(E) == 8 && Kp > 64;
    dim3 grid(assign_grid(n, wg_per_cu, WAVES * (wide ? 2 : PT) * 16)), block(256);
#define DA_ASSIGN(KCHV, PTV)                                                                     \
    {                                                                                            \
        auto kern = kmeans_assign_mfma_kernel<E, KCHV, PTV>;                                     \
        if (lds > 65536)                                                                         \
            (void)hipFuncSetAttribute((const void *)kern,                                        \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);     \
        hipLaunchKernelGGL(kern, grid, block, lds, s, X, n, Kp, CP, chalf, k, kch, cmax, labels, \
                           prev, changed, list_count, list);                                     \
    }
    switch (Kp / 4) {
    case 4: DA_ASSIGN(4, PT) break;
    case 8: DA_ASSIGN(8, PT) break;
    case 12: DA_ASSIGN(12, PT) break;
    case 16: DA_ASSIGN(16, PT) break;
    case 24:
        if (wide)
            DA_ASSIGN(24, 2)
        else
            DA_ASSIGN(24, PT)
        break;
    case 32:
        if (wide)
            DA_ASSIGN(32, 2)
        else
            DA_ASSIGN(32, PT)
        break;
    default:
        return hipErrorInvalidValue;
    }
#undef DA_ASSIGN
    return hipGetLastError();
}
template hipError_t launch_kmeans_assign_mfma<float>(const float *, int, int, const float *,
                                                     const float *, int, double, int *,
                                                     const int *, int *, int *, int *,
                                                     hipStream_t, int);
template hipError_t launch_kmeans_assign_mfma<double>(const double *, int, int, const double *,
                                                      const double *, int, double, int *,
                                                      const int *, int *, int *, int *,
                                                      hipStream_t, int);

__global__ __launch_bounds__(256) void transpose_centres_kernel(const double *__restrict__ C,
                                                                int k, int d,
                                                                double *__restrict__ CT) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)k * d)
        return;
    const int c = (int)(t / d), l = (int)(t - (int64_t)c * d);
    CT[(int64_t)l * k + c] = C[t];
}

template <typename T>
hipError_t launch_kmeans_recheck(const T *X, int ldx, int d, const double *C, int k,
                                 const int *list_count, const int *list, int *labels,
                                 const int *prev, int *changed, double *CT, hipStream_t s,
                                 int sph) {
    const int64_t kd = (int64_t)k * d;
    hipLaunchKernelGGL(transpose_centres_kernel, dim3((unsigned)((kd + 255) / 256)), dim3(256), 0,
                       s, C, k, d, CT);
    if (sph)
        hipLaunchKernelGGL((kmeans_recheck_kernel<T, true>), dim3(1024), dim3(256), 0, s, X, ldx,
                           d, CT, k, list_count, list, labels, prev, changed);
    else
        hipLaunchKernelGGL((kmeans_recheck_kernel<T, false>), dim3(1024), dim3(256), 0, s, X, ldx,
                           d, CT, k, list_count, list, labels, prev, changed);
    return hipGetLastError();
}
template hipError_t launch_kmeans_recheck<float>(const float *, int, int, const double *, int,
                                                 const int *, const int *, int *, const int *,
                                                 int *, double *, hipStream_t, int);
template hipError_t launch_kmeans_recheck<double>(const double *, int, int, const double *, int,
                                                  const int *, const int *, int *, const int *,
                                                  int *, double *, hipStream_t, int);

} // namespace dagpu
