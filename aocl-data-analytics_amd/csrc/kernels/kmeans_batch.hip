// Batched Lloyd k-means for gfx950: many independent instances of uneven size in one set of
// launches. There is no batched entry point in the reference; each instance follows the
// single-instance algorithm (kmeans.cpp:996-1124, kmeans_lloyd.hpp:105-405) with supplied
// initial centres, and the results match running the instances one by one.
//
// Work lists instead of per-instance launches:
//  * assignment tiles (instance, first sample) of 64*PT samples, dealt round-robin to a
//    persistent grid sized to the device, most expensive instances first;
//  * sample chunks (instance, first sample, count) for the centroid sums, the column maxima
//    and the inertia.
// The assignment runs the cross term on MFMA (fp32 v_mfma_f32_16x16x4f32 or bf16
// v_mfma_f32_16x16x32_bf16) against centres staged in LDS per instance, and re-checks every
// sample whose two best scores are within the error bound with exact double distances, so
// the labels are the exact argmin (ties to the lowest centre) in either form. Centroid sums are 64-bit fixed point per instance and feature (the same
// scheme as the single-instance kernels), reduced in a fixed order: the result is bitwise
// reproducible. Convergence is tested per instance on the device; converged instances drop
// out of the later launches, so the whole run can be captured in one graph.
#include "launch.hpp"
#include <climits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Operand forms of the assignment cross term. E: elements per lane per MFMA; NS: parts per
// value (2: hi + lo in bf16, rows stored [hi kp][lo kp], products hi*hi + hi*lo + lo*hi);
// UX, UC: relative error of a product term charged to the data and to the centres (fp32
// data are exact in fp32; centres are rounded from double).
constexpr double KB_U32 = 1.0 / 16777216.0, KB_UB16 = 1.0 / 256.0;
struct KbBF16 {
    using T = __bf16;
    using frag = v8bf;
    static constexpr int E = 8, NS = 1;
    static constexpr double UX = KB_UB16, UC = KB_UB16;
    __device__ static v4f mfma(frag a, frag b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
struct KbF32 {
    using T = float;
    using frag = float;
    static constexpr int E = 1, NS = 1;
    static constexpr double UX = 0.0, UC = KB_U32;
    __device__ static v4f mfma(frag a, frag b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
// hi + lo split: |x - (hi + lo)| <= ub16^2 |x| and the dropped lo*lo <= ub16^2 (1 + ub16)^2
// |x||c|, so a product carries at most 3 ub16^2 (1 + ub16)^2 |x||c|, charged half to each
// operand; the centres add their rounding to fp32.
struct KbBF16X3 {
    using T = __bf16;
    using frag = v8bf;
    static constexpr int E = 8, NS = 2;
    static constexpr double UX = 1.5 * KB_UB16 * KB_UB16 * (1 + KB_UB16) * (1 + KB_UB16),
                            UC = UX + KB_U32;
    __device__ static v4f mfma(frag a, frag b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// NF fragments of a lane from contiguous operand elements (16- or 8-byte vector loads).
template <class Op, int NF>
__device__ __forceinline__ void kb_load(const typename Op::T *p, typename Op::frag *f) {
    if constexpr (Op::E == 8) {
#pragma unroll
        for (int v = 0; v < NF; v++)
            f[v] = *reinterpret_cast<const v8bf *>(p + 8 * v);
    } else if constexpr (NF % 4 == 0) {
#pragma unroll
        for (int v = 0; v < NF; v += 4) {
            const v4f x = *reinterpret_cast<const v4f *>(p + v);
            f[v] = x[0];
            f[v + 1] = x[1];
            f[v + 2] = x[2];
            f[v + 3] = x[3];
        }
    } else {
        static_assert(NF == 2, "fp32 fragments come in pairs or fours");
        const v2f x = *reinterpret_cast<const v2f *>(p);
        f[0] = x[0];
        f[1] = x[1];
    }
}


// Column maxima |x| per instance and feature (atomic max on the bit pattern of a
// non-negative double).
__global__ __launch_bounds__(256) void kb_absmax_kernel(const float *__restrict__ X, KBatchDesc b,
                                                        const KChunk *__restrict__ chunks,
                                                        unsigned long long *__restrict__ amax) {
    const KChunk ch = chunks[blockIdx.x];
    const int inst = ch.inst, d = b.d[inst];
    const float *x = X + b.xoff[inst] + (int64_t)ch.start * d;
    __shared__ double m[64];
    if (threadIdx.x < 64)
        m[threadIdx.x] = 0.0;
    __syncthreads();
    // Thread t takes feature t % d of samples t / d, t / d + spp, ... (d <= 64).
    const int spp = 256 / d, sub = (int)threadIdx.x / d, l = (int)threadIdx.x - sub * d;
    if (sub < spp) {
        double mx = 0.0;
        for (int i = sub; i < ch.count; i += spp) {
            const double v = fabs((double)x[(int64_t)i * d + l]);
            if (v == v && v < __builtin_huge_val())
                mx = fmax(mx, v);
        }
        atomicMax(reinterpret_cast<unsigned long long *>(&m[l]),
                  (unsigned long long)__double_as_longlong(mx));
    }
    __syncthreads();
    if ((int)threadIdx.x < d)
        atomicMax(&amax[(int64_t)inst * 64 + threadIdx.x],
                  (unsigned long long)__double_as_longlong(m[threadIdx.x]));
}

// Fixed-point exponents per instance: feature l gets e_l = 60 - ceil(log2(n+1)) - E(amax_l)
// (so that no sum of n values exceeds 2^60), and the low word is scaled by a further
// 2^(61 - ceil(log2(n+1))). Stored at escale[inst*65 + l], the low-word shift at +64.
__global__ void kb_escale_kernel(KBatchDesc b, int n_inst,
                                 const unsigned long long *__restrict__ amax,
                                 int *__restrict__ escale) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= n_inst)
        return;
    const int n = b.n[inst], d = b.d[inst];
    int l2n = 0;
    while (((int64_t)1 << l2n) < (int64_t)n + 1)
        l2n++;
    for (int l = 0; l < 64; l++) {
        int e = 0;
        if (l < d) {
            const double a = __longlong_as_double((long long)amax[(int64_t)inst * 64 + l]);
            if (a > 0.0) {
                int E = 0;
                (void)frexp(a, &E);
                e = 60 - l2n - E;
            }
        }
        escale[inst * 65 + l] = e;
    }
    escale[inst * 65 + 64] = 61 - l2n;
}

// Copy of each instance's rows in the operand type, padded to kp features.
template <class Op>
__global__ __launch_bounds__(256) void kb_to_op_kernel(const float *__restrict__ X, KBatchDesc b,
                                                       const KChunk *__restrict__ chunks,
                                                       typename Op::T *__restrict__ Xo) {
    using T = typename Op::T;
    const KChunk ch = chunks[blockIdx.x];
    const int inst = ch.inst, d = b.d[inst], kp = b.kp[inst];
    const float *x = X + b.xoff[inst] + (int64_t)ch.start * d;
    const int rs = kp * Op::NS;
    T *y = Xo + b.xboff[inst] + (int64_t)ch.start * rs;
    const int spp = 256 / kp, sub = (int)threadIdx.x / kp, l = (int)threadIdx.x - sub * kp;
    for (int i = sub; i < ch.count; i += spp) {
        const float v = l < d ? x[(int64_t)i * d + l] : 0.0f;
        const T hi = (T)v;
        y[(int64_t)i * rs + l] = hi;
        if constexpr (Op::NS == 2)
            y[(int64_t)i * rs + kp + l] = (T)(v - (float)hi);
    }
}

// Initial centres (fp32, packed k x d per instance) into the double master copy.
__global__ void kb_centres_in_kernel(const float *__restrict__ Cin, int64_t total,
                                     double *__restrict__ C) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total)
        C[t] = (double)Cin[t];
}

// Per-pass reset: the re-check list and the label-change counters.
__global__ void kb_reset_kernel(int *__restrict__ list_count, int *__restrict__ changed,
                                int n_inst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0)
        *list_count = 0;
    if (t < n_inst)
        changed[t] = 0;
}

// Whether instance inst takes part in a pass: the main loop (mode 0) runs the active ones;
// the final relabel (mode 1) runs those that stopped on the centre shift or at max_iter.
__device__ __forceinline__ bool kb_runs(const int *state, int inst, int mode) {
    const int s = state[inst];
    return mode == 0 ? s == KB_ACTIVE : (s == KB_CONV_SHIFT || s == KB_MAXIT);
}

// Which centre / label buffers an instance uses: in iteration j (mode 0) the centres of
// parity p = j & 1 are the input and labels are written to buffer p; the final relabel uses
// the centres of the last update (1 - fpar) and writes labels to buffer 1 - fpar.
__device__ __forceinline__ int kb_in_parity(const int *fpar, int inst, int mode, int par) {
    return mode == 0 ? par : 1 - fpar[inst];
}

// Per instance and selected parity: centres in the operand type (kpad x kp, zero padded),
// -|c|^2/2 in fp32 and max |c|^2. One workgroup per instance.
template <class Op>
__global__ __launch_bounds__(64) void kb_centre_prep_kernel(KBatchDesc b, const double *C0,
                                                            const double *C1,
                                                            const int *__restrict__ state,
                                                            const int *__restrict__ fpar, int mode,
                                                            int par, typename Op::T *__restrict__ Cb,
                                                            float *__restrict__ chalf,
                                                            double *__restrict__ cmax2) {
    using T = typename Op::T;
    const int inst = blockIdx.x;
    if (!kb_runs(state, inst, mode))
        return;
    const int k = b.k[inst], d = b.d[inst], kp = b.kp[inst];
    const double *C = (kb_in_parity(fpar, inst, mode, par) ? C1 : C0) + b.coff[inst];
    const int kpad = (k + 15) / 16 * 16;
    __shared__ double mx[64];
    const int c = threadIdx.x;
    double s = 0.0;
    if (c < kpad) {
        T *row = Cb + b.cboff[inst] + (int64_t)c * kp * Op::NS;
        for (int l = 0; l < kp; l++) {
            const double v = (c < k && l < d) ? C[(int64_t)c * d + l] : 0.0;
            const float f = (float)v;
            const T hi = (T)f;
            row[l] = hi;
            if constexpr (Op::NS == 2)
                row[kp + l] = (T)(f - (float)hi);
            s = fma(v, v, s);
        }
        if (c < k)
            chalf[inst * 64 + c] = (float)(-0.5 * s);
    }
    mx[c] = c < k ? s : 0.0;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (c < w)
            mx[c] = fmax(mx[c], mx[c + w]);
        __syncthreads();
    }
    if (c == 0)
        cmax2[inst] = mx[0];
}

// Assignment of one tile: 16*PT samples per wave against the instance's centres in LDS.
// Lane group g holds features g*kp/4 .. of its sample (the k-permutation), NF fragments per
// part (NS parts: the value, or its hi and lo).
template <class Op, int NF, int PT>
__device__ __forceinline__ void kb_assign_tile(const typename Op::T *__restrict__ Xo, int n,
                                               int kp, int start, const typename Op::T *Cs,
                                               const float *Hs, int k, double cmax,
                                               int *__restrict__ labels,
                                               const int *__restrict__ prev, int *changed,
                                               int *list_count, KTile *list, int inst) {
    using frag = typename Op::frag;
    constexpr int KCH = NF * Op::E;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int kpad = (k + 15) / 16 * 16;
    const int s0 = start + w * PT * 16;
    constexpr int NS = Op::NS;
    frag bb[PT][NS][NF];
    float xn[PT];
#pragma unroll
    for (int pt = 0; pt < PT; pt++) {
        const int i = s0 + pt * 16 + r;
#pragma unroll
        for (int h = 0; h < NS; h++) {
            kb_load<Op, NF>(Xo + (int64_t)(i < n ? i : 0) * kp * NS + h * kp + g * KCH,
                            bb[pt][h]);
            if (i >= n) {
#pragma unroll
                for (int f = 0; f < NF; f++)
                    bb[pt][h][f] = frag{};
            }
        }
        float s = 0.f;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if constexpr (Op::E == 8) {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    float x = (float)bb[pt][0][f][e];
                    if constexpr (NS == 2)
                        x += (float)bb[pt][NS - 1][f][e];
                    s = fmaf(x, x, s);
                }
            } else {
                s = fmaf(bb[pt][0][f], bb[pt][0][f], s);
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
        frag a[NS][NF];
#pragma unroll
        for (int h = 0; h < NS; h++)
            kb_load<Op, NF>(Cs + (c0 + r) * kp * NS + h * kp + g * KCH, a[h]);
        v4f h;
#pragma unroll
        for (int i = 0; i < 4; i++)
            h[i] = Hs[c0 + 4 * g + i];
#pragma unroll
        for (int pt = 0; pt < PT; pt++) {
            v4f acc = h;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                acc = Op::mfma(a[0][f], bb[pt][0][f], acc);
                if constexpr (NS == 2) {
                    acc = Op::mfma(a[0][f], bb[pt][NS - 1][f], acc);
                    acc = Op::mfma(a[NS - 1][f], bb[pt][0][f], acc);
                }
            }
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
    const double ux = Op::UX, uc = Op::UC, u32 = 1.0 / 16777216.0;
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
            // Score error: operand rounding (ux + uc + ux uc)|x| cmax, fp32 accumulation
            // gamma_{kp+2} over the rounded products and -|c|^2/2, the rounded -|c|^2/2;
            // |x| from the operand copy (relative error <= ux) with a margin for the sum.
            const double xnorm = sqrt((double)xs) / (1.0 - ux) * (1.0 + 1e-6);
            // fp32 accumulation of NS == 2 ? 3 kp : kp products and -|c|^2/2.
            const int terms = (NS == 2 ? 3 : 1) * kp + 2;
            const double gam = terms * u32 / (1.0 - terms * u32);
            const double S = xnorm * cmax;
            const double e = (ux + uc + ux * uc) * S +
                             gam * (S * (1.0 + ux) * (1.0 + uc) + 0.5 * cmax * cmax) +
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
                list[base + __popcll(lm & ((1ull << lane) - 1))] = KTile{inst, i};
        }
        if (cm && lane == 0)
            atomicAdd(changed, __popcll(cm));
    }
}

// Persistent assignment: the tile list (sorted by decreasing cost, the tiles of an instance
// adjacent) is dealt in runs of KB_RUN tiles, workgroup b taking runs b, b + G, b + 2G, ...:
// the round-robin deal keeps the grid balanced as instances converge and drop out, and the
// centres of an instance are staged once per run.
constexpr int KB_RUN = 4;
template <class Op, int PT>
__global__ __launch_bounds__(256) void kb_assign_kernel(
    const typename Op::T *__restrict__ Xo, KBatchDesc b, const KTile *__restrict__ tiles,
    int ntiles, const int *__restrict__ state, const int *__restrict__ fpar, int mode, int par,
    const typename Op::T *__restrict__ Cb, const float *__restrict__ chalf,
    const double *__restrict__ cmax2, int *__restrict__ L0, int *__restrict__ L1,
    int *__restrict__ changed, int *__restrict__ list_count, KTile *__restrict__ list) {
    using T = typename Op::T;
    __shared__ __attribute__((aligned(16))) T Cs[64 * 64 * Op::NS];
    __shared__ float Hs[64];
    int staged = -1;
    for (int t = blockIdx.x * KB_RUN; t < ntiles;
         t += (t % KB_RUN == KB_RUN - 1) ? (gridDim.x - 1) * KB_RUN + 1 : 1) {
        const KTile tl = tiles[t];
        const int inst = tl.inst;
        if (!kb_runs(state, inst, mode))
            continue;
        const int k = b.k[inst], kp = b.kp[inst];
        if (inst != staged) {
            // Every wave is done with the previous instance's centres.
            __syncthreads();
            const int kpad = (k + 15) / 16 * 16;
            const uint4 *src = reinterpret_cast<const uint4 *>(Cb + b.cboff[inst]);
            uint4 *dst = reinterpret_cast<uint4 *>(Cs);
            for (int q = threadIdx.x; q < kpad * kp * Op::NS * (int)sizeof(T) / 16; q += 256)
                dst[q] = src[q];
            if (threadIdx.x < 64)
                Hs[threadIdx.x] = (int)threadIdx.x < k ? chalf[inst * 64 + threadIdx.x]
                                                       : -__builtin_huge_valf();
            __syncthreads();
            staged = inst;
        }
        const int p = kb_in_parity(fpar, inst, mode, par);
        int *lab_new = (p ? L1 : L0) + b.loff[inst];
        const int *lab_old = mode == 0 ? (p ? L0 : L1) + b.loff[inst] : nullptr;
        const double cmax = sqrt(cmax2[inst]);
        const T *xo = Xo + b.xboff[inst];
        const int n = b.n[inst];
#define KB_TILE(NFV)                                                                             \
    kb_assign_tile<Op, NFV, PT>(xo, n, kp, tl.start, Cs, Hs, k, cmax, lab_new, lab_old,        \
                                &changed[inst], list_count, list, inst)
        if constexpr (Op::E == 8) {
            if (kp == 32)
                KB_TILE(1);
            else
                KB_TILE(2);
        } else {
            if (kp == 8)
                KB_TILE(2);
            else if (kp == 16)
                KB_TILE(4);
            else if (kp == 32)
                KB_TILE(8);
            else
                KB_TILE(16);
        }
#undef KB_TILE
    }
}

// Exact re-check of listed samples (fixed grid, count read on the device): one wave per
// listed sample, lane c takes centre c (k <= 64), then a wave argmin by (distance, index).
__global__ __launch_bounds__(256) void kb_recheck_kernel(
    const float *__restrict__ X, KBatchDesc b, const double *C0, const double *C1,
    const int *__restrict__ fpar, int mode, int par, const int *__restrict__ list_count,
    const KTile *__restrict__ list, int *__restrict__ L0, int *__restrict__ L1,
    int *__restrict__ changed) {
    const int cnt = *list_count;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int t = wave; t < cnt; t += nwaves) {
        const KTile e = list[t];
        const int inst = e.inst, i = e.start;
        const int d = b.d[inst], k = b.k[inst];
        const int p = kb_in_parity(fpar, inst, mode, par);
        const double *C = (p ? C1 : C0) + b.coff[inst];
        const float *x = X + b.xoff[inst] + (int64_t)i * d;
        double best = __builtin_huge_val();
        int bl = INT_MAX;
        if (lane < k) {
            double acc = 0.0;
            for (int l = 0; l < d; l++) {
                const double v = (double)x[l] - C[(int64_t)lane * d + l];
                acc = fma(v, v, acc);
            }
            best = acc;
            bl = lane;
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
            int *lab_new = (p ? L1 : L0) + b.loff[inst];
            lab_new[i] = bl;
            if (mode == 0) {
                const int *lab_old = (p ? L0 : L1) + b.loff[inst];
                if (lab_old[i] != bl)
                    atomicAdd(&changed[inst], 1);
            }
        }
    }
}

// Fixed-point partial sums of one chunk: per (cluster, feature) a high and a low word, and
// the cluster counts, written to the chunk's slot.
__global__ __launch_bounds__(1024) void kb_partial_kernel(
    const float *__restrict__ X, KBatchDesc b, const KChunk *__restrict__ chunks,
    const int64_t *__restrict__ poff, const int *__restrict__ state, int par,
    const int *__restrict__ L0, const int *__restrict__ L1, const int *__restrict__ escale,
    long long *__restrict__ part, int *__restrict__ pcnt) {
    // The chunk's samples are counting-sorted by label in LDS (two 32-bit LDS atomics per
    // sample); then thread pairs (cluster, feature) sum their cluster's values in registers
    // and write the chunk's partial, which the reduce adds up in chunk order.
    extern __shared__ int sh[];
    int *off = sh;        // k + 1 (k <= 64)
    int *fill = sh + 65;  // k
    int *list = sh + 130; // chunk samples sorted by label (8-byte aligned)
    const KChunk ch = chunks[blockIdx.x];
    const int inst = ch.inst;
    if (state[inst] != KB_ACTIVE)
        return;
    const int d = b.d[inst], k = b.k[inst];
    const int *lab = (par ? L1 : L0) + b.loff[inst] + ch.start;
    const float *x = X + b.xoff[inst] + (int64_t)ch.start * d;
    const int *es = escale + inst * 65;
    const int lshift = es[64];
    if (threadIdx.x < 65)
        off[threadIdx.x] = 0;
    if (threadIdx.x < 64)
        fill[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < ch.count; i += 1024)
        atomicAdd(&off[lab[i]], 1);
    __syncthreads();
    if (threadIdx.x < 64) {
        // Exclusive scan of the k <= 64 counts in one wave; off[k] = chunk count.
        const int lane = threadIdx.x;
        const int v = lane < k ? off[lane] : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o)
                incl += y;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < k)
            off[lane] = incl - v;
        if (lane == k - 1)
            off[k] = incl;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ch.count; i += 1024) {
        const int c = lab[i];
        list[off[c] + atomicAdd(&fill[c], 1)] = i;
    }
    __syncthreads();
    long long *pb = part + poff[blockIdx.x];
    int *cb = pcnt + (int64_t)blockIdx.x * 64;
    // R threads per pair when the pairs are fewer than the threads: thread (p, r) takes every
    // R-th sample of the cluster from the r-th on.
    const int P = k * d;
    const int R = P >= 1024 ? 1 : 1024 / P;
    auto pair_sum = [&](int p, int r, long long &H, long long &Lw) {
        const int c = p / d, l = p - c * d;
        const int j1 = off[c + 1];
        const int el = es[l];
        H = 0;
        Lw = 0;
        auto add = [&](float xv) {
            const double v = (double)xv;
            if (v == v && v - v == 0.0) {
                const double sv = ldexp(v, el);
                const double h = rint(sv);
                H += (long long)h;
                Lw += (long long)rint(ldexp(sv - h, lshift));
            }
        };
        const float *xs = x + l;
        int j = off[c] + r;
        // Eight independent gathers in flight per thread.
        for (; j + 7 * R < j1; j += 8 * R) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                v[u] = xs[(int64_t)list[j + u * R] * d];
#pragma unroll
            for (int u = 0; u < 8; u++)
                add(v[u]);
        }
        for (; j < j1; j += R)
            add(xs[(int64_t)list[j] * d]);
    };
    if (R == 1) {
        for (int p = threadIdx.x; p < P; p += 1024) {
            long long H, Lw;
            pair_sum(p, 0, H, Lw);
            pb[2 * p] = H;
            pb[2 * p + 1] = Lw;
        }
    } else {
        const int t = threadIdx.x;
        long long H = 0, Lw = 0;
        if (t < P * R)
            pair_sum(t % P, t / P, H, Lw);
        // The sample list is dead: its space holds the per-thread sums (16 KB).
        __syncthreads();
        long long *red = reinterpret_cast<long long *>(list);
        red[t] = H;
        red[1024 + t] = Lw;
        __syncthreads();
        if (t < P) {
            long long sh_ = 0, sl = 0;
            for (int r = 0; r < R; r++) {
                sh_ += red[r * P + t];
                sl += red[1024 + r * P + t];
            }
            pb[2 * t] = sh_;
            pb[2 * t + 1] = sl;
        }
    }
    if ((int)threadIdx.x < k)
        cb[threadIdx.x] = off[threadIdx.x + 1] - off[threadIdx.x];
}

// Per instance: new centres from its chunks' partials (fixed order), the centre shift and
// norm, and the convergence test of kmeans.cpp:1359-1397 (iteration j > 1 with no label
// change: stop with labels unchanged; shift below tol * norm: stop on the shift).
__global__ __launch_bounds__(256) void kb_reduce_kernel(
    KBatchDesc b, const int *__restrict__ cbeg, const int *__restrict__ cend,
    const int64_t *__restrict__ poff, const long long *__restrict__ part,
    const int *__restrict__ pcnt, const int *__restrict__ escale, int par, int iter,
    double tol, double *C0, double *C1, const int *__restrict__ changed,
    int *__restrict__ state, int *__restrict__ fpar, int *__restrict__ n_iter) {
    const int inst = blockIdx.x;
    if (state[inst] != KB_ACTIVE)
        return;
    const int d = b.d[inst], k = b.k[inst];
    const double *Cin = (par ? C1 : C0) + b.coff[inst];
    double *Cout = (par ? C0 : C1) + b.coff[inst];
    const int *es = escale + inst * 65;
    const int lshift = es[64];
    __shared__ double s0[256], s1[256];
    double x0 = 0.0, x1 = 0.0;
    for (int e = threadIdx.x; e < k * d; e += 256) {
        const int c = e / d, l = e - c * d;
        long long H = 0, Lw = 0;
        int count = 0;
        for (int q = cbeg[inst]; q < cend[inst]; q++) {
            H += part[poff[q] + 2 * e];
            Lw += part[poff[q] + 2 * e + 1];
            count += pcnt[(int64_t)q * 64 + c];
        }
        const double sum = ldexp((double)H, -es[l]) + ldexp((double)Lw, -(es[l] + lshift));
        // An empty cluster keeps count 1 and a zero sum: the origin
        // (scale_current_cluster_centres, kmeans_lloyd.hpp:363-373).
        const double v = sum / (double)(count > 0 ? count : 1);
        Cout[e] = v;
        const double a = Cin[e];
        x0 = fma(a, a, x0);
        x1 = fma(a - v, a - v, x1);
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        fpar[inst] = par;
        n_iter[inst] = iter + 1;
        int st = KB_ACTIVE;
        if (iter > 1 && changed[inst] == 0)
            st = KB_CONV_LABELS;
        else if (sqrt(s1[0]) < tol * sqrt(s0[0]))
            st = KB_CONV_SHIFT;
        if (st != KB_ACTIVE) {
            n_iter[inst] = iter;
            state[inst] = st;
        }
    }
}

// After the last iteration: instances still active stopped at max_iter.
__global__ void kb_maxit_kernel(int *__restrict__ state, int n_inst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_inst && state[t] == KB_ACTIVE)
        state[t] = KB_MAXIT;
}

// Final labels of an instance: the last assignment for KB_CONV_LABELS, else the relabel.
__device__ __forceinline__ const int *kb_final_labels(const int *L0, const int *L1,
                                                      const int *state, const int *fpar,
                                                      int inst) {
    const int p = state[inst] == KB_CONV_LABELS ? fpar[inst] : 1 - fpar[inst];
    return p ? L1 : L0;
}

// Inertia per chunk: sum of exact squared distances to the final centres in a fixed order.
__global__ __launch_bounds__(256) void kb_inertia_kernel(
    const float *__restrict__ X, KBatchDesc b, const KChunk *__restrict__ chunks,
    const double *C0, const double *C1, const int *__restrict__ state,
    const int *__restrict__ fpar, const int *__restrict__ L0, const int *__restrict__ L1,
    double *__restrict__ cin) {
    const KChunk ch = chunks[blockIdx.x];
    const int inst = ch.inst, d = b.d[inst];
    const double *C = (1 - fpar[inst] ? C1 : C0) + b.coff[inst];
    const int *lab = kb_final_labels(L0, L1, state, fpar, inst) + b.loff[inst];
    const float *x = X + b.xoff[inst];
    double s = 0.0;
    for (int i = ch.start + threadIdx.x; i < ch.start + ch.count; i += 256) {
        const int c = lab[i];
        double acc = 0.0;
        for (int l = 0; l < d; l++) {
            const double v = (double)x[(int64_t)i * d + l] - C[(int64_t)c * d + l];
            acc = fma(v, v, acc);
        }
        s += acc;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        cin[blockIdx.x] = sh[0];
}

// Outputs: final centres (fp32) and labels in the caller's packed arrays, inertia per
// instance (its chunks summed in order).
__global__ __launch_bounds__(256) void kb_output_kernel(
    KBatchDesc b, const double *C0, const double *C1, const int *__restrict__ state,
    const int *__restrict__ fpar, const int *__restrict__ L0, const int *__restrict__ L1,
    const int *__restrict__ cbeg, const int *__restrict__ cend, const double *__restrict__ cin,
    float *__restrict__ Cout, int *__restrict__ labels, double *__restrict__ inertia) {
    const int inst = blockIdx.x;
    const int d = b.d[inst], k = b.k[inst], n = b.n[inst];
    const double *C = (1 - fpar[inst] ? C1 : C0) + b.coff[inst];
    for (int e = threadIdx.x; e < k * d; e += 256)
        Cout[b.coff[inst] + e] = (float)C[e];
    const int *lab = kb_final_labels(L0, L1, state, fpar, inst) + b.loff[inst];
    for (int i = threadIdx.x; i < n; i += 256)
        labels[b.loff[inst] + i] = lab[i];
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = cbeg[inst]; q < cend[inst]; q++)
            s += cin[q];
        inertia[inst] = s;
    }
}

__global__ void kb_fill_int_kernel(int *__restrict__ p, int n, int v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        p[t] = v;
}

__global__ void kb_fill_u64_kernel(unsigned long long *__restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        p[t] = 0ull;
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

} // namespace

hipError_t kb_launch_prepare(const float *X, const float *Cinit, const KBatchDesc &b, int n_inst,
                             const KChunk *chunks, int nchunks, int64_t ctotal,
                             unsigned long long *amax, int *escale, int prec, void *Xo,
                             double *C0, int *state, int *fpar, int *n_iter, hipStream_t s) {
    hipLaunchKernelGGL(kb_fill_u64_kernel, dim3(blocks_for((int64_t)n_inst * 64, 256)), dim3(256),
                       0, s, amax, (int64_t)n_inst * 64);
    hipLaunchKernelGGL(kb_absmax_kernel, dim3(nchunks), dim3(256), 0, s, X, b, chunks, amax);
    hipLaunchKernelGGL(kb_escale_kernel, dim3(blocks_for(n_inst, 64)), dim3(64), 0, s, b, n_inst,
                       amax, escale);
    if (prec == 1)
        hipLaunchKernelGGL(kb_to_op_kernel<KbBF16>, dim3(nchunks), dim3(256), 0, s, X, b, chunks,
                           reinterpret_cast<__bf16 *>(Xo));
    else if (prec == 2)
        hipLaunchKernelGGL(kb_to_op_kernel<KbBF16X3>, dim3(nchunks), dim3(256), 0, s, X, b,
                           chunks, reinterpret_cast<__bf16 *>(Xo));
    else
        hipLaunchKernelGGL(kb_to_op_kernel<KbF32>, dim3(nchunks), dim3(256), 0, s, X, b, chunks,
                           reinterpret_cast<float *>(Xo));
    hipLaunchKernelGGL(kb_centres_in_kernel, dim3(blocks_for(ctotal, 256)), dim3(256), 0, s, Cinit,
                       ctotal, C0);
    hipLaunchKernelGGL(kb_fill_int_kernel, dim3(blocks_for(n_inst, 256)), dim3(256), 0, s, state,
                       n_inst, (int)KB_ACTIVE);
    hipLaunchKernelGGL(kb_fill_int_kernel, dim3(blocks_for(n_inst, 256)), dim3(256), 0, s, fpar,
                       n_inst, 0);
    hipLaunchKernelGGL(kb_fill_int_kernel, dim3(blocks_for(n_inst, 256)), dim3(256), 0, s, n_iter,
                       n_inst, 0);
    return hipGetLastError();
}

template <class Op>
static void kb_assign_op(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int mode,
                         int par, int pt, int grid, hipStream_t s) {
    using T = typename Op::T;
    T *Cb = static_cast<T *>(w.Cb);
    const T *Xo = static_cast<const T *>(w.Xo);
    hipLaunchKernelGGL(kb_centre_prep_kernel<Op>, dim3(n_inst), dim3(64), 0, s, b, w.C0, w.C1,
                       w.state, w.fpar, mode, par, Cb, w.chalf, w.cmax2);
    if (pt == 2)
        hipLaunchKernelGGL((kb_assign_kernel<Op, 2>), dim3(grid), dim3(256), 0, s, Xo, b, w.tiles,
                           w.ntiles, w.state, w.fpar, mode, par, Cb, w.chalf, w.cmax2, w.L0, w.L1,
                           w.changed, w.list_count, w.list);
    else
        hipLaunchKernelGGL((kb_assign_kernel<Op, 4>), dim3(grid), dim3(256), 0, s, Xo, b, w.tiles,
                           w.ntiles, w.state, w.fpar, mode, par, Cb, w.chalf, w.cmax2, w.L0, w.L1,
                           w.changed, w.list_count, w.list);
}

hipError_t kb_launch_assign(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int mode,
                            int par, int pt, int grid, hipStream_t s) {
    hipLaunchKernelGGL(kb_reset_kernel, dim3(blocks_for(n_inst, 256)), dim3(256), 0, s,
                       w.list_count, w.changed, n_inst);
    if (w.prec == 1)
        kb_assign_op<KbBF16>(w, b, n_inst, mode, par, pt, grid, s);
    else if (w.prec == 2)
        kb_assign_op<KbBF16X3>(w, b, n_inst, mode, par, pt, grid, s);
    else
        kb_assign_op<KbF32>(w, b, n_inst, mode, par, pt, grid, s);
    hipLaunchKernelGGL(kb_recheck_kernel, dim3(1024), dim3(256), 0, s, w.X, b, w.C0, w.C1, w.fpar,
                       mode, par, w.list_count, w.list, w.L0, w.L1, w.changed);
    return hipGetLastError();
}

hipError_t kb_launch_update(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int par,
                            int iter, double tol, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(kb_partial_kernel, dim3(w.nchunks), dim3(1024), lds, s, w.X, b, w.chunks,
                       w.poff, w.state, par, w.L0, w.L1, w.escale, w.part, w.pcnt);
    hipLaunchKernelGGL(kb_reduce_kernel, dim3(n_inst), dim3(256), 0, s, b, w.cbeg, w.cend, w.poff,
                       w.part, w.pcnt, w.escale, par, iter, tol, w.C0, w.C1, w.changed,
                       w.state, w.fpar, w.n_iter);
    return hipGetLastError();
}

hipError_t kb_launch_finish(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int pt,
                            int grid, float *Cout, int *labels, double *inertia, hipStream_t s) {
    hipLaunchKernelGGL(kb_maxit_kernel, dim3(blocks_for(n_inst, 256)), dim3(256), 0, s, w.state,
                       n_inst);
    hipError_t e = kb_launch_assign(w, b, n_inst, 1, 0, pt, grid, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(kb_inertia_kernel, dim3(w.nchunks), dim3(256), 0, s, w.X, b, w.chunks, w.C0,
                       w.C1, w.state, w.fpar, w.L0, w.L1, w.cin);
    hipLaunchKernelGGL(kb_output_kernel, dim3(n_inst), dim3(256), 0, s, b, w.C0, w.C1, w.state,
                       w.fpar, w.L0, w.L1, w.cbeg, w.cend, w.cin, Cout, labels, inertia);
    return hipGetLastError();
}

hipError_t kb_set_partial_lds(size_t lds) {
    if (lds > 160 * 1024)
        return hipErrorInvalidValue;
    return hipFuncSetAttribute((const void *)kb_partial_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int kb_assign_blocks_per_cu(int prec, int pt) {
    int per_cu = 0;
    hipError_t e;
    if (prec == 1)
        e = pt == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbBF16, 2>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbBF16, 4>, 256, 0);
    else if (prec == 2)
        e = pt == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbBF16X3, 2>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbBF16X3, 4>, 256, 0);
    else
        e = pt == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbF32, 2>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_assign_kernel<KbF32, 4>, 256, 0);
    return e == hipSuccess && per_cu > 0 ? per_cu : 1;
}

} // namespace dagpu
