// Fused brute-force kNN for gfx950: MFMA distance tiles with per-query candidate buffers,
// then an exact re-rank that certifies the result.
//
// The reference computes, per (training block x query block) tile, the squared distances
// and keeps each query's k smallest with a replace-the-maximum scan
// (source/core/nearest_neighbors/nearest_neighbors.cpp:567-816,
// nearest_neighbors_utils.hpp:44-101). Here the ranking key of training row y for query q
// is ||y||^2 - 2 q.y (||q||^2 is common to the row), computed with the
// v_mfma_f32_16x16x4f32 contraction, and never materialised:
//
//  knn_tile: a 256-thread workgroup owns 128 queries (each wave 32, kept in registers as
//    MFMA A operands) and one contiguous split of the training rows. It walks the split
//    32 rows at a time; every key below its query's running threshold is appended to that
//    query's LDS buffer. When a buffer nears capacity the owning wave bitonic-sorts it in
//    registers and keeps the KC best, which sets the new threshold. At the end each query's
//    KC best keys and the split threshold tau (every rejected key is >= tau) are written.
//  knn_merge: one wave per query merges the splits' candidates, bounds the error E of the
//    fp32 keys, re-ranks every candidate within 2E of the k-th key with exact double
//    distances and orders them by (distance, index). The result is certified exact when
//    min(tau) > k-th key + 2E; otherwise the query is listed for the exact fallback.
//
// K-permutation: for MFMA step s, lane group g = lane/16 supplies feature g*(Kp/4) + s, so a
// lane's operands are one contiguous slice of its row and load as 16-byte vectors.
#include "launch.hpp"
#include <climits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;
constexpr int ROWT = 2;                   // 16-row query tiles per wave
constexpr int COLT = 2;                   // 16-row training tiles per step
constexpr int QB = WAVES * ROWT * 16;     // queries per workgroup
constexpr int STEP = COLT * 16;           // training rows per step

__device__ __forceinline__ void wave_sort64(float &key, int &idx) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const float ok = __shfl_xor(key, stride, 64);
            const int oi = __shfl_xor(idx, stride, 64);
            const bool up = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            const bool other_less = (ok < key) || (ok == key && oi < idx);
            const bool take = lower ? (up ? other_less : !other_less) : (up ? !other_less : other_less);
            if (take) {
                key = ok;
                idx = oi;
            }
        }
    }
}

// Compacts query q's buffer (count c <= CAP = 64) to its KC best entries.
template <int KC>
__device__ __forceinline__ void compact_query(float *bkey, int *bidx, int *cnt, float *thr, int q) {
    const int lane = threadIdx.x & 63;
    const int c = cnt[q];
    float key = lane < c ? bkey[q * 64 + lane] : __builtin_huge_valf();
    int idx = lane < c ? bidx[q * 64 + lane] : INT_MAX;
    wave_sort64(key, idx);
    if (lane < KC) {
        bkey[q * 64 + lane] = key;
        bidx[q * 64 + lane] = idx;
    }
    const float kth = __shfl(key, KC - 1, 64);
    if (lane == 0) {
        cnt[q] = c < KC ? c : KC;
        thr[q] = c >= KC ? kth : __builtin_huge_valf();
    }
}

template <int KCH, int KC>
__global__ __launch_bounds__(256) void knn_tile_f32_kernel(
    const float *__restrict__ Q, int nq, const float *__restrict__ Y, const float *__restrict__ ny,
    int n, int Kp, int split_len, int nsplit, float *__restrict__ part_key,
    int *__restrict__ part_idx, float *__restrict__ part_tau) {
    __shared__ float bkey[QB * 64];
    __shared__ int bidx[QB * 64];
    __shared__ int cnt[QB];
    __shared__ float thr[QB];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int qblock = blockIdx.x / nsplit;
    const int split = blockIdx.x - qblock * nsplit;
    const int q0 = qblock * QB;
    const int t_begin = split * split_len;
    int t_end = t_begin + split_len;
    if (t_end > n)
        t_end = n;

    for (int t = threadIdx.x; t < QB; t += 256) {
        cnt[t] = 0;
        thr[t] = __builtin_huge_valf();
    }

    // A operands: this lane's slice of its query rows, resident for the whole sweep.
    float a[ROWT][KCH];
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
        const int qr = q0 + (w * ROWT + rt) * 16 + r;
        const float *src = Q + (int64_t)qr * Kp + g * KCH;
#pragma unroll
        for (int f = 0; f < KCH; f += 4) {
            v4f v = {0.f, 0.f, 0.f, 0.f};
            if (qr < nq)
                v = *reinterpret_cast<const v4f *>(src + f);
            a[rt][f] = v[0];
            a[rt][f + 1] = v[1];
            a[rt][f + 2] = v[2];
            a[rt][f + 3] = v[3];
        }
    }
    __syncthreads();

    // Query slots of this lane's accumulator rows: 4*g + i of each row tile.
    const int qbase = (w * ROWT) * 16;
    for (int tb = t_begin; tb < t_end; tb += STEP) {
        v4f acc[ROWT][COLT];
        float nyv[COLT];
        int col[COLT];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++) {
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
                acc[rt][ct] = (v4f){0.f, 0.f, 0.f, 0.f};
            const int j = tb + ct * 16 + r;
            col[ct] = j;
            nyv[ct] = j < t_end ? ny[j] : __builtin_huge_valf();
            const float *src = Y + (int64_t)(j < t_end ? j : t_begin) * Kp + g * KCH;
#pragma unroll
            for (int f = 0; f < KCH; f += 4) {
                const v4f b = *reinterpret_cast<const v4f *>(src + f);
#pragma unroll
                for (int e = 0; e < 4; e++)
#pragma unroll
                    for (int rt = 0; rt < ROWT; rt++)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][f + e], b[e],
                                                                           acc[rt][ct], 0, 0, 0);
            }
        }
        // Epilogue: keys against the running thresholds.
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int q = qbase + rt * 16 + 4 * g + i;
                const float th = thr[q];
#pragma unroll
                for (int ct = 0; ct < COLT; ct++) {
                    const float key = fmaf(-2.0f, acc[rt][ct][i], nyv[ct]);
                    if (key < th) {
                        const int pos = atomicAdd(&cnt[q], 1);
                        bkey[q * 64 + pos] = key;
                        bidx[q * 64 + pos] = col[ct];
                    }
                }
            }
        }
        __syncthreads();
        // Compact the buffers that could overflow in the next step.
        const int myq = qbase + (lane & 31);
        const bool need = (lane < ROWT * 16) && (cnt[myq] > 64 - STEP);
        unsigned long long mask = __ballot(need);
        while (mask) {
            const int b = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            compact_query<KC>(bkey, bidx, cnt, thr, qbase + b);
        }
        __syncthreads();
    }
    // Final compaction and output of every query of this wave.
    for (int b = 0; b < ROWT * 16; b++)
        compact_query<KC>(bkey, bidx, cnt, thr, qbase + b);
    __syncthreads();
    for (int b = 0; b < ROWT * 16; b++) {
        const int q = q0 + qbase + b;
        if (q >= nq)
            continue;
        const int ql = qbase + b;
        const int64_t o = ((int64_t)q * nsplit + split) * KC;
        if (lane < KC) {
            const bool valid = lane < cnt[ql];
            part_key[o + lane] = valid ? bkey[ql * 64 + lane] : __builtin_huge_valf();
            part_idx[o + lane] = valid ? bidx[ql * 64 + lane] : -1;
        }
        if (lane == 0)
            part_tau[(int64_t)q * nsplit + split] = thr[ql];
    }
}

template <typename T>
__global__ __launch_bounds__(64) void knn_merge_kernel(
    const float *__restrict__ part_key, const int *__restrict__ part_idx,
    const float *__restrict__ part_tau, int nsplit, int kc, const T *__restrict__ Q,
    const T *__restrict__ Yrows, int d, int k, double unit, double ymax, int post,
    T *__restrict__ out_dist, int *__restrict__ out_idx, int *__restrict__ fb_count,
    int *__restrict__ fb_list) {
    __shared__ float skey[256];
    __shared__ int sidx[256];
    __shared__ double qrow[1024];
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const int m = nsplit * kc; // <= 256
    const int64_t base = (int64_t)q * m;
    for (int t = lane; t < 256; t += 64) {
        skey[t] = t < m ? part_key[base + t] : __builtin_huge_valf();
        sidx[t] = t < m ? part_idx[base + t] : INT_MAX;
        if (skey[t] != skey[t]) {
            skey[t] = __builtin_huge_valf();
        }
    }
    float tau = __builtin_huge_valf();
    for (int s = lane; s < nsplit; s += 64)
        tau = fminf(tau, part_tau[(int64_t)q * nsplit + s]);
    for (int off = 32; off > 0; off >>= 1)
        tau = fminf(tau, __shfl_xor(tau, off, 64));
    double qn = 0.0;
    for (int l = lane; l < d; l += 64) {
        const double v = (double)Q[(int64_t)q * d + l];
        qrow[l] = v;
        qn = fma(v, v, qn);
    }
    for (int off = 32; off > 0; off >>= 1)
        qn += __shfl_xor(qn, off, 64);
    __syncthreads();
    // Bitonic sort of the 256 (key, index) candidates, ascending.
    for (int size = 2; size <= 256; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < 128; t += 64) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const float ki = skey[i], kj = skey[j];
                const int ii = sidx[i], ij = sidx[j];
                const bool gt = (kj < ki) || (kj == ki && ij < ii);
                if (gt == asc) {
                    skey[i] = kj;
                    skey[j] = ki;
                    sidx[i] = ij;
                    sidx[j] = ii;
                }
            }
            __syncthreads();
        }
    }
    // Error bound of the fp32 key ||y||^2 - 2 q.y: products and a d-term accumulation in
    // fp32 (gamma_{d+1}), the rounded norm and the final fma.
    const double qnorm = sqrt(qn);
    const double gam = (d + 1) * unit / (1.0 - (d + 1) * unit);
    const double E = (2.0 * gam + 2.5 * unit) * qnorm * ymax + 2.5 * unit * ymax * ymax;
    const float ak = skey[k - 1];
    const double lim = (double)ak + 2.0 * E;
    // Candidates within 2E of the k-th key, re-ranked with exact distances.
    int nr = 0;
    while (nr < 256 && (double)skey[nr] <= lim)
        nr++;
    const bool certified = (double)tau > lim && nr <= 128;
    if (!certified) {
        if (lane == 0) {
            const int p = atomicAdd(fb_count, 1);
            fb_list[p] = q;
        }
        return;
    }
    double ex[2];
    int ei[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int t = lane + 64 * h;
        ex[h] = __builtin_huge_val();
        ei[h] = INT_MAX;
        if (t < nr) {
            const int j = sidx[t];
            const T *y = Yrows + (int64_t)j * d;
            double acc = 0.0;
            for (int l = 0; l < d; l++) {
                const double dv = qrow[l] - (double)y[l];
                acc = fma(dv, dv, acc);
            }
            ex[h] = acc;
            ei[h] = j;
        }
    }
    __syncthreads();
    // Sort the <= 128 exact candidates by (distance, index) through LDS (reusing qrow).
    double *xk = qrow;
    int *xi = sidx;
    for (int h = 0; h < 2; h++) {
        xk[lane + 64 * h] = ex[h];
        xi[lane + 64 * h] = ei[h];
    }
    __syncthreads();
    for (int size = 2; size <= 128; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int t = lane;
            const int i = 2 * t - (t & (stride - 1));
            const int j = i + stride;
            const bool asc = (i & size) == 0;
            const double ki = xk[i], kj = xk[j];
            const int ii = xi[i], ij = xi[j];
            const bool gt = (kj < ki) || (kj == ki && ij < ii);
            if (gt == asc) {
                xk[i] = kj;
                xk[j] = ki;
                xi[i] = ij;
                xi[j] = ii;
            }
            __syncthreads();
        }
    }
    for (int t = lane; t < k; t += 64) {
        double v = xk[t];
        if (post == 1)
            v = sqrt(v);
        if (out_dist)
            out_dist[(int64_t)q * k + t] = (T)v;
        out_idx[(int64_t)q * k + t] = xi[t];
    }
}

} // namespace

int knn_tile_kc(int k) { return k <= 16 ? 16 : 32; }

hipError_t launch_knn_tile_f32(const float *Q, int nq, const float *Y, const float *ny, int n,
                               int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                               float *part_tau, hipStream_t s) {
    int split_len = (n + nsplit - 1) / nsplit;
    split_len = (split_len + STEP - 1) / STEP * STEP;
    const int qblocks = (nq + QB - 1) / QB;
    dim3 grid(qblocks * nsplit), block(256);
    const int kch = Kp / 4;
#define DA_TILE(KCHV, KCV)                                                                     \
    hipLaunchKernelGGL((knn_tile_f32_kernel<KCHV, KCV>), grid, block, 0, s, Q, nq, Y, ny, n, Kp, \
                       split_len, nsplit, part_key, part_idx, part_tau)
#define DA_TILE_K(KCHV)                                                                        \
    if (kc == 16)                                                                              \
        DA_TILE(KCHV, 16);                                                                     \
    else                                                                                       \
        DA_TILE(KCHV, 32);
    switch (kch) {
    case 4: DA_TILE_K(4) break;
    case 8: DA_TILE_K(8) break;
    case 12: DA_TILE_K(12) break;
    case 16: DA_TILE_K(16) break;
    case 20: DA_TILE_K(20) break;
    case 24: DA_TILE_K(24) break;
    case 28: DA_TILE_K(28) break;
    case 32: DA_TILE_K(32) break;
    default:
        return hipErrorInvalidValue;
    }
#undef DA_TILE_K
#undef DA_TILE
    return hipGetLastError();
}

template <typename T>
hipError_t launch_knn_merge(const float *part_key, const int *part_idx, const float *part_tau,
                            int nsplit, int kc, const T *Q, const T *Y, int nq, int d, int k,
                            double unit, double ymax, int post, T *out_dist, int *out_idx,
                            int *fb_count, int *fb_list, hipStream_t s) {
    if (nsplit * kc > 256 || d > 1024 || k > 128)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(knn_merge_kernel<T>, dim3(nq), dim3(64), 0, s, part_key, part_idx,
                       part_tau, nsplit, kc, Q, Y, d, k, unit, ymax, post, out_dist, out_idx,
                       fb_count, fb_list);
    return hipGetLastError();
}

template hipError_t launch_knn_merge<float>(const float *, const int *, const float *, int, int,
                                            const float *, const float *, int, int, int, double,
                                            double, int, float *, int *, int *, int *,
                                            hipStream_t);

} // namespace dagpu
