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
#include <algorithm>
#include <climits>
#include <type_traits>
#include <hip/hip_runtime.h>

namespace dagpu {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int COLT = 2;                   // 16-row training tiles per step
constexpr int STEP = COLT * 16;           // training rows per step

// Tile shapes per candidate-list length KC: buffer capacity CAP per query, waves per
// workgroup WV and 16-query row tiles per wave RT. The LDS buffers (WV*RT*16*CAP*8 bytes)
// stay at or below 64 KB so that two workgroups fit a CU.
template <int KC> struct TileShape;
template <> struct TileShape<16> { static constexpr int CAP = 64, WV = 4, RT = 2; };
template <> struct TileShape<32> { static constexpr int CAP = 128, WV = 4, RT = 1; };
template <> struct TileShape<64> { static constexpr int CAP = 128, WV = 4, RT = 1; };
template <> struct TileShape<128> { static constexpr int CAP = 256, WV = 2, RT = 1; };

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

// Compacts query q's buffer (count c <= CAP) to its KC best entries, sorted by (key,
// index). CAP = 64: one entry per lane, sorted in registers; larger buffers are sorted in
// place in LDS by the owning wave.
template <int KC, int CAP>
__device__ __forceinline__ void compact_query(float *bkey, int *bidx, int *cnt, float *thr, int q) {
    const int lane = threadIdx.x & 63;
    const int c = cnt[q];
    float *kk = bkey + q * CAP;
    int *ii = bidx + q * CAP;
    if constexpr (CAP == 64) {
        float key = lane < c ? kk[lane] : __builtin_huge_valf();
        int idx = lane < c ? ii[lane] : INT_MAX;
        wave_sort64(key, idx);
        if (lane < KC) {
            kk[lane] = key;
            ii[lane] = idx;
        }
        const float kth = __shfl(key, KC - 1, 64);
        if (lane == 0) {
            cnt[q] = c < KC ? c : KC;
            thr[q] = c >= KC ? kth : __builtin_huge_valf();
        }
    } else {
        for (int t = c + lane; t < CAP; t += 64) {
            kk[t] = __builtin_huge_valf();
            ii[t] = INT_MAX;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int size = 2; size <= CAP; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = lane; t < CAP / 2; t += 64) {
                    const int i = 2 * t - (t & (stride - 1));
                    const int j = i + stride;
                    const bool asc = (i & size) == 0;
                    const float ki = kk[i], kj = kk[j];
                    const int xi = ii[i], xj = ii[j];
                    const bool gt = (kj < ki) || (kj == ki && xj < xi);
                    if (gt == asc) {
                        kk[i] = kj;
                        kk[j] = ki;
                        ii[i] = xj;
                        ii[j] = xi;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (lane == 0) {
            cnt[q] = c < KC ? c : KC;
            thr[q] = c >= KC ? kk[KC - 1] : __builtin_huge_valf();
        }
    }
}

// MFMA operand traits. E = elements per lane per instruction; rows_strided: the fp64
// instruction returns rows lane/16 + 4i, the others 4*(lane/16) + i. load<NF, CS> reads a
// lane's NF fragments as 16-byte chunks CS elements apart (CS = 16 bytes: contiguous row
// slice; CS = 1024 bytes: the tiled training layout).
struct OpF32 {
    using in_t = float;
    using frag_t = float;
    using acc_t = v4f;
    using acc_e = float;
    static constexpr int E = 1;
    static constexpr bool rows_strided = false;
    __device__ static acc_t mfma(frag_t a, frag_t b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    template <int NF, int CS> __device__ static void load(const in_t *p, frag_t *f) {
#pragma unroll
        for (int v = 0; v < NF; v += 4) {
            const v4f x = *reinterpret_cast<const v4f *>(p + (v / 4) * CS);
            f[v] = x[0];
            f[v + 1] = x[1];
            f[v + 2] = x[2];
            f[v + 3] = x[3];
        }
    }
};
struct OpF64 {
    using in_t = double;
    using frag_t = double;
    using acc_t = v4d;
    using acc_e = double;
    static constexpr int E = 1;
    static constexpr bool rows_strided = true;
    __device__ static acc_t mfma(frag_t a, frag_t b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    template <int NF, int CS> __device__ static void load(const in_t *p, frag_t *f) {
#pragma unroll
        for (int v = 0; v < NF; v += 2) {
            const v2d x = *reinterpret_cast<const v2d *>(p + (v / 2) * CS);
            f[v] = x[0];
            f[v + 1] = x[1];
        }
    }
};
struct OpF16 {
    using in_t = _Float16;
    using frag_t = v8h;
    using acc_t = v4f;
    using acc_e = float;
    static constexpr int E = 8;
    static constexpr bool rows_strided = false;
    __device__ static acc_t mfma(frag_t a, frag_t b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    template <int NF, int CS> __device__ static void load(const in_t *p, frag_t *f) {
#pragma unroll
        for (int v = 0; v < NF; v++)
            f[v] = *reinterpret_cast<const v8h *>(p + v * CS);
    }
};
struct OpBF16 {
    using in_t = __bf16;
    using frag_t = v8bf;
    using acc_t = v4f;
    using acc_e = float;
    static constexpr int E = 8;
    static constexpr bool rows_strided = false;
    __device__ static acc_t mfma(frag_t a, frag_t b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    template <int NF, int CS> __device__ static void load(const in_t *p, frag_t *f) {
#pragma unroll
        for (int v = 0; v < NF; v++)
            f[v] = *reinterpret_cast<const v8bf *>(p + v * CS);
    }
};

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Sweep of one training split for 128 queries. Scores are x.y - |y|^2/2 (the accumulator
// starts at -|y|^2/2), so a larger score is a nearer row; keys -2*score are what the
// candidate buffers, thresholds and the merge use. Each wave owns its 32 queries' buffers,
// so no workgroup barrier is needed inside the sweep. SPLIT operands carry each value as
// hi + lo in the 16-bit type (query rows stored as [hi Kp][lo Kp]); the product is
// hi*hi + hi*lo + lo*hi. Training operands are loaded from the tiled layout straight into
// registers, one step at a time (PF: the next step's loads are issued before the current
// step's MFMAs, into a second register buffer).
template <class Op, int KCH, int KC, bool SPLIT, bool PF>
__device__ __forceinline__ void knn_tile_body(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau) {
    using frag_t = typename Op::frag_t;
    using acc_t = typename Op::acc_t;
    using ae = typename Op::acc_e;
    using in_t = typename Op::in_t;
    constexpr int NF = KCH / Op::E;
    constexpr int NS = SPLIT ? 2 : 1;
    constexpr int CSR = 16 / (int)sizeof(in_t);   // chunk stride, row-major query rows
    constexpr int CST = 1024 / (int)sizeof(in_t); // chunk stride, tiled training rows
    constexpr int CAP = TileShape<KC>::CAP;
    constexpr int ROWT = TileShape<KC>::RT;
    constexpr int QB = TileShape<KC>::WV * ROWT * 16;
    __shared__ float bkey[QB * CAP];
    __shared__ int bidx[QB * CAP];
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
    const int qbase = (w * ROWT) * 16;
    const int64_t rowlen = (int64_t)Kp * NS;
    if (lane < ROWT * 16) {
        cnt[qbase + lane] = 0;
        thr[qbase + lane] = __builtin_huge_valf();
    }

    frag_t a[NS][ROWT][NF];
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
        const int qr = q0 + qbase + rt * 16 + r;
#pragma unroll
        for (int h = 0; h < NS; h++) {
            Op::template load<NF, CSR>(Q + (int64_t)(qr < nq ? qr : 0) * rowlen + h * Kp + g * KCH,
                                       a[h][rt]);
            if (qr >= nq) {
#pragma unroll
                for (int f = 0; f < NF; f++)
                    a[h][rt][f] = frag_t{};
            }
        }
    }
    ae ts[ROWT][4];
    auto row_of = [&](int i) { return Op::rows_strided ? g + 4 * i : 4 * g + i; };
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            ts[rt][i] = -__builtin_huge_val();
    wave_fence();

    // Training tile t (rows 16t .. 16t+15) is stored as [half][chunk][lane][16 bytes]: the
    // 16-byte chunk c of lane l = 16g + r is row 16t + r, features g*KCH + c*16/sizeof(in_t)
    // onward, so each load instruction of the wave reads 1 KB of contiguous memory.
    auto load_step = [&](int tb, frag_t (&b)[COLT][NS][NF], ae (&hv)[COLT]) {
#pragma unroll
        for (int ct = 0; ct < COLT; ct++) {
            const int j = tb + ct * 16 + r;
            hv[ct] = j < t_end ? nyh[j] : (ae)(-__builtin_huge_val());
            const in_t *tile = Y + (int64_t)(tb + ct * 16) * rowlen + lane * CSR;
#pragma unroll
            for (int h = 0; h < NS; h++)
                Op::template load<NF, CST>(tile + h * 16 * Kp, b[ct][h]);
        }
    };

    auto compute_step = [&](int tb, const frag_t (&b)[COLT][NS][NF], const ae (&hv)[COLT]) {
        acc_t acc[ROWT][COLT];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++)
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
                acc[rt][ct] = acc_t{hv[ct], hv[ct], hv[ct], hv[ct]};
        // All ROWT x COLT accumulators advance together: independent MFMA chains keep the
        // matrix pipeline busy while each result is still in flight.
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int ct = 0; ct < COLT; ct++)
#pragma unroll
                for (int rt = 0; rt < ROWT; rt++) {
                    acc[rt][ct] = Op::mfma(a[0][rt][f], b[ct][0][f], acc[rt][ct]);
                    if constexpr (SPLIT) {
                        acc[rt][ct] = Op::mfma(a[0][rt][f], b[ct][NS - 1][f], acc[rt][ct]);
                        acc[rt][ct] = Op::mfma(a[NS - 1][rt][f], b[ct][0][f], acc[rt][ct]);
                    }
                }
        }
        bool hit = false;
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
                    hit |= acc[rt][ct][i] > ts[rt][i];
        if (__ballot(hit) == 0)
            return;
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int q = qbase + rt * 16 + row_of(i);
#pragma unroll
                for (int ct = 0; ct < COLT; ct++) {
                    if (acc[rt][ct][i] > ts[rt][i]) {
                        const int pos = atomicAdd(&cnt[q], 1);
                        bkey[q * CAP + pos] = (float)(-2.0 * (double)acc[rt][ct][i]);
                        bidx[q * CAP + pos] = tb + ct * 16 + r;
                    }
                }
            }
        }
        wave_fence();
        const bool need = (lane < ROWT * 16) && (cnt[qbase + (lane & 31)] > CAP - STEP);
        unsigned long long mask = __ballot(need);
        if (mask) {
            while (mask) {
                const int bq = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                compact_query<KC, CAP>(bkey, bidx, cnt, thr, qbase + bq);
            }
            wave_fence();
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    ts[rt][i] = (ae)(-0.5 * (double)thr[qbase + rt * 16 + row_of(i)]);
        }
    };

    if constexpr (!PF) {
        for (int tb = t_begin; tb < t_end; tb += STEP) {
            frag_t b0[COLT][NS][NF];
            ae h0[COLT];
            load_step(tb, b0, h0);
            // Keep every load of the step ahead of its first MFMA.
            __builtin_amdgcn_sched_barrier(0);
            compute_step(tb, b0, h0);
        }
    } else if (t_begin < t_end) {
        // Two register buffers: the loads of step i+1 are issued before the MFMAs of step i.
        frag_t b0[COLT][NS][NF], b1[COLT][NS][NF];
        ae h0[COLT], h1[COLT];
        load_step(t_begin, b0, h0);
        for (int tb = t_begin;; tb += 2 * STEP) {
            const bool more1 = tb + STEP < t_end;
            if (more1)
                load_step(tb + STEP, b1, h1);
            __builtin_amdgcn_sched_barrier(0);
            compute_step(tb, b0, h0);
            if (!more1)
                break;
            const bool more2 = tb + 2 * STEP < t_end;
            if (more2)
                load_step(tb + 2 * STEP, b0, h0);
            __builtin_amdgcn_sched_barrier(0);
            compute_step(tb + STEP, b1, h1);
            if (!more2)
                break;
        }
    }
    for (int b = 0; b < ROWT * 16; b++)
        compact_query<KC, CAP>(bkey, bidx, cnt, thr, qbase + b);
    wave_fence();
    for (int b = 0; b < ROWT * 16; b++) {
        const int q = q0 + qbase + b;
        if (q >= nq)
            continue;
        const int ql = qbase + b;
        const int64_t o = ((int64_t)q * nsplit + split) * KC;
        for (int t = lane; t < KC; t += 64) {
            const bool valid = t < cnt[ql];
            part_key[o + t] = valid ? bkey[ql * CAP + t] : __builtin_huge_valf();
            part_idx[o + t] = valid ? bidx[ql * CAP + t] : -1;
        }
        if (lane == 0)
            part_tau[(int64_t)q * nsplit + split] = thr[ql];
    }
}

// The sweep as a kernel: K2 asks the compiler for at most 256 registers per lane, two waves
// per SIMD (the LDS of two workgroups fits a CU); the fp64 operands need more registers
// and run one wave per SIMD without spilling.
template <class Op, int KCH, int KC, bool SPLIT, bool PF>
__global__ __launch_bounds__(64 * TileShape<KC>::WV) void knn_tile_kernel(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau) {
    knn_tile_body<Op, KCH, KC, SPLIT, PF>(Q, nq, Y, nyh, n, Kp, split_len, nsplit, part_key,
                                          part_idx, part_tau);
}
template <class Op, int KCH, int KC, bool SPLIT, bool PF>
__global__ __launch_bounds__(64 * TileShape<KC>::WV)
__attribute__((amdgpu_waves_per_eu(2))) void knn_tile_kernel_k2(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau) {
    knn_tile_body<Op, KCH, KC, SPLIT, PF>(Q, nq, Y, nyh, n, Kp, split_len, nsplit, part_key,
                                          part_idx, part_tau);
}

// Wave-uniform wait for this wave's outstanding vector-memory operations to drop to N and
// for its LDS operations to finish, then a workgroup barrier. Inline assembly keeps the
// compiler from moving LDS accesses across the barrier.
template <int N> __device__ __forceinline__ void wait_vm_barrier() {
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

__device__ __forceinline__ void wait_vm_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Direct global-to-LDS copy of 16 bytes per lane (4 bytes per lane for the dword form): lane
// l's bytes land at LDS offset lds + 16 l (4 l). lds must be wave-uniform. Issued as inline
// assembly so that the compiler, which does not see the LDS write, adds no waits of its own
// for it: every read of the destination follows a wait_vm_barrier that covers the copy.
__device__ __forceinline__ void lds_copy16(const void *src, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                 ::"s"(lds), "v"(src) : "memory", "m0");
}
__device__ __forceinline__ void lds_copy4(const void *src, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off"
                 ::"s"(lds), "v"(src) : "memory", "m0");
}
template <typename P> __device__ __forceinline__ uint32_t lds_offset(P *p) {
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) char *)(p);
}

// LDS-staged sweep. The same work as knn_tile_kernel, but the training rows stream through
// an NSTAGE-deep ring of LDS stages (32 rows each) filled by direct global-to-LDS loads
// issued NSTAGE-1 steps ahead and shared by all waves of the workgroup. Each wave copies
// CPW 1 KB chunks of every stage (one 16-byte load per lane and chunk: the tiled layout
// makes a chunk contiguous in memory and in LDS) and its share of the stage's -|y|^2/2
// values. A step starts with a wait for this wave's copies of the stage and a barrier; the
// slot refilled after that barrier is the one every wave finished reading in the previous
// step.
template <class Op, int KCH, int KC, bool SPLIT, int NSTAGE>
__global__ __launch_bounds__(64 * TileShape<KC>::WV) void knn_tile_staged_kernel(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau) {
    using in_t = typename Op::in_t;
    using frag_t = typename Op::frag_t;
    using acc_t = typename Op::acc_t;
    using ae = typename Op::acc_e;
    constexpr int WV = TileShape<KC>::WV;
    constexpr int NF = KCH / Op::E;
    constexpr int NS = SPLIT ? 2 : 1;
    constexpr int KPC = 4 * KCH;                                 // padded features
    constexpr int RB = KPC * (int)sizeof(in_t) * NS;             // bytes per row
    constexpr int SBY = STEP * RB;                               // bytes per stage
    constexpr int CPW = SBY / 1024 / WV;                         // 1 KB chunks per wave
    constexpr int NYL = STEP * (int)sizeof(ae) / 4 / WV;         // nyh dwords per wave
    constexpr int CSR = 16 / (int)sizeof(in_t);
    constexpr int CST = 1024 / (int)sizeof(in_t);
    constexpr int CAP = TileShape<KC>::CAP;
    constexpr int ROWT = TileShape<KC>::RT;
    constexpr int QB = WV * ROWT * 16;
    constexpr int VMW = (NSTAGE - 2) * (CPW + 1);
    static_assert(CPW >= 1 && CPW * WV * 1024 == SBY, "stage must split into whole chunks");
    static_assert(NYL >= 1 && NYL <= 64, "nyh share");
    static_assert((NSTAGE - 1) * (CPW + 1) <= 63, "vmcnt range");
    __shared__ float bkey[QB * CAP];
    __shared__ int bidx[QB * CAP];
    __shared__ int cnt[QB];
    __shared__ float thr[QB];
    __shared__ __attribute__((aligned(1024))) char ring[NSTAGE * SBY];
    __shared__ __attribute__((aligned(16))) ae nh[NSTAGE * STEP];
    (void)Kp;

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
    const int nsteps = t_end > t_begin ? (t_end - t_begin + STEP - 1) / STEP : 0;
    const int qbase = (w * ROWT) * 16;
    const int64_t rowlen = (int64_t)KPC * NS;
    if (lane < ROWT * 16) {
        cnt[qbase + lane] = 0;
        thr[qbase + lane] = __builtin_huge_valf();
    }

    frag_t a[NS][ROWT][NF];
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
        const int qr = q0 + qbase + rt * 16 + r;
#pragma unroll
        for (int h = 0; h < NS; h++) {
            Op::template load<NF, CSR>(Q + (int64_t)(qr < nq ? qr : 0) * rowlen + h * KPC + g * KCH,
                                       a[h][rt]);
            if (qr >= nq) {
#pragma unroll
                for (int f = 0; f < NF; f++)
                    a[h][rt][f] = frag_t{};
            }
        }
    }
    // Consume the query operands here, so that the compiler's waits for these loads come
    // before the sweep and not inside it, where they would also wait for the stage copies.
#pragma unroll
    for (int h = 0; h < NS; h++)
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
            for (int f = 0; f < NF; f++)
                asm volatile("" ::"v"(a[h][rt][f]));
    ae ts[ROWT][4];
    auto row_of = [&](int i) { return Op::rows_strided ? g + 4 * i : 4 * g + i; };
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            ts[rt][i] = -__builtin_huge_val();
    wave_fence();
    const int t_last = t_begin + (nsteps > 0 ? nsteps - 1 : 0) * STEP;
    const uint32_t ring_lds = lds_offset(ring);
    const uint32_t nh_lds = lds_offset(nh);

    // Copies stage s of the split (rows t_begin + 32 s ..) into ring slot s % NSTAGE.
    // Stages past the end re-read the last stage, so every wave issues the same number of
    // copies per step.
    auto issue = [&](int s) {
        int tb = t_begin + s * STEP;
        if (tb > t_last)
            tb = t_last;
        const int slot = s % NSTAGE;
        const char *src = reinterpret_cast<const char *>(Y) + (int64_t)tb * RB + lane * 16;
#pragma unroll
        for (int c = 0; c < CPW; c++) {
            const int chunk = w * CPW + c;
            lds_copy16(src + chunk * 1024,
                       __builtin_amdgcn_readfirstlane(ring_lds + slot * SBY + chunk * 1024));
        }
        const uint32_t ndst =
            __builtin_amdgcn_readfirstlane(nh_lds + slot * STEP * (int)sizeof(ae) + w * NYL * 4);
        const char *nsrc = reinterpret_cast<const char *>(nyh + tb) + (w * NYL + (lane < NYL ? lane : 0)) * 4;
        if (lane < NYL)
            lds_copy4(nsrc, ndst);
    };

    auto compute = [&](int tb, int slot) {
        const char *stage = ring + slot * SBY;
        const ae *nhs = nh + slot * STEP;
        acc_t acc[ROWT][COLT];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++) {
            const ae hv = nhs[ct * 16 + r];
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
                acc[rt][ct] = acc_t{hv, hv, hv, hv};
        }
        frag_t b[COLT][NS][NF];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++)
#pragma unroll
            for (int h = 0; h < NS; h++)
                Op::template load<NF, CST>(
                    reinterpret_cast<const in_t *>(stage + ct * 16 * RB) + h * 16 * KPC + lane * CSR,
                    b[ct][h]);
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int ct = 0; ct < COLT; ct++)
#pragma unroll
                for (int rt = 0; rt < ROWT; rt++) {
                    acc[rt][ct] = Op::mfma(a[0][rt][f], b[ct][0][f], acc[rt][ct]);
                    if constexpr (SPLIT) {
                        acc[rt][ct] = Op::mfma(a[0][rt][f], b[ct][NS - 1][f], acc[rt][ct]);
                        acc[rt][ct] = Op::mfma(a[NS - 1][rt][f], b[ct][0][f], acc[rt][ct]);
                    }
                }
        }
        bool hit = false;
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
                    hit |= acc[rt][ct][i] > ts[rt][i];
        if (__ballot(hit) == 0)
            return;
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int q = qbase + rt * 16 + row_of(i);
#pragma unroll
                for (int ct = 0; ct < COLT; ct++) {
                    if (acc[rt][ct][i] > ts[rt][i]) {
                        const int pos = atomicAdd(&cnt[q], 1);
                        bkey[q * CAP + pos] = (float)(-2.0 * (double)acc[rt][ct][i]);
                        bidx[q * CAP + pos] = tb + ct * 16 + r;
                    }
                }
            }
        }
        wave_fence();
        const bool need = (lane < ROWT * 16) && (cnt[qbase + (lane & 31)] > CAP - STEP);
        unsigned long long mask = __ballot(need);
        if (mask) {
            while (mask) {
                const int bq = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                compact_query<KC, CAP>(bkey, bidx, cnt, thr, qbase + bq);
            }
            wave_fence();
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    ts[rt][i] = (ae)(-0.5 * (double)thr[qbase + rt * 16 + row_of(i)]);
        }
    };

    if (nsteps > 0) {
        for (int s = 0; s < NSTAGE - 1; s++)
            issue(s);
        for (int i = 0; i < nsteps; i++) {
            wait_vm_barrier<VMW>();
            issue(i + NSTAGE - 1);
            compute(t_begin + i * STEP, i % NSTAGE);
        }
        // Copies of stages past the end may still be in flight into LDS.
        wait_vm_all();
    }
    for (int b = 0; b < ROWT * 16; b++)
        compact_query<KC, CAP>(bkey, bidx, cnt, thr, qbase + b);
    wave_fence();
    for (int b = 0; b < ROWT * 16; b++) {
        const int q = q0 + qbase + b;
        if (q >= nq)
            continue;
        const int ql = qbase + b;
        const int64_t o = ((int64_t)q * nsplit + split) * KC;
        for (int t = lane; t < KC; t += 64) {
            const bool valid = t < cnt[ql];
            part_key[o + t] = valid ? bkey[ql * CAP + t] : __builtin_huge_valf();
            part_idx[o + t] = valid ? bidx[ql * CAP + t] : -1;
        }
        if (lane == 0)
            part_tau[(int64_t)q * nsplit + split] = thr[ql];
    }
}

constexpr int MMAX = 2048; // candidates merged per query (nsplit * kc)
constexpr int RMAX = 256; // candidates re-ranked exactly per query

// MM: candidate capacity of this instance (a power of two >= nsplit * kc).
template <typename T, int MM>
__global__ __launch_bounds__(64) void knn_merge_kernel(
    const float *__restrict__ part_key, const int *__restrict__ part_idx,
    const float *__restrict__ part_tau, int nsplit, int kc, const T *__restrict__ Q,
    const T *__restrict__ Yrows, int d, int k, double u_in, double abs_in, double u_acc,
    double ymax, int post,
    T *__restrict__ out_dist, int *__restrict__ out_idx, int *__restrict__ fb_count,
    int *__restrict__ fb_list) {
    __shared__ float skey[MM];
    __shared__ int sidx[MM];
    __shared__ double qrow[1024];
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const int m = nsplit * kc; // <= MM
    const int64_t base = (int64_t)q * m;
    for (int t = lane; t < MM; t += 64) {
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
    // Bitonic sort of the (key, index) candidates, ascending.
    for (int size = 2; size <= MM; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < MM / 2; t += 64) {
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
    // Error bound of the stored key |y|^2 - 2 q.y. With S = sum|q_l y_l| <= |q| ymax: the
    // operands rounded to the input precision (2u_in + u_in^2) S; the d+1-term accumulation
    // from -|y|^2/2 (gamma_{d+2}); the rounded -|y|^2/2; the float storage of the key.
    const double qnorm = sqrt(qn);
    const double S = qnorm * ymax;
    const double gam = (d + 2) * u_acc / (1.0 - (d + 2) * u_acc);
    const double u32 = 1.0 / 16777216.0;
    const double E = 2.0 * ((2.0 * u_in + u_in * u_in) * S +
                            gam * (S * (1.0 + u_in) * (1.0 + u_in) + 0.5 * ymax * ymax) +
                            u_acc * 0.5 * ymax * ymax) +
                     u32 * (ymax * ymax + 2.01 * S) +
                     // operands below the normal range of fp16 carry an absolute error
                     2.0 * abs_in * sqrt((double)d) * (qnorm + ymax) * 1.01;
    const float ak = skey[k - 1];
    const double lim = (double)ak + 2.0 * E;
    // Candidates within 2E of the k-th key, re-ranked with exact distances.
    int nr = 0;
    while (nr < MM && (double)skey[nr] <= lim)
        nr++;
    const bool certified = (double)tau > lim && nr <= RMAX;
    if (!certified) {
        if (lane == 0) {
            const int p = atomicAdd(fb_count, 1);
            fb_list[p] = q;
        }
        return;
    }
    double ex[RMAX / 64];
    int ei[RMAX / 64];
#pragma unroll
    for (int h = 0; h < RMAX / 64; h++) {
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
    // Sort the <= RMAX exact candidates by (distance, index) through LDS (reusing qrow).
    double *xk = qrow;
    int *xi = sidx;
    for (int h = 0; h < RMAX / 64; h++) {
        xk[lane + 64 * h] = ex[h];
        xi[lane + 64 * h] = ei[h];
    }
    __syncthreads();
    for (int size = 2; size <= RMAX; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < RMAX / 2; t += 64) {
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

namespace {

// Rearranges row-major operand rows (row_bytes each, ns halves of four lane-group slices)
// into the tiled training layout: tile t, half h, chunk c, lane l = 16g + r holds the 16
// bytes at row 16t + r, byte h*row_bytes/ns + g*row_bytes/(4ns) + 16c. Rows at or beyond
// `rows` are written as zeros. One thread per destination chunk.
__global__ void tile_operand_kernel(const uint4 *__restrict__ src, int rows, int row_bytes,
                                    int ns, uint4 *__restrict__ dst, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total)
        return;
    const int ru = row_bytes / 16;       // chunks per row = chunks per lane per tile * 64 / 16
    const int64_t t = i / (16 * ru);     // tile
    const int within = (int)(i - t * 16 * ru);
    const int l = within & 63, rest = within >> 6;
    const int nc = ru / (4 * ns);        // chunks per lane-group slice
    const int h = rest / nc, c = rest - h * nc;
    const int g = l >> 4, r = l & 15;
    const int64_t row = t * 16 + r;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (row < rows)
        v = src[row * ru + h * (ru / ns) + g * nc + c];
    dst[i] = v;
}

} // namespace

hipError_t launch_knn_tile_operand(const void *src, int rows, int row_bytes, int ns, void *dst,
                                   int rows_pad, hipStream_t s) {
    if (rows_pad % 16 != 0 || row_bytes % (64 * ns) != 0)
        return hipErrorInvalidValue;
    const int64_t total = (int64_t)rows_pad * row_bytes / 16;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(tile_operand_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       static_cast<const uint4 *>(src), rows, row_bytes, ns,
                       static_cast<uint4 *>(dst), total);
    return hipGetLastError();
}

int knn_tile_kc(int k) { return k <= 16 ? 16 : k <= 32 ? 32 : k <= 64 ? 64 : 128; }

int knn_tile_qb(int kc) {
    switch (kc) {
    case 16: return TileShape<16>::WV * TileShape<16>::RT * 16;
    case 32: return TileShape<32>::WV * TileShape<32>::RT * 16;
    case 64: return TileShape<64>::WV * TileShape<64>::RT * 16;
    default: return TileShape<128>::WV * TileShape<128>::RT * 16;
    }
}

int knn_merge_max_candidates() { return MMAX; }


// Geometry of the LDS-staged kernel: stage bytes, and whether a stage splits into whole
// 1 KB chunks per wave.
template <class Op, int KCH, int KC, bool SPLIT> struct StagedGeom {
    static constexpr int RB = 4 * KCH * (int)sizeof(typename Op::in_t) * (SPLIT ? 2 : 1);
    static constexpr int SBY = STEP * RB;
    static constexpr bool ok = SBY % (1024 * TileShape<KC>::WV) == 0;
    static constexpr int NSTAGE = SBY * 4 <= 65536 ? 4 : SBY * 3 <= 65536 ? 3 : 2;
};

// Launches the fused tile for one operand type, candidate-list length, width and variant
// (0: register operands, 1: LDS-staged where the geometry allows, else 0; 2: register
// operands with the next step's loads in flight) or, with `slots`
// set, reports how many of its workgroups the device runs at once instead.
template <class Op, int KCH, int KC, bool SPLIT>
static hipError_t launch_tile_w(const void *Q, int nq, const void *Y, const void *nyh, int n,
                                int Kp, int nsplit, float *part_key, int *part_idx,
                                float *part_tau, int variant, int *slots, hipStream_t s) {
    using in_t = typename Op::in_t;
    using ae = typename Op::acc_e;
    using G = StagedGeom<Op, KCH, KC, SPLIT>;
    using kernel_t = void (*)(const in_t *, int, const in_t *, const ae *, int, int, int, int,
                              float *, int *, float *);
    constexpr int QB = TileShape<KC>::WV * TileShape<KC>::RT * 16;
    kernel_t kern;
    if constexpr (std::is_same_v<Op, OpF64>)
        kern = variant == 2 ? knn_tile_kernel<Op, KCH, KC, SPLIT, true>
                            : knn_tile_kernel<Op, KCH, KC, SPLIT, false>;
    else
        kern = variant == 2 ? knn_tile_kernel_k2<Op, KCH, KC, SPLIT, true>
                            : knn_tile_kernel_k2<Op, KCH, KC, SPLIT, false>;
    if constexpr (G::ok) {
        if (variant == 1)
            kern = knn_tile_staged_kernel<Op, KCH, KC, SPLIT, G::NSTAGE>;
    }
    const dim3 block(64 * TileShape<KC>::WV);
    if (slots) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess)
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block.x, 0);
        *slots = e == hipSuccess ? std::max(1, cus * per_cu) : 0;
        return e;
    }
    int split_len = (n + nsplit - 1) / nsplit;
    split_len = (split_len + STEP - 1) / STEP * STEP;
    const int qblocks = (nq + QB - 1) / QB;
    const dim3 grid(qblocks * nsplit);
    hipLaunchKernelGGL(kern, grid, block, 0, s, static_cast<const in_t *>(Q), nq,
                       static_cast<const in_t *>(Y), static_cast<const ae *>(nyh), n, Kp,
                       split_len, nsplit, part_key, part_idx, part_tau);
    return hipGetLastError();
}

template <class Op, int KC, bool SPLIT>
static hipError_t launch_tile_op(const void *Q, int nq, const void *Y, const void *nyh, int n,
                                 int Kp, int nsplit, float *part_key, int *part_idx,
                                 float *part_tau, int variant, int *slots, hipStream_t s) {
    switch (Kp / 4) {
    case 8:
        return launch_tile_w<Op, 8, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 16:
        return launch_tile_w<Op, 16, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 24:
        return launch_tile_w<Op, 24, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 32:
        return launch_tile_w<Op, 32, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    default:
        return hipErrorInvalidValue;
    }
}

int knn_tile_kp(int d) {
    // Padded feature count: a multiple of 32 so that every operand type gets whole 16-byte
    // slices per lane group.
    return (d + 31) / 32 * 32;
}

template <int KC>
static hipError_t launch_tile_kc(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                                 int n, int Kp, int nsplit, float *part_key, int *part_idx,
                                 float *part_tau, int variant, int *slots, hipStream_t s) {
#define DA_OP(OP, SPLIT)                                                                          \
    launch_tile_op<OP, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau,     \
                                  variant, slots, s)
    switch (prec) {
    case KP_F32: return DA_OP(OpF32, false);
    case KP_F64: return DA_OP(OpF64, false);
    case KP_F16: return DA_OP(OpF16, false);
    case KP_BF16: return DA_OP(OpBF16, false);
    case KP_F16X3: return DA_OP(OpF16, true);
    case KP_BF16X3: return DA_OP(OpBF16, true);
    default: return hipErrorInvalidValue;
    }
#undef DA_OP
}

static hipError_t tile_dispatch(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                                int n, int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                                float *part_tau, int variant, int *slots, hipStream_t s) {
    switch (kc) {
    case 16:
        return launch_tile_kc<16>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 32:
        return launch_tile_kc<32>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 64:
        return launch_tile_kc<64>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    case 128:
        return launch_tile_kc<128>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, variant, slots, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_knn_tile(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                           int n, int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                           float *part_tau, int variant, hipStream_t s) {
    return tile_dispatch(prec, Q, nq, Y, nyh, n, Kp, nsplit, kc, part_key, part_idx, part_tau,
                         variant, nullptr, s);
}

hipError_t knn_tile_slots(int prec, int Kp, int kc, int variant, int *slots) {
    return tile_dispatch(prec, nullptr, 0, nullptr, nullptr, 0, Kp, 1, kc, nullptr, nullptr,
                         nullptr, variant, slots, nullptr);
}

template <typename T>
hipError_t launch_knn_merge(const float *part_key, const int *part_idx, const float *part_tau,
                            int nsplit, int kc, const T *Q, const T *Y, int nq, int d, int k,
                            double u_in, double abs_in, double u_acc, double ymax, int post,
                            T *out_dist, int *out_idx, int *fb_count, int *fb_list, hipStream_t s) {
    const int m = nsplit * kc;
    if (m > MMAX || d > 1024 || k > 128 || k > kc)
        return hipErrorInvalidValue;
    if (nq == 0)
        return hipSuccess;
    auto kern = m <= 256 ? knn_merge_kernel<T, 256>
              : m <= 512 ? knn_merge_kernel<T, 512>
              : m <= 1024 ? knn_merge_kernel<T, 1024>
                          : knn_merge_kernel<T, 2048>;
    hipLaunchKernelGGL(kern, dim3(nq), dim3(64), 0, s, part_key, part_idx, part_tau, nsplit, kc,
                       Q, Y, d, k, u_in, abs_in, u_acc, ymax, post, out_dist, out_idx, fb_count,
                       fb_list);
    return hipGetLastError();
}

template hipError_t launch_knn_merge<float>(const float *, const int *, const float *, int, int,
                                            const float *, const float *, int, int, int, double,
                                            double, double, double, int, float *, int *, int *,
                                            int *, hipStream_t);
template hipError_t launch_knn_merge<double>(const float *, const int *, const float *, int, int,
                                             const double *, const double *, int, int, int,
                                             double, double, double, double, int, double *, int *,
                                             int *, int *, hipStream_t);

} // namespace dagpu
