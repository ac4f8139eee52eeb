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

constexpr int STEP = 32; // split lengths are multiples of 32 rows (whole steps of any sweep)

// Candidate-buffer shape per candidate-list length: CAP entries per query, WV waves of RT
// row tiles (16 queries each) per workgroup.
template <int KC> struct TileShape;
template <> struct TileShape<16> { static constexpr int CAP = 64, WV = 4, RT = 2; };
template <> struct TileShape<32> { static constexpr int CAP = 128, WV = 4, RT = 1; };
template <> struct TileShape<64> { static constexpr int CAP = 128, WV = 4, RT = 1; };
template <> struct TileShape<128> { static constexpr int CAP = 256, WV = 2, RT = 1; };

// Whether a sweep keeps its buffered row offsets in 16 bits (splits of at most 65,536 rows):
// that cuts a workgroup's buffers from 64 to 48 KB, so three workgroups fit a CU's LDS,
// which pays only where the registers also leave room for three waves per SIMD -- the query
// fragments plus the current and prefetched training fragments (ns * nf fragments of
// frag_regs 32-bit registers, for rt query tiles and 2 x colt training tiles, colt 16-row
// tiles per step) within 144 of the 168 registers a wave then gets (the sweep's other state
// takes about 24). kc <= 64. The longest lists (kc = 128) keep 32-bit offsets: the extra
// splits that 65,536-row splits would force cost more (each split restarts a 128-entry list)
// than half a wave more per SIMD gains. The fp64 sweep runs one wave per SIMD and never
// qualifies. Evaluated by the kernels and by the host's bound on the split count.
constexpr bool tile_short_index(int kc, int rt, int ns, int nf, int frag_regs, bool f64,
                                int colt) {
    return !f64 && kc <= 64 && ns * nf * frag_regs * (rt + 2 * colt) <= 144;
}
constexpr int SPLIT16_MAX = 65536; // rows per split with 16-bit offsets

// Value of lane (lane ^ S) for S in {1, 2, 4, 8, 16, 32}, by register-to-register lane
// moves (DPP within rows of 16, the gfx950 permlane swaps across rows) instead of the LDS
// crossbar: quad_perm for 1 and 2; the two row shifts by 4, selected by lane bit 2, for 4;
// a row rotation by 8 for 8; v_permlane16_swap / v_permlane32_swap for 16 and 32.
template <int S> __device__ __forceinline__ int lane_xor(int v) {
    const int lane = threadIdx.x & 63;
    if constexpr (S == 1) {
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false); // quad_perm [1,0,3,2]
    } else if constexpr (S == 2) {
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false); // quad_perm [2,3,0,1]
    } else if constexpr (S == 4) {
        const int up = __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, false);   // row_shl:4
        const int down = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false); // row_shr:4
        return (lane & 4) ? down : up;
    } else if constexpr (S == 8) {
        return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false); // row_ror:8
    } else if constexpr (S == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16) ? (int)p[0] : (int)p[1];
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane & 32) ? (int)p[0] : (int)p[1];
    }
}

template <int SIZE, int S>
__device__ __forceinline__ void sort_stage(float &key, int &idx, int lane) {
    const float ok = __int_as_float(lane_xor<S>(__float_as_int(key)));
    const int oi = lane_xor<S>(idx);
    const bool up = (lane & SIZE) == 0;
    const bool lower = (lane & S) == 0;
    const bool other_less = (ok < key) || (ok == key && oi < idx);
    const bool take = lower ? (up ? other_less : !other_less) : (up ? !other_less : other_less);
    if (take) {
        key = ok;
        idx = oi;
    }
}

template <int SIZE, int S> __device__ __forceinline__ void sort_merge(float &key, int &idx, int lane) {
    sort_stage<SIZE, S>(key, idx, lane);
    if constexpr (S > 1)
        sort_merge<SIZE, S / 2>(key, idx, lane);
}

// Bitonic sort of one (key, index) pair per lane across the wave, ascending by (key, index).
__device__ __forceinline__ void wave_sort64(float &key, int &idx) {
    const int lane = threadIdx.x & 63;
    sort_merge<2, 1>(key, idx, lane);
    sort_merge<4, 2>(key, idx, lane);
    sort_merge<8, 4>(key, idx, lane);
    sort_merge<16, 8>(key, idx, lane);
    sort_merge<32, 16>(key, idx, lane);
    sort_merge<64, 32>(key, idx, lane);
}

// Compacts the buffer of query q (c entries, c <= CAP) to its KC best entries sorted by
// (key, index) and returns the new threshold: the KC-th key, or +inf while fewer than KC
// entries have been seen. Wave-uniform; not inlined, so that the sweep's many call sites
// share one copy.
template <int KC, int CAP, typename IT>
__device__ __attribute__((noinline)) float compact_count(float *bkey, IT *bidx, int c, int q) {
    const int lane = threadIdx.x & 63;
    float *kk = bkey + q * CAP;
    IT *ii = bidx + q * CAP;
    if constexpr (CAP == 64) {
        float key = lane < c ? kk[lane] : __builtin_huge_valf();
        int idx = lane < c ? ii[lane] : INT_MAX;
        wave_sort64(key, idx);
        if (lane < KC) {
            kk[lane] = key;
            ii[lane] = (IT)idx;
        }
        const float kth = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(key), KC - 1));
        return c >= KC ? kth : __builtin_huge_valf();
    } else {
        for (int t = c + lane; t < CAP; t += 64) {
            kk[t] = __builtin_huge_valf();
            ii[t] = (IT)INT_MAX;
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
                    const IT xi = ii[i], xj = ii[j];
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
        return c >= KC ? kk[KC - 1] : __builtin_huge_valf();
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
    static constexpr int FPC = 4; // fragments per 16-byte chunk
    template <int CS, int NF> __device__ static void load_chunk(const in_t *p, int c, frag_t *f) {
        const v4f x = *reinterpret_cast<const v4f *>(p + c * CS);
        f[4 * c] = x[0];
        f[4 * c + 1] = x[1];
        f[4 * c + 2] = x[2];
        f[4 * c + 3] = x[3];
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
    static constexpr int FPC = 2;
    template <int CS, int NF> __device__ static void load_chunk(const in_t *p, int c, frag_t *f) {
        const v2d x = *reinterpret_cast<const v2d *>(p + c * CS);
        f[2 * c] = x[0];
        f[2 * c + 1] = x[1];
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
    static constexpr int FPC = 1;
    template <int CS, int NF> __device__ static void load_chunk(const in_t *p, int c, frag_t *f) {
        f[c] = *reinterpret_cast<const v8h *>(p + c * CS);
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
    static constexpr int FPC = 1;
    template <int CS, int NF> __device__ static void load_chunk(const in_t *p, int c, frag_t *f) {
        f[c] = *reinterpret_cast<const v8bf *>(p + c * CS);
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

// Candidate bookkeeping of one wave after a step. Lane (r, g) holds, for row tile rt and
// accumulator element i, the scores of query rt*16 + row(g, i) against the training rows at
// offsets tb + ct*16 + r of the split. Scores above the query's threshold are appended to its
// LDS buffer at
// positions from a ballot: the 16 lanes of a lane group serve one query per (rt, i), so a
// hit's slot is the group's running count (cq, held in registers by the group's lanes)
// plus the number of hits of the lower lanes of the group. A buffer with more than
// CAP - COLT*16 entries (room for one more step) is compacted to its KC best, which sets
// the new threshold.
template <class Op, int KC, int CAP, int ROWT, int COLT, typename IT>
__device__ __forceinline__ void take_hits(const typename Op::acc_t (&acc)[ROWT][COLT],
                                          typename Op::acc_e (&ts)[ROWT][4],
                                          int (&cq)[ROWT][4], float *bkey, IT *bidx,
                                          int qbase, int tb) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    auto row_of = [&](int i, int gg) { return Op::rows_strided ? gg + 4 * i : 4 * gg + i; };
    const unsigned long long below = (1ull << lane) - 1;
    const unsigned long long gmask = 0xFFFFull << (16 * g);
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = qbase + rt * 16 + row_of(i, g);
#pragma unroll
            for (int ct = 0; ct < COLT; ct++) {
                const bool h = acc[rt][ct][i] > ts[rt][i];
                // The compare's own lane mask (no materialised boolean).
                const unsigned long long m = __builtin_amdgcn_ballot_w64(h);
                if (m == 0)
                    continue;
                const unsigned long long mg = m & gmask;
                if (h) {
                    const int pos = cq[rt][i] + __popcll(mg & below);
                    bkey[q * CAP + pos] = (float)(-2.0 * (double)acc[rt][ct][i]);
                    bidx[q * CAP + pos] = (IT)(tb + ct * 16 + r);
                }
                cq[rt][i] += __popcll(mg);
            }
        }
    }
    bool need = false;
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            need |= cq[rt][i] > CAP - COLT * 16;
    if (__builtin_amdgcn_ballot_w64(need) == 0)
        return;
    wave_fence();
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            // one representative lane (r = 0) per group
            unsigned long long m2 =
                __builtin_amdgcn_ballot_w64(cq[rt][i] > CAP - COLT * 16) & 0x0001000100010001ull;
            while (m2) {
                const int l0 = __ffsll((long long)m2) - 1;
                m2 &= m2 - 1;
                const int gg = l0 >> 4;
                const int c = __builtin_amdgcn_readlane(cq[rt][i], l0);
                const float t =
                    compact_count<KC, CAP>(bkey, bidx, c, qbase + rt * 16 + row_of(i, gg));
                if (g == gg) {
                    cq[rt][i] = c < KC ? c : KC;
                    ts[rt][i] = (typename Op::acc_e)(-0.5 * (double)t);
                }
            }
        }
    }
    wave_fence();
}

// Final compaction of a wave's queries and the per-(query, split) output: the KC best keys
// and row indices and the split threshold tau (every key of the split not kept is >= tau):
// the KC-th kept key, or the seeded threshold tau0 where that is lower (a row at or above the
// seed was never buffered). A buffered offset b is row t_begin + b of a contiguous split, or,
// with interleaved splits (split s sweeps the 32-row steps s, s + nsplit, ...), row
// ((b / 32) * nsplit + s) * 32 + b % 32.
template <class Op, int KC, int CAP, int ROWT, typename IT>
__device__ __forceinline__ void write_split(const int (&cq)[ROWT][4], float *bkey, IT *bidx,
                                            int t_begin, bool interleaved, int qbase, int q0,
                                            int nq, int nsplit, int split,
                                            float *__restrict__ part_key,
                                            int *__restrict__ part_idx,
                                            float *__restrict__ part_tau,
                                            const float *__restrict__ tau0,
                                            const int *__restrict__ perm) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            for (int gg = 0; gg < 4; gg++) {
                const int ql = qbase + rt * 16 + (Op::rows_strided ? gg + 4 * i : 4 * gg + i);
                const int c = __builtin_amdgcn_readlane(cq[rt][i], 16 * gg);
                const float t = compact_count<KC, CAP>(bkey, bidx, c, ql);
                wave_fence();
                const int q = q0 + ql;
                if (q >= nq)
                    continue;
                const int64_t o = ((int64_t)q * nsplit + split) * KC;
                const int valid = c < KC ? c : KC;
                for (int u = lane; u < KC; u += 64) {
                    part_key[o + u] = u < valid ? bkey[ql * CAP + u] : __builtin_huge_valf();
                    const int bo = (int)bidx[ql * CAP + u];
                    const int row = interleaved
                                        ? ((bo / STEP) * nsplit + split) * STEP + (bo & (STEP - 1))
                                        : t_begin + bo;
                    part_idx[o + u] = u < valid ? (perm ? perm[row] : row) : -1;
                }
                if (lane == 0)
                    part_tau[(int64_t)q * nsplit + split] = tau0 ? fminf(t, tau0[q]) : t;
            }
        }
    }
}

// Per-kernel configuration: COLT 16-row training tiles per step; 16-bit buffered offsets and
// three waves per SIMD, or 32-bit offsets and two.
template <class Op, int KCH, int KC, bool SPLIT, int COLT> struct TileCfg {
    static constexpr bool SHORT =
        tile_short_index(KC, TileShape<KC>::RT, SPLIT ? 2 : 1, KCH / Op::E,
                         (int)sizeof(typename Op::frag_t) / 4, std::is_same_v<Op, OpF64>, COLT);
    static constexpr int WPE = SHORT ? 3 : 2;
    using idx_t = std::conditional_t<SHORT, unsigned short, int>;
};

// Sweep of one training split for 128 queries. Scores are x.y - |y|^2/2 (the accumulator
// starts at -|y|^2/2), so a larger score is a nearer row; keys -2*score are what the
// candidate buffers, thresholds and the merge use. Each wave owns its 32 queries' buffers,
// so no workgroup barrier is needed inside the sweep. SPLIT operands carry each value as
// hi + lo in the 16-bit type (query rows stored as [hi Kp][lo Kp]); the product is
// hi*hi + hi*lo + lo*hi. Training operands are loaded from the tiled layout straight into
// registers. PF selects the schedule: 0, each step loads, then computes; 1, the next step's
// loads are issued right after the current step's MFMAs and overlap its candidate
// bookkeeping; 2 (rolling), the loads of the next step's fragment f are issued as soon as
// the current step's MFMAs that read fragment f have issued, so that every load has about a
// step of MFMA work and bookkeeping in front of its first use, with no extra registers.
template <class Op, int KCH, int KC, bool SPLIT, int PF, int COLT, bool CEN = false>
__device__ __forceinline__ void knn_tile_body(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau,
    const float *__restrict__ tau0, const KnnCentring cen) {
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
    using idx_t = typename TileCfg<Op, KCH, KC, SPLIT, COLT>::idx_t;
    constexpr int TSTEP = COLT * 16; // training rows per step
    __shared__ float bkey[QB * CAP];
    __shared__ idx_t bidx[QB * CAP];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    // XCD-aware order: workgroup b runs on XCD b % 8 and the work items (split-major: all
    // query blocks of split 0, then split 1, ...) are dealt to the XCDs in 8 contiguous
    // ranges, so the workgroups sharing an XCD's L2 sweep the same one or two splits at
    // about the same time. The grid is padded to a multiple of 8; extra workgroups exit.
    const int qblocks = (nq + QB - 1) / QB;
    const int nwork = qblocks * nsplit;
    const int per_xcd = (nwork + 7) / 8;
    const int work = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (work >= nwork)
        return;
    const int split = work / qblocks;
    const int qblock = work - split * qblocks;
    const int q0 = qblock * QB;
    const int t_begin = split * split_len;
    int t_end = t_begin + split_len;
    if (t_end > n)
        t_end = n;
    const int qbase = (w * ROWT) * 16;
    const int64_t rowlen = (int64_t)Kp * NS;
    int cq[ROWT][4] = {};

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
    // Thresholds in score space (score > ts <=> key < -2 ts): -inf, or the seeded key bound
    // of the query when one is given.
    ae ts[ROWT][4];
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = q0 + qbase + rt * 16 + (Op::rows_strided ? g + 4 * i : 4 * g + i);
            ts[rt][i] = (tau0 && q < nq) ? (ae)(-0.5 * (double)tau0[q]) : (ae)-__builtin_huge_val();
        }
    wave_fence();

    // Training tile t (rows 16t .. 16t+15) is stored as [half][chunk][lane][16 bytes]: the
    // 16-byte chunk c of lane l = 16g + r is row 16t + r, features g*KCH + c*16/sizeof(in_t)
    // onward, so each load instruction of the wave reads 1 KB of contiguous memory.
    auto load_step = [&](int tb, frag_t (&b)[COLT][NS][NF], ae (&hv)[COLT]) {
#pragma unroll
        for (int ct = 0; ct < COLT; ct++) {
            // Unconditional: a step never crosses the end of a split except at the end of
            // the data, where the padded rows have -|y|^2/2 = -inf.
            hv[ct] = nyh[tb + ct * 16 + r];
            const in_t *tile = Y + (int64_t)(tb + ct * 16) * rowlen + lane * CSR;
#pragma unroll
            for (int h = 0; h < NS; h++)
                Op::template load<NF, CST>(tile + h * 16 * Kp, b[ct][h]);
        }
    };

    auto mfma_step = [&](const frag_t (&b)[COLT][NS][NF], const ae (&hv)[COLT],
                         acc_t (&acc)[ROWT][COLT]) {
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
    };
    // Whether any score of the step beats its query's threshold (wave-uniform).
    auto any_hit = [&](const acc_t (&acc)[ROWT][COLT]) {
        bool hit = false;
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
                    hit |= acc[rt][ct][i] > ts[rt][i];
        return __builtin_amdgcn_ballot_w64(hit) != 0;
    };
    auto check_step = [&](int tb, const acc_t (&acc)[ROWT][COLT]) {
        if (any_hit(acc))
            take_hits<Op, KC, CAP, ROWT, COLT>(acc, ts, cq, bkey, bidx, qbase, tb - t_begin);
    };

    if constexpr (PF == 2) {
        // Rolling prefetch. The address of the next step's tile is computed before the
        // step's MFMAs (the last step reloads its own tile, in bounds), so that only loads
        // with immediate offsets sit between the MFMA groups: no VALU instruction writes a
        // register while the MFMAs are in flight. Each MFMA reads its A and B operands when
        // it issues, so a fragment can be reloaded once every MFMA that reads it has issued.
        // Centred rows: the splits are interleaved (split s sweeps the steps s, s + nsplit,
        // ... of the whole set), so that the rows of one group, which are stored together
        // and hold most of a query's nearest rows, are spread over the splits. The products
        // of this wave's queries with the centre of a step's group (a step lies in one group)
        // are loaded a step ahead, with the step's operands, when the group changes.
        const int first = CEN ? split * TSTEP : t_begin;
        const int last = CEN ? n : t_end;
        const int stride = CEN ? nsplit * TSTEP : TSTEP;
        if (first < last) {
            frag_t b[COLT][NS][NF];
            ae hv[COLT];
            load_step(first, b, hv);
            ae qcv[ROWT][4] = {};
            int cur_g = -1;
            auto load_qc = [&](int grp, ae (&dst)[ROWT][4]) {
#pragma unroll
                for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int q = q0 + qbase + rt * 16 + 4 * g + i;
                        dst[rt][i] = q < nq ? cen.qc[(int64_t)grp * nq + q] : (ae)0;
                    }
            };
            if constexpr (CEN) {
                cur_g = cen.step_group[first / TSTEP];
                load_qc(cur_g, qcv);
            }
            int off = 0; // buffered offset of the step's first row
            for (int tb = first; tb < last; tb += stride, off += TSTEP) {
                const int nb = tb + stride < last ? tb + stride : tb;
                const in_t *ntile = Y + (int64_t)nb * rowlen + lane * CSR;
                acc_t acc[ROWT][COLT];
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
#pragma unroll
                    for (int rt = 0; rt < ROWT; rt++) {
                        if constexpr (CEN)
                            acc[rt][ct] = acc_t{hv[ct] + qcv[rt][0], hv[ct] + qcv[rt][1],
                                                hv[ct] + qcv[rt][2], hv[ct] + qcv[rt][3]};
                        else
                            acc[rt][ct] = acc_t{hv[ct], hv[ct], hv[ct], hv[ct]};
                    }
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
                    hv[ct] = nyh[nb + ct * 16 + r];
                if constexpr (CEN) {
                    const int ng = cen.step_group[nb / TSTEP];
                    if (ng != cur_g) {
                        cur_g = ng;
                        load_qc(ng, qcv);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                constexpr int FPC = Op::FPC;
#pragma unroll
                for (int c = 0; c < NF / FPC; c++) {
#pragma unroll
                    for (int f = c * FPC; f < (c + 1) * FPC; f++)
#pragma unroll
                        for (int ct = 0; ct < COLT; ct++)
#pragma unroll
                            for (int rt = 0; rt < ROWT; rt++) {
                                acc[rt][ct] = Op::mfma(a[0][rt][f], b[ct][0][f], acc[rt][ct]);
                                if constexpr (SPLIT) {
                                    acc[rt][ct] =
                                        Op::mfma(a[0][rt][f], b[ct][NS - 1][f], acc[rt][ct]);
                                    acc[rt][ct] =
                                        Op::mfma(a[NS - 1][rt][f], b[ct][0][f], acc[rt][ct]);
                                }
                            }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int ct = 0; ct < COLT; ct++)
#pragma unroll
                        for (int h = 0; h < NS; h++)
                            Op::template load_chunk<CST, NF>(
                                ntile + (int64_t)ct * 16 * rowlen + h * 16 * Kp, c, b[ct][h]);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (any_hit(acc))
                    take_hits<Op, KC, CAP, ROWT, COLT>(acc, ts, cq, bkey, bidx, qbase, off);
            }
        }
    } else if constexpr (PF == 1) {
        // The next step's operands are loaded after this step's MFMAs and run under its
        // candidate bookkeeping. The hit test comes first: it reads the accumulators, so it
        // waits for the last MFMA, and only then may the loads and their address arithmetic
        // reuse registers. An MFMA still in flight can read its accumulator input after the
        // instructions that follow it have issued, so a register it reads must not be written
        // before its result has been consumed.
        if (t_begin < t_end) {
            frag_t b[COLT][NS][NF];
            ae hv[COLT];
            load_step(t_begin, b, hv);
            for (int tb = t_begin; tb < t_end; tb += TSTEP) {
                acc_t acc[ROWT][COLT];
                mfma_step(b, hv, acc);
                const bool hit = any_hit(acc);
                __builtin_amdgcn_sched_barrier(0);
                if (tb + TSTEP < t_end)
                    load_step(tb + TSTEP, b, hv);
                __builtin_amdgcn_sched_barrier(0);
                if (hit)
                    take_hits<Op, KC, CAP, ROWT, COLT>(acc, ts, cq, bkey, bidx, qbase,
                                                       tb - t_begin);
            }
        }
    } else {
        for (int tb = t_begin; tb < t_end; tb += TSTEP) {
            frag_t b[COLT][NS][NF];
            ae hv[COLT];
            load_step(tb, b, hv);
            // Keep every load of the step ahead of its first MFMA.
            __builtin_amdgcn_sched_barrier(0);
            acc_t acc[ROWT][COLT];
            mfma_step(b, hv, acc);
            check_step(tb, acc);
        }
    }
    write_split<Op, KC, CAP, ROWT>(cq, bkey, bidx, t_begin, CEN, qbase, q0, nq, nsplit, split,
                                   part_key, part_idx, part_tau, tau0, CEN ? cen.perm : nullptr);
}

// The sweep as a kernel: K2 asks the compiler for registers that leave room for WPE waves
// per SIMD (as many workgroups as fit a CU's LDS); the fp64 operands need more registers
// and run one wave per SIMD without spilling.
template <class Op, int KCH, int KC, bool SPLIT, int PF>
__global__ __launch_bounds__(64 * TileShape<KC>::WV) void knn_tile_kernel(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit,
    float *__restrict__ part_key, int *__restrict__ part_idx, float *__restrict__ part_tau,
    const float *__restrict__ tau0, const KnnCentring cen) {
    knn_tile_body<Op, KCH, KC, SPLIT, PF, 2>(Q, nq, Y, nyh, n, Kp, split_len, nsplit, part_key,
                                             part_idx, part_tau, tau0, cen);
}
template <class Op, int KCH, int KC, bool SPLIT, int PF, int COLT, bool CEN = false>
__global__ __launch_bounds__(64 * TileShape<KC>::WV)
__attribute__((amdgpu_waves_per_eu(TileCfg<Op, KCH, KC, SPLIT, COLT>::WPE))) void
knn_tile_kernel_k2(const typename Op::in_t *__restrict__ Q, int nq,
                   const typename Op::in_t *__restrict__ Y,
                   const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len,
                   int nsplit, float *__restrict__ part_key, int *__restrict__ part_idx,
                   float *__restrict__ part_tau, const float *__restrict__ tau0,
                   const KnnCentring cen) {
    knn_tile_body<Op, KCH, KC, SPLIT, PF, COLT, CEN>(Q, nq, Y, nyh, n, Kp, split_len, nsplit,
                                                     part_key, part_idx, part_tau, tau0, cen);
}

// Threshold seeds by group minima. The same operand layout and MFMA products as the sweep,
// over a small sample of training rows split like the sweep, but no candidate lists: lane
// (r, g) keeps, for each of its queries, the largest score over the sample rows whose offset
// in the split is r mod 16, so every query has 16 disjoint row groups and the best row of
// each. The k-th smallest of the 16 group-best keys is attained by k distinct rows, so it
// bounds the query's k-th key over the whole training set from above (k <= 16); written to
// part_tau[q * nsplit + split] for launch_knn_seed. Fixed shape: four waves of 32 queries,
// two 16-row training tiles per step, loads then products.
template <class Op, int KCH, bool SPLIT>
__global__ __launch_bounds__(256) void knn_seed_min_kernel(
    const typename Op::in_t *__restrict__ Q, int nq, const typename Op::in_t *__restrict__ Y,
    const typename Op::acc_e *__restrict__ nyh, int n, int Kp, int split_len, int nsplit, int k,
    float *__restrict__ part_tau) {
    using frag_t = typename Op::frag_t;
    using acc_t = typename Op::acc_t;
    using ae = typename Op::acc_e;
    using in_t = typename Op::in_t;
    constexpr int NF = KCH / Op::E;
    constexpr int NS = SPLIT ? 2 : 1;
    constexpr int CSR = 16 / (int)sizeof(in_t);
    constexpr int CST = 1024 / (int)sizeof(in_t);
    constexpr int ROWT = 2, COLT = 2, QB = 4 * ROWT * 16, TSTEP = COLT * 16;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int qblocks = (nq + QB - 1) / QB;
    const int work = (int)blockIdx.x;
    if (work >= qblocks * nsplit)
        return;
    const int split = work / qblocks;
    const int q0 = (work - split * qblocks) * QB;
    const int t_begin = split * split_len;
    const int t_end = t_begin + split_len < n ? t_begin + split_len : n;
    const int qbase = (w * ROWT) * 16;
    const int64_t rowlen = (int64_t)Kp * NS;
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
    ae best[ROWT][4];
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            best[rt][i] = (ae)-__builtin_huge_val();
    for (int tb = t_begin; tb < t_end; tb += TSTEP) {
        frag_t b[COLT][NS][NF];
        ae hv[COLT];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++) {
            hv[ct] = nyh[tb + ct * 16 + r];
            const in_t *tile = Y + (int64_t)(tb + ct * 16) * rowlen + lane * CSR;
#pragma unroll
            for (int h = 0; h < NS; h++)
                Op::template load<NF, CST>(tile + h * 16 * Kp, b[ct][h]);
        }
        __builtin_amdgcn_sched_barrier(0);
        acc_t acc[ROWT][COLT];
#pragma unroll
        for (int ct = 0; ct < COLT; ct++)
#pragma unroll
            for (int rt = 0; rt < ROWT; rt++)
                acc[rt][ct] = acc_t{hv[ct], hv[ct], hv[ct], hv[ct]};
#pragma unroll
        for (int f = 0; f < NF; f++)
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
#pragma unroll
        for (int rt = 0; rt < ROWT; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int ct = 0; ct < COLT; ct++)
                    best[rt][i] = acc[rt][ct][i] > best[rt][i] ? acc[rt][ct][i] : best[rt][i];
    }
    // Sort each 16-lane group's keys ascending (bitonic; the last merge ascending in every
    // group) and take the k-th.
#pragma unroll
    for (int rt = 0; rt < ROWT; rt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float key = (float)(-2.0 * (double)best[rt][i]);
            int idx = r;
            sort_merge<2, 1>(key, idx, lane);
            sort_merge<4, 2>(key, idx, lane);
            sort_merge<8, 4>(key, idx, lane);
            sort_merge<64, 8>(key, idx, lane);
            const float kth = __shfl(key, (lane & 48) + k - 1, 64);
            const int q = q0 + qbase + rt * 16 + (Op::rows_strided ? g + 4 * i : 4 * g + i);
            if (r == 0 && q < nq)
                part_tau[(int64_t)q * nsplit + split] = kth;
        }
    }
}

constexpr int MMAX = 2048; // candidates merged per query (nsplit * kc)
constexpr int RMAX = 256; // candidates re-ranked exactly per query

// Error bound of a stored key |y|^2 - 2 q.y of the fused tile. The tile accumulates the
// products of the query with the operand rows r (the rows themselves, or their residuals about
// a group centre c) from h = -|y|^2/2 (+ q.c for a centred form). With S = sum|q_l r_l| <=
// |q| rmax and |h| <= H = ymax^2/2 + |q| cmax: the operands rounded to the input precision
// (2u_in + u_in^2) S; the d+1-term accumulation from h (gamma_{d+2}); the rounding of h (of
// -|y|^2/2 alone, or of -|y|^2/2, q.c and their sum); the float storage of the key; the
// absolute error of operands below the normal range of fp16.
__device__ double key_error(double qnorm, int d, const KnnKeyError &e) {
    const double S = qnorm * e.rmax;
    const double H = 0.5 * e.ymax * e.ymax + qnorm * e.cmax;
    const double gam = (d + 2) * e.u_acc / (1.0 - (d + 2) * e.u_acc);
    const double u32 = 1.0 / 16777216.0;
    const double herr = e.cmax > 0.0 ? 2.0 * u32 * H : e.u_acc * 0.5 * e.ymax * e.ymax;
    return 2.0 * ((2.0 * e.u_in + e.u_in * e.u_in) * S +
                  gam * (S * (1.0 + e.u_in) * (1.0 + e.u_in) + H) + herr) +
           u32 * (2.0 * H + 2.01 * S) + 2.0 * e.abs_in * sqrt((double)d) * (qnorm + e.rmax) * 1.01;
}

// MM: candidate capacity of this instance (a power of two >= nsplit * kc).
template <typename T, int MM>
__global__ __launch_bounds__(64) void knn_merge_kernel(
    const float *__restrict__ part_key, const int *__restrict__ part_idx,
    const float *__restrict__ part_tau, int nsplit, int kc, const T *__restrict__ Q,
    const T *__restrict__ Yrows, int d, int k, double u_in, double abs_in, double u_acc,
    double ymax, double rmax, double cmax, int post,
    T *__restrict__ out_dist, int *__restrict__ out_idx, int *__restrict__ fb_count,
    int *__restrict__ fb_list, float *__restrict__ fb_bound) {
    __shared__ float skey[MM];
    __shared__ int sidx[MM];
    __shared__ double qrow[1024];
    __shared__ int scount;
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const int m = nsplit * kc; // <= MM
    const int64_t base = (int64_t)q * m;
    // The stored candidates (finite keys; unfilled list entries are +inf) packed to the front,
    // in any order: the sort below orders them by (key, index), and a row appears in one
    // list at most, so the result does not depend on the packing order. Only the next power
    // of two (at least 64 and at least k) is sorted.
    if (lane == 0)
        scount = 0;
    __syncthreads();
    for (int t = lane; t < m; t += 64) {
        const float kv = part_key[base + t];
        if (kv < __builtin_huge_valf()) {
            const int p = atomicAdd(&scount, 1);
            skey[p] = kv;
            sidx[p] = part_idx[base + t];
        }
    }
    __syncthreads();
    const int cnt = scount;
    int ms = 64;
    while (ms < cnt || ms < k)
        ms <<= 1;
    for (int t = cnt + lane; t < ms; t += 64) {
        skey[t] = __builtin_huge_valf();
        sidx[t] = INT_MAX;
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
    for (int size = 2; size <= ms; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < ms / 2; t += 64) {
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
    // Error bound of the stored keys.
    KnnKeyError ke;
    ke.u_in = u_in;
    ke.abs_in = abs_in;
    ke.u_acc = u_acc;
    ke.ymax = ymax;
    ke.rmax = rmax;
    ke.cmax = cmax;
    const double E = key_error(sqrt(qn), d, ke);
    const float ak = skey[k - 1];
    const double lim = (double)ak + 2.0 * E;
    // Candidates within 2E of the k-th key, re-ranked with exact distances.
    int nr = 0;
    while (nr < ms && (double)skey[nr] <= lim)
        nr++;
    const bool certified = (double)tau > lim && nr <= RMAX;
    if (!certified) {
        if (lane == 0) {
            const int p = atomicAdd(fb_count, 1);
            fb_list[p] = q;
            if (fb_bound) {
                const double b = (double)ak + 4.0 * E;
                float bf = (float)b;
                if ((double)bf < b)
                    bf = nextafterf(bf, __builtin_huge_valf());
                fb_bound[p] = bf;
            }
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
            if constexpr (sizeof(T) == 4) {
                if ((d & 3) == 0) {
                    // The same sums in the same order, the row read in 16-byte pieces.
                    const float4 *y4 = reinterpret_cast<const float4 *>(y);
                    for (int l4 = 0; l4 < d / 4; l4++) {
                        const float4 v = y4[l4];
                        double dv = qrow[4 * l4] - (double)v.x;
                        acc = fma(dv, dv, acc);
                        dv = qrow[4 * l4 + 1] - (double)v.y;
                        acc = fma(dv, dv, acc);
                        dv = qrow[4 * l4 + 2] - (double)v.z;
                        acc = fma(dv, dv, acc);
                        dv = qrow[4 * l4 + 3] - (double)v.w;
                        acc = fma(dv, dv, acc);
                    }
                } else {
                    for (int l = 0; l < d; l++) {
                        const double dv = qrow[l] - (double)y[l];
                        acc = fma(dv, dv, acc);
                    }
                }
            } else {
                for (int l = 0; l < d; l++) {
                    const double dv = qrow[l] - (double)y[l];
                    acc = fma(dv, dv, acc);
                }
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
    // Only the next power of two (at least 64 and at least k) holds re-ranked entries.
    int xs = 64;
    while (xs < nr || xs < k)
        xs <<= 1;
    for (int size = 2; size <= xs; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < xs / 2; t += 64) {
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

int knn_tile_min_splits(int prec, int Kp, int n, int kc, int variant) {
    const bool split = prec == KP_F16X3 || prec == KP_BF16X3;
    const bool wide = prec == KP_F16 || prec == KP_F16C || prec == KP_BF16 || split; // 8 elements
    const int frag_regs = wide ? 4 : prec == KP_F64 ? 2 : 1;
    const int nf = wide ? Kp / 4 / 8 : Kp / 4;
    const int rt = kc <= 16 ? TileShape<16>::RT : 1;
    const int colt = variant == 2 && prec != KP_F64 && prec != KP_F16C ? 1 : 2;
    if (!tile_short_index(kc, rt, split ? 2 : 1, nf, frag_regs, prec == KP_F64, colt))
        return 1;
    return (n + SPLIT16_MAX - 1) / SPLIT16_MAX;
}


// Launches the fused tile for one operand type, candidate-list length, width and variant
// (0: two 16-row training tiles per step, the next step's loads under the current step's
// bookkeeping; 1: two tiles per step, load then compute; 2: as 0 with one tile per step,
// which halves the training fragments held in registers; 3: two tiles per step with the
// rolling prefetch, each fragment reloaded for the next step as soon as its MFMAs have
// issued) or, with `slots` set, reports how many of its workgroups the device runs at once
// instead.
template <class Op, int KCH, int KC, bool SPLIT>
static hipError_t launch_tile_w(const void *Q, int nq, const void *Y, const void *nyh, int n,
                                int Kp, int nsplit, float *part_key, int *part_idx,
                                float *part_tau, const float *tau0, int variant, int *slots,
                                hipStream_t s, const KnnCentring *cen) {
    using in_t = typename Op::in_t;
    using ae = typename Op::acc_e;
    using kernel_t = void (*)(const in_t *, int, const in_t *, const ae *, int, int, int, int,
                              float *, int *, float *, const float *, const KnnCentring);
    constexpr int QB = TileShape<KC>::WV * TileShape<KC>::RT * 16;
    kernel_t kern;
    bool short_index = false;
    // Centred operand rows: the fp16 form with the rolling sweep.
    constexpr bool CEN_OK = std::is_same_v<Op, OpF16> && !SPLIT;
    if (cen && !CEN_OK)
        return hipErrorInvalidValue;
    if constexpr (CEN_OK) {
        if (cen) {
            kern = knn_tile_kernel_k2<Op, KCH, KC, SPLIT, 2, 2, true>;
            short_index = TileCfg<Op, KCH, KC, SPLIT, 2>::SHORT;
        }
    }
    if (cen) {
        // chosen above
    } else if constexpr (std::is_same_v<Op, OpF64>) {
        kern = variant == 1   ? knn_tile_kernel<Op, KCH, KC, SPLIT, 0>
               : variant == 3 ? knn_tile_kernel<Op, KCH, KC, SPLIT, 2>
                              : knn_tile_kernel<Op, KCH, KC, SPLIT, 1>;
    } else if (variant == 3) {
        kern = knn_tile_kernel_k2<Op, KCH, KC, SPLIT, 2, 2>;
        short_index = TileCfg<Op, KCH, KC, SPLIT, 2>::SHORT;
    } else if (variant == 2) {
        kern = knn_tile_kernel_k2<Op, KCH, KC, SPLIT, 1, 1>;
        short_index = TileCfg<Op, KCH, KC, SPLIT, 1>::SHORT;
    } else {
        kern = variant == 1 ? knn_tile_kernel_k2<Op, KCH, KC, SPLIT, 0, 2>
                            : knn_tile_kernel_k2<Op, KCH, KC, SPLIT, 1, 2>;
        short_index = TileCfg<Op, KCH, KC, SPLIT, 2>::SHORT;
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
    // Buffered row offsets must fit the index type.
    if (short_index && split_len > SPLIT16_MAX)
        return hipErrorInvalidValue;
    const int qblocks = (nq + QB - 1) / QB;
    const dim3 grid((qblocks * nsplit + 7) / 8 * 8);
    const KnnCentring none{};
    hipLaunchKernelGGL(kern, grid, block, 0, s, static_cast<const in_t *>(Q), nq,
                       static_cast<const in_t *>(Y), static_cast<const ae *>(nyh), n, Kp,
                       split_len, nsplit, part_key, part_idx, part_tau, tau0, cen ? *cen : none);
    return hipGetLastError();
}

template <class Op, int KC, bool SPLIT>
static hipError_t launch_tile_op(const void *Q, int nq, const void *Y, const void *nyh, int n,
                                 int Kp, int nsplit, float *part_key, int *part_idx,
                                 float *part_tau, const float *tau0, int variant, int *slots,
                                 hipStream_t s, const KnnCentring *cen) {
    switch (Kp / 4) {
    case 8:
        return launch_tile_w<Op, 8, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 16:
        return launch_tile_w<Op, 16, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 24:
        return launch_tile_w<Op, 24, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 32:
        return launch_tile_w<Op, 32, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
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
                                 float *part_tau, const float *tau0, int variant, int *slots,
                                 hipStream_t s, const KnnCentring *cen) {
#define DA_OP(OP, SPLIT)                                                                          \
    launch_tile_op<OP, KC, SPLIT>(Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau,     \
                                  tau0, variant, slots, s, cen)
    if ((prec == KP_F16C) != (cen != nullptr))
        return hipErrorInvalidValue;
    switch (prec) {
    case KP_F32: return DA_OP(OpF32, false);
    case KP_F64: return DA_OP(OpF64, false);
    case KP_F16:
    case KP_F16C: return DA_OP(OpF16, false);
    case KP_BF16: return DA_OP(OpBF16, false);
    case KP_F16X3: return DA_OP(OpF16, true);
    case KP_BF16X3: return DA_OP(OpBF16, true);
    default: return hipErrorInvalidValue;
    }
#undef DA_OP
}

static hipError_t tile_dispatch(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                                int n, int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                                float *part_tau, const float *tau0, int variant, int *slots,
                                hipStream_t s, const KnnCentring *cen) {
    switch (kc) {
    case 16:
        return launch_tile_kc<16>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 32:
        return launch_tile_kc<32>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 64:
        return launch_tile_kc<64>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    case 128:
        return launch_tile_kc<128>(prec, Q, nq, Y, nyh, n, Kp, nsplit, part_key, part_idx, part_tau, tau0, variant, slots, s, cen);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_knn_tile(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                           int n, int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                           float *part_tau, const float *tau0, int variant, hipStream_t s,
                           const KnnCentring *cen) {
    return tile_dispatch(prec, Q, nq, Y, nyh, n, Kp, nsplit, kc, part_key, part_idx, part_tau,
                         tau0, variant, nullptr, s, cen);
}

template <class Op, bool SPLIT>
static hipError_t launch_seed_op(const void *Q, int nq, const void *Y, const void *nyh, int n,
                                 int Kp, int nsplit, int k, float *part_tau, hipStream_t s) {
    using in_t = typename Op::in_t;
    using ae = typename Op::acc_e;
    using kernel_t = void (*)(const in_t *, int, const in_t *, const ae *, int, int, int, int,
                              int, float *);
    kernel_t kern;
    switch (Kp / 4) {
    case 8: kern = knn_seed_min_kernel<Op, 8, SPLIT>; break;
    case 16: kern = knn_seed_min_kernel<Op, 16, SPLIT>; break;
    case 24: kern = knn_seed_min_kernel<Op, 24, SPLIT>; break;
    case 32: kern = knn_seed_min_kernel<Op, 32, SPLIT>; break;
    default: return hipErrorInvalidValue;
    }
    int split_len = (n + nsplit - 1) / nsplit;
    split_len = (split_len + STEP - 1) / STEP * STEP;
    const int qblocks = (nq + 127) / 128;
    hipLaunchKernelGGL(kern, dim3(qblocks * nsplit), dim3(256), 0, s,
                       static_cast<const in_t *>(Q), nq, static_cast<const in_t *>(Y),
                       static_cast<const ae *>(nyh), n, Kp, split_len, nsplit, k, part_tau);
    return hipGetLastError();
}

hipError_t launch_knn_seed_min(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                               int n, int Kp, int nsplit, int k, float *part_tau,
                               hipStream_t s) {
    if (k < 1 || k > 16 || n <= 0 || n % STEP != 0 || nsplit < 1 || Kp % 32 != 0)
        return hipErrorInvalidValue;
    if (nq <= 0)
        return hipSuccess;
    switch (prec) {
    case KP_F32: return launch_seed_op<OpF32, false>(Q, nq, Y, nyh, n, Kp, nsplit, k, part_tau, s);
    case KP_F16: return launch_seed_op<OpF16, false>(Q, nq, Y, nyh, n, Kp, nsplit, k, part_tau, s);
    case KP_BF16: return launch_seed_op<OpBF16, false>(Q, nq, Y, nyh, n, Kp, nsplit, k, part_tau, s);
    case KP_F16X3: return launch_seed_op<OpF16, true>(Q, nq, Y, nyh, n, Kp, nsplit, k, part_tau, s);
    case KP_BF16X3: return launch_seed_op<OpBF16, true>(Q, nq, Y, nyh, n, Kp, nsplit, k, part_tau, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t knn_tile_slots(int prec, int Kp, int kc, int variant, int *slots) {
    // The centred form's kernel has the registers and LDS of the plain fp16 sweep.
    static const KnnCentring probe{};
    return tile_dispatch(prec, nullptr, 0, nullptr, nullptr, 0, Kp, 1, kc, nullptr, nullptr,
                         nullptr, nullptr, variant, slots, nullptr,
                         prec == KP_F16C ? &probe : nullptr);
}

namespace {
// One wave per stored row: the residual of its original row about its group centre, rounded
// to fp16 and zero-padded to Kp features, and the residual's squared norm (in double).
template <typename T>
__global__ __launch_bounds__(256) void knn_centre_rows_kernel(
    const T *__restrict__ X, int d, const int *__restrict__ perm,
    const int *__restrict__ step_group, const double *__restrict__ C, int rows_pad, int Kp,
    _Float16 *__restrict__ dst, double *__restrict__ rnorm2) {
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= rows_pad)
        return;
    const int j = perm[p];
    const int g = step_group[p / STEP];
    double acc = 0.0;
    for (int l = lane; l < Kp; l += 64) {
        double v = 0.0;
        if (j >= 0 && l < d)
            v = (double)X[(int64_t)j * d + l] - C[(int64_t)g * d + l];
        dst[(int64_t)p * Kp + l] = (_Float16)v;
        acc = fma(v, v, acc);
    }
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    if (lane == 0)
        rnorm2[p] = acc;
}

// tau0[q] = the smallest split threshold of a sweep over a subset of the rows: the subset
// holds kc rows with keys at most that value, so the kc-th key of the whole set is no larger.
// With Q, raised by the error margins of the two sweeps' keys (rounded up to float).
template <typename T>
__global__ void knn_seed_kernel(const float *__restrict__ part_tau, int nq, int nsplit,
                                const T *__restrict__ Q, int d, const KnnKeyError et,
                                const KnnKeyError es, float *__restrict__ tau0) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    float t = __builtin_huge_valf();
    for (int s = 0; s < nsplit; s++)
        t = fminf(t, part_tau[(int64_t)q * nsplit + s]);
    if (Q && t < __builtin_huge_valf()) {
        double qn = 0.0;
        for (int l = 0; l < d; l++) {
            const double v = (double)Q[(int64_t)q * d + l];
            qn = fma(v, v, qn);
        }
        const double qnorm = sqrt(qn);
        const double v = (double)t + 4.0 * key_error(qnorm, d, et) + 2.0 * key_error(qnorm, d, es);
        t = (float)v;
        if ((double)t < v)
            t = nextafterf(t, __builtin_huge_valf());
    }
    tau0[q] = t;
}
} // namespace

template <typename T>
hipError_t launch_knn_centre_rows(const T *X, int d, const int *perm, const int *step_group,
                                  const double *C, int rows_pad, int Kp, void *dst,
                                  double *rnorm2, hipStream_t s) {
    if (rows_pad % STEP != 0 || Kp < d)
        return hipErrorInvalidValue;
    if (rows_pad == 0)
        return hipSuccess;
    hipLaunchKernelGGL(knn_centre_rows_kernel<T>, dim3((rows_pad + 3) / 4), dim3(256), 0, s, X, d,
                       perm, step_group, C, rows_pad, Kp, static_cast<_Float16 *>(dst), rnorm2);
    return hipGetLastError();
}
template hipError_t launch_knn_centre_rows<float>(const float *, int, const int *, const int *,
                                                  const double *, int, int, void *, double *,
                                                  hipStream_t);
template hipError_t launch_knn_centre_rows<double>(const double *, int, const int *, const int *,
                                                   const double *, int, int, void *, double *,
                                                   hipStream_t);

// The products of every query with every group centre: the fp64 matrix-core inner-product
// tile (LDS-staged 64 x 64 blocks).
template <typename T>
hipError_t launch_knn_query_centre(const T *Q, int nq, int d, const double *C, int G, float *qc,
                                   hipStream_t s) {
    if (nq <= 0 || G <= 0)
        return hipSuccess;
    return launch_inner_mfma<double, T>(C, Q, G, nq, d, qc, nq, s);
}
template hipError_t launch_knn_query_centre<float>(const float *, int, int, const double *, int,
                                                   float *, hipStream_t);
template hipError_t launch_knn_query_centre<double>(const double *, int, int, const double *, int,
                                                    float *, hipStream_t);

template <typename T>
hipError_t launch_knn_seed(const float *part_tau, int nq, int nsplit, const T *Q, int d,
                           const KnnKeyError &tile_err, const KnnKeyError &seed_err,
                           float *tau0, hipStream_t s) {
    if (nq <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(knn_seed_kernel<T>, dim3((nq + 255) / 256), dim3(256), 0, s, part_tau, nq,
                       nsplit, Q, d, tile_err, seed_err, tau0);
    return hipGetLastError();
}
template hipError_t launch_knn_seed<float>(const float *, int, int, const float *, int,
                                           const KnnKeyError &, const KnnKeyError &, float *,
                                           hipStream_t);
template hipError_t launch_knn_seed<double>(const float *, int, int, const double *, int,
                                            const KnnKeyError &, const KnnKeyError &, float *,
                                            hipStream_t);

template <typename T>
hipError_t launch_knn_merge(const float *part_key, const int *part_idx, const float *part_tau,
                            int nsplit, int kc, const T *Q, const T *Y, int nq, int d, int k,
                            double u_in, double abs_in, double u_acc, double ymax, double rmax,
                            double cmax, int post, T *out_dist, int *out_idx, int *fb_count,
                            int *fb_list, float *fb_bound, hipStream_t s) {
    const int m = nsplit * kc;
    // The lists may be shorter than k; their union must hold k candidates.
    if (m > MMAX || d > 1024 || k > 128 || k > m)
        return hipErrorInvalidValue;
    if (nq == 0)
        return hipSuccess;
    auto kern = m <= 256 ? knn_merge_kernel<T, 256>
              : m <= 512 ? knn_merge_kernel<T, 512>
              : m <= 1024 ? knn_merge_kernel<T, 1024>
                          : knn_merge_kernel<T, 2048>;
    hipLaunchKernelGGL(kern, dim3(nq), dim3(64), 0, s, part_key, part_idx, part_tau, nsplit, kc,
                       Q, Y, d, k, u_in, abs_in, u_acc, ymax, rmax, cmax, post, out_dist, out_idx,
                       fb_count, fb_list, fb_bound);
    return hipGetLastError();
}

template hipError_t launch_knn_merge<float>(const float *, const int *, const float *, int, int,
                                            const float *, const float *, int, int, int, double,
                                            double, double, double, double, double, int, float *,
                                            int *, int *, int *, float *, hipStream_t);
template hipError_t launch_knn_merge<double>(const float *, const int *, const float *, int, int,
                                             const double *, const double *, int, int, int,
                                             double, double, double, double, double, double, int,
                                             double *, int *, int *, int *, float *, hipStream_t);

} // namespace dagpu
