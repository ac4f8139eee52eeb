/*
 * CPU baseline of the brute-force kNN: a restatement of AOCL-DA's host algorithm, timed
 * beside the GPU path. Test and measurement infrastructure only: used by bench.py's
 * cpu_baseline leg and by tests/ as a checked baseline; nothing in the product library
 * links or calls it.
 *
 * It follows the reference's structure step by step:
 *   - kneighbors_compute_brute_force (source/core/nearest_neighbors/nearest_neighbors.cpp:
 *     567-816): training blocks of 512 rows (float) or 256 (double) times query blocks of
 *     1,024; OpenMP over the (query block, training block) pairs (collapse(2),
 *     schedule(guided)); a per-thread distance tile, column-major with ldd = training
 *     block; per query a fill to k, then the replace-the-maximum scan
 *     (smaller_values_and_indices, nearest_neighbors_utils.hpp:44-68,187-240, with the
 *     first-maximum rule of inline_iamax :106-116); per-thread candidate lists merged
 *     into thread 0's with the same rule; then a stable sort by distance and the square
 *     root (sorted_n_dist_n_ind :72-101).
 *   - the distance tile is the packed direct-difference form
 *     (partitioning_infrastructure.hpp:45-387, metrics_kernels.hpp:413-558): the BLIS
 *     five-loop blocking with the block sizes MR/NR/MC/NC/KC of :45-58 (AVX-512: float
 *     16/8/256/1024/256, double 8/8/128/1024/512; AVX2: float 8/8/512/1024/64, double
 *     4/4/256/1024/32), packed zero-padded micro-panels, and an MR x NR register tile of
 *     (a - b)^2 accumulations with two accumulator sets (k-rows 0,1 and 2,3 of each step
 *     of four) merged at the end of a KC slice; the C tile is reloaded per KC slice.
 *     Inside the kNN the tile routine runs on one thread (it is called inside the parallel
 *     region), as in the reference.
 * The training matrix is column-major (the reference keeps a transposed copy of row-major
 * data, nearest_neighbors.cpp:469-514), and so are the queries.
 * The vector code is written with GCC vector extensions; built -O3 -march=native -fopenmp.
 */
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <numeric>
#include <omp.h>
#include <vector>

namespace {

template <typename T> struct Blocks;
#if defined(__AVX512F__)
template <> struct Blocks<float> { static constexpr int MR = 16, NR = 8, MC = 256, NC = 1024, KC = 256; };
template <> struct Blocks<double> { static constexpr int MR = 8, NR = 8, MC = 128, NC = 1024, KC = 512; };
#else
template <> struct Blocks<float> { static constexpr int MR = 8, NR = 8, MC = 512, NC = 1024, KC = 64; };
template <> struct Blocks<double> { static constexpr int MR = 4, NR = 4, MC = 256, NC = 1024, KC = 32; };
#endif

template <typename T, int W> struct Vec {
    typedef T type __attribute__((vector_size(W * sizeof(T))));
};

// Packs an m x k block (column-major, ld) into BLOCK-row micro-panels: per micro-panel,
// column after column of BLOCK contiguous entries, the rows past m zero.
template <typename T, int BLOCK>
void pack_block(int m, int k, const T *X, int64_t ld, T *Xt) {
    for (int i = 0; i < m; i += BLOCK) {
        const int ib = std::min(BLOCK, m - i);
        for (int p = 0; p < k; p++) {
            std::memcpy(Xt, X + i + p * ld, ib * sizeof(T));
            if (ib < BLOCK)
                std::memset(Xt + ib, 0, (BLOCK - ib) * sizeof(T));
            Xt += BLOCK;
        }
    }
}

// MR x NR tile of squared distances over k packed steps: C(i, j) (+)= sum_p (a_pi - b_pj)^2.
// Two accumulator sets take steps 0,1 and 2,3 of every four.
template <typename T, bool FIRST>
inline void micro_tile(int m, int n, int k, const T *At, const T *Bt, T *C, int64_t ldc) {
    constexpr int MR = Blocks<T>::MR, NR = Blocks<T>::NR;
    typedef typename Vec<T, MR>::type V;
    V g[NR], g2[NR];
    for (int j = 0; j < NR; j++) {
        g2[j] = V{};
        if (FIRST || j >= n || m < MR) {
            g[j] = V{};
            if (!FIRST && j < n)
                for (int i = 0; i < m; i++)
                    g[j][i] = C[i + j * ldc];
        } else {
            std::memcpy(&g[j], C + j * ldc, sizeof(V));
        }
    }
    int p = 0;
    for (; p + 4 <= k; p += 4) {
        V a0, a1, a2, a3;
        std::memcpy(&a0, At, sizeof(V));
        std::memcpy(&a1, At + MR, sizeof(V));
        std::memcpy(&a2, At + 2 * MR, sizeof(V));
        std::memcpy(&a3, At + 3 * MR, sizeof(V));
        for (int j = 0; j < NR; j++) {
            const V t0 = a0 - Bt[j], t1 = a1 - Bt[NR + j];
            g[j] += t0 * t0;
            g[j] += t1 * t1;
            const V t2 = a2 - Bt[2 * NR + j], t3 = a3 - Bt[3 * NR + j];
            g2[j] += t2 * t2;
            g2[j] += t3 * t3;
        }
        At += 4 * MR;
        Bt += 4 * NR;
    }
    for (int j = 0; j < NR; j++)
        g[j] += g2[j];
    for (; p < k; p++) {
        V a0;
        std::memcpy(&a0, At, sizeof(V));
        for (int j = 0; j < NR; j++) {
            const V t0 = a0 - Bt[j];
            g[j] += t0 * t0;
        }
        At += MR;
        Bt += NR;
    }
    for (int j = 0; j < n; j++) {
        if (m == MR)
            std::memcpy(C + j * ldc, &g[j], sizeof(V));
        else
            for (int i = 0; i < m; i++)
                C[i + j * ldc] = g[j][i];
    }
}

// The five loops (LoopFive..LoopOne_packed) for an m x n tile of squared distances between
// the columns-major A (m rows, lda) and B (n rows, ldb), single thread. C column-major, ldc.
template <typename T>
void packed_sqdist(int m, int n, int k, const T *A, int64_t lda, const T *B, int64_t ldb, T *C,
                   int64_t ldc, T *Atilde, T *Btilde) {
    constexpr int MR = Blocks<T>::MR, NR = Blocks<T>::NR, MC = Blocks<T>::MC,
                  NC = Blocks<T>::NC, KC = Blocks<T>::KC;
    for (int j = 0; j < n; j += NC) {
        const int jb = std::min(NC, n - j);
        for (int p = 0; p < k; p += KC) {
            const int pb = std::min(KC, k - p);
            pack_block<T, NR>(jb, pb, B + j + p * ldb, ldb, Btilde);
            for (int i = 0; i < m; i += MC) {
                const int ib = std::min(MC, m - i);
                pack_block<T, MR>(ib, pb, A + i + p * lda, lda, Atilde);
                for (int jj = 0; jj < jb; jj += NR) {
                    const int nb = std::min(NR, jb - jj);
                    for (int ii = 0; ii < ib; ii += MR) {
                        const int mb = std::min(MR, ib - ii);
                        T *Ct = C + (i + ii) + (int64_t)(j + jj) * ldc;
                        if (p == 0)
                            micro_tile<T, true>(mb, nb, pb, Atilde + ii * pb, Btilde + jj * pb,
                                                Ct, ldc);
                        else
                            micro_tile<T, false>(mb, nb, pb, Atilde + ii * pb, Btilde + jj * pb,
                                                 Ct, ldc);
                    }
                }
            }
        }
    }
}

// inline_iamax: the first index of the largest value (strict >).
template <typename T> inline int iamax(int k, const T *v) {
    int b = 0;
    for (int i = 1; i < k; i++)
        if (v[i] > v[b])
            b = i;
    return b;
}

// smaller_values_and_indices: a value <= the current maximum replaces it. Blocks of 16 are
// skipped when none of their values is <= the maximum (the vectorised pre-test).
template <typename T>
void replace_max(int n, const T *D, int k, int *kind, T *kdist, int offset) {
    int mi = iamax(k, kdist);
    T mv = kdist[mi];
    int i = 0;
    for (; i + 16 <= n; i += 16) {
        bool any = false;
        for (int l = 0; l < 16; l++)
            any |= D[i + l] <= mv;
        if (!any)
            continue;
        for (int l = 0; l < 16; l++) {
            if (D[i + l] <= mv) {
                kdist[mi] = D[i + l];
                kind[mi] = i + l + offset;
                mi = iamax(k, kdist);
                mv = kdist[mi];
            }
        }
    }
    for (; i < n; i++) {
        if (D[i] <= mv) {
            kdist[mi] = D[i];
            kind[mi] = i + offset;
            mi = iamax(k, kdist);
            mv = kdist[mi];
        }
    }
}

template <typename T>
int knn(int n, int d, const T *Xt, int nq, const T *Xq, int k, int nthreads, int *ind,
        T *dist) {
    const int tb_max = std::is_same<T, float>::value ? 512 : 256;
    const int qb_max = 1024;
    const int tb = std::min(tb_max, n), qb = std::min(qb_max, nq);
    const int ntb = (n + tb - 1) / tb, nqb = (nq + qb - 1) / qb;
    if (nthreads < 1)
        nthreads = 1;
    nthreads = std::min(nthreads, ntb * nqb);
    constexpr int MC = Blocks<T>::MC, NC = Blocks<T>::NC, KC = Blocks<T>::KC;
    std::vector<std::vector<int>> tind(nthreads), tcnt(nthreads);
    std::vector<std::vector<T>> tdist(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        const int me = omp_get_thread_num();
        std::vector<T> D((size_t)tb * qb);
        std::vector<T> At((size_t)MC * KC), Bt((size_t)NC * KC);
        tind[me].assign((size_t)nq * k, 0);
        tdist[me].assign((size_t)nq * k, (T)0);
        tcnt[me].assign(nq, 0);
        int *kind = tind[me].data();
        T *kdist = tdist[me].data();
        int *cnt = tcnt[me].data();
#pragma omp for collapse(2) schedule(guided) nowait
        for (int jb = 0; jb < nqb; jb++) {
            for (int ib = 0; ib < ntb; ib++) {
                const int q0 = jb * qb, t0 = ib * tb;
                const int qn = std::min(qb, nq - q0), tn = std::min(tb, n - t0);
                packed_sqdist<T>(tn, qn, d, Xt + t0, n, Xq + q0, nq, D.data(), tb, At.data(),
                                 Bt.data());
                for (int jj = 0; jj < qn; jj++) {
                    const int q = q0 + jj;
                    int *qi = kind + (size_t)q * k;
                    T *qd = kdist + (size_t)q * k;
                    int &c = cnt[q];
                    int ii = 0;
                    for (; ii < tn && c < k; ii++, c++) {
                        qi[c] = t0 + ii;
                        qd[c] = D[ii + (size_t)jj * tb];
                    }
                    if (ii < tn)
                        replace_max(tn - ii, D.data() + ii + (size_t)jj * tb, k, qi, qd, t0 + ii);
                }
            }
        }
#pragma omp barrier
        std::vector<int> perm(k);
#pragma omp for schedule(guided)
        for (int q = 0; q < nq; q++) {
            int *qi = tind[0].data() + (size_t)q * k;
            T *qd = tdist[0].data() + (size_t)q * k;
            int c = tcnt[0][q];
            for (int t = 1; t < nthreads; t++) {
                const int oc = tcnt[t][q];
                const int *oi = tind[t].data() + (size_t)q * k;
                const T *od = tdist[t].data() + (size_t)q * k;
                int o = 0;
                for (; o < oc && c < k; o++, c++) {
                    qi[c] = oi[o];
                    qd[c] = od[o];
                }
                if (o < oc) {
                    int mi = iamax(k, qd);
                    T mv = qd[mi];
                    for (; o < oc; o++)
                        if (od[o] <= mv) {
                            qi[mi] = oi[o];
                            qd[mi] = od[o];
                            mi = iamax(k, qd);
                            mv = qd[mi];
                        }
                }
            }
            std::iota(perm.begin(), perm.end(), 0);
            std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return qd[a] < qd[b]; });
            for (int j = 0; j < k; j++) {
                ind[(size_t)q * k + j] = qi[perm[j]];
                dist[(size_t)q * k + j] = std::sqrt(qd[perm[j]]);
            }
        }
    }
    return 0;
}

} // namespace

extern "C" {

/* Brute-force euclidean kNN: Xt n x d and Xq nq x d column-major (ld n and nq); ind and dist
 * nq x k row per query, ascending; nthreads OpenMP threads. */
int cpu_knn_f32(int n, int d, const float *Xt, int nq, const float *Xq, int k, int nthreads,
                int *ind, float *dist) {
    return knn<float>(n, d, Xt, nq, Xq, k, nthreads, ind, dist);
}
int cpu_knn_f64(int n, int d, const double *Xt, int nq, const double *Xq, int k, int nthreads,
                int *ind, double *dist) {
    return knn<double>(n, d, Xt, nq, Xq, k, nthreads, ind, dist);
}
/* The vector width the library was built for: 512 (AVX-512 block sizes) or 256 (AVX2). */
int cpu_knn_vector_bits(void) {
#if defined(__AVX512F__)
    return 512;
#else
    return 256;
#endif
}

}
