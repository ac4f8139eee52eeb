// Host-callable launchers of the HIP kernels. Every launcher enqueues on the given stream
// and returns the launch status; none of them synchronizes.
#pragma once
#include <cstdint>
#include <hip/hip_runtime_api.h>

namespace dagpu {

// Metric codes used by the distance kernels (values are internal).
enum KMetric : int {
    KM_SQEUCLIDEAN = 0, // sum (x-y)^2
    KM_EUCLIDEAN = 1,   // sqrt(sum (x-y)^2)
    KM_MANHATTAN = 2,   // sum |x-y|
    KM_MINKOWSKI = 3,   // (sum |x-y|^p)^(1/p)
    KM_COSINE = 4,      // 1 - x.y / (|x||y|), norms <= eps replaced by 1
    KM_INNER = 5,       // x.y
    KM_SQEUCLIDEAN_GEMM = 6, // max(0, |x|^2 + |y|^2 - 2 x.y)
    KM_EUCLIDEAN_GEMM = 7,   // sqrt of the above
    KM_MINKOWSKI_KEY = 8,    // sum |x-y|^p (monotone key of minkowski)
};

// Layout conversion: column-major (ld) -> row-major compact, and row-major (ld) -> compact.
template <typename T>
hipError_t launch_colmajor_to_rowmajor(const T *src, int64_t ld, int rows, int cols, T *dst,
                                       hipStream_t s);
// Row-major rows -> row-major with padded leading dimension ldd, zero padding on the right
// and zero rows up to rows_pad; output element type U.
template <typename T, typename U>
hipError_t launch_pack_rows(const T *src, int64_t lds, int rows, int cols, U *dst, int64_t ldd,
                            int rows_pad, hipStream_t s);

// launch_pack_rows into the operand type of a fused-tile precision code (KP_* below):
// double, float, _Float16 or __bf16.
template <typename T>
hipError_t launch_pack_rows_prec(const T *src, int64_t lds, int rows, int cols, void *dst,
                                 int64_t ldd, int rows_pad, int prec, hipStream_t s);

// Squared row norms in double: out[i] = sum_j x[i*ld+j]^2.
template <typename T>
hipError_t launch_row_sqnorms(const T *X, int64_t ld, int rows, int cols, double *out,
                              hipStream_t s);

// Cosine norms: sqrt of the squared norms, replaced by 1 when <= eps (in place).
hipError_t launch_cosine_norms(double *sq, int rows, double eps, hipStream_t s);

// Distance tile: D[i*ldd + j] = metric(X_i, Y_j) for i<m, j<n, X and Y row-major compact
// (ld = k), accumulated in double and stored as U. nx/ny: cosine norms or squared norms for
// the GEMM-form metrics (may be null otherwise). zero_diag forces D[i][j] = 0 where
// i + diag_offset == j.
template <typename T, typename U>
hipError_t launch_pairwise_direct(int metric, const T *X, const T *Y, int m, int n, int k,
                                  U *D, int64_t ldd, const double *nx, const double *ny,
                                  double p, int zero_diag, int diag_offset, hipStream_t s);

// GEMM-form distance tile on the fp64 matrix cores for the metrics that are functions of x.y
// (pairwise_mfma_metric): nx/ny are the squared row norms (cosine: the cosine norms, sqrt
// replaced by 1 when <= eps). Squared euclidean and euclidean elements whose GEMM-form
// rounding bound could exceed 5e-13 (double output) or 5e-7 (float output) of the value are
// recomputed with direct differences. Same layout and diagonal rule as launch_pairwise_direct.
bool pairwise_mfma_metric(int metric);
// D[i*ldd + j] = (float) X_i . Y_j for X (m x k, type TX) and Y (n x k, type TY), row-major
// compact: the same tile with the products and sums in double, rounded once to float.
template <typename TX, typename TY>
hipError_t launch_inner_mfma(const TX *X, const TY *Y, int m, int n, int k, float *D,
                             int64_t ldd, hipStream_t s);
template <typename T, typename U>
hipError_t launch_pairwise_mfma(int metric, const T *X, const T *Y, int m, int n, int k, U *D,
                                int64_t ldd, const double *nx, const double *ny, int zero_diag,
                                int diag_offset, hipStream_t s);

// Per-row top-k of a row-major key matrix (rows x n, leading dimension ld) by
// (key, column index), ascending (descending when `largest`). Writes the k best keys and
// indices per row, sorted, to out_key[row*k + j] / out_idx[row*k + j]. k <= 3840.
template <typename K>
hipError_t launch_row_topk(const K *keys, int64_t ld, int rows, int n, int k, int largest,
                           K *out_key, int *out_idx, hipStream_t s);

} // namespace dagpu

namespace dagpu {
// Converts selected keys into reported distances and scatters them into the output
// arrays: out_dist[q*ldo + j] (post: 0 key, 1 sqrt(max(key,0)), 2 key^(1/p)).
template <typename T>
hipError_t launch_knn_finalize(const double *keys, const int *idx, int rows, int k, int post,
                               double p, T *out_dist, int *out_idx, int64_t ldo,
                               hipStream_t s);
} // namespace dagpu

namespace dagpu {
// ---- k-means ----
// Per-feature maximum absolute value (as double), out[j] = max_i |X[i*d+j]|; out must be
// zeroed beforehand.
template <typename T>
hipError_t launch_col_absmax(const T *X, int n, int d, double *out, hipStream_t s);

// Exact assignment: labels[i] = argmin_c sum_l (X[i,l]-C[c,l])^2 in double, ties to the
// lowest c (sph: argmin of -X_i.C_c, the largest dot product). mind[i] receives the minimum
// key (may be null). When prev is non-null, *changed is incremented for every i with
// labels[i] != prev[i].
template <typename T>
hipError_t launch_kmeans_assign(const T *X, int n, int d, const double *C, int k, int *labels,
                                const int *prev, int *changed, double *mind, hipStream_t s,
                                int sph = 0);

// Fixed-point centroid sums, one pass over the clusters [c0, c0+kc) and the features
// [l0, l0+dl), kc * dl <= KMEANS_PASS_PAIRS: `grid` workgroups take the chunks of ppb points in
// turn, sort each chunk by label in LDS and sum every (cluster, feature) pair in registers,
// then write their totals to wpart[(b * kc * dl + p) * 2 + w] (w = 0: the value scaled by
// 2^e[l] and rounded to int64, w = 1: the remainder scaled by 2^(e[l]+lshift)) and their
// cluster sizes to wcnt[b * kc + c]. wpart needs grid * kc * dl * 2 words, wcnt grid * kc
// integers; both are overwritten, never accumulated into. With winv (n doubles) sample i
// contributes X_i * winv[i].
constexpr int KMEANS_PASS_PAIRS = 16384;
template <typename T>
hipError_t launch_kmeans_partial_sums(const T *X, int n, int d, const int *labels, int c0,
                                      int kc, int l0, int dl, const int *escale, int ppb,
                                      int grid, long long *wpart, int *wcnt, hipStream_t s,
                                      const double *winv = nullptr);
// C[(c0+c)*d + l] for l in [l0, l0+dl) = the sum of the nb workgroup partials of a pass / the
// cluster size (sph: the sum, to be normalised by the caller once every feature is summed),
// counts[c0+c] = the cluster size.
hipError_t launch_kmeans_reduce(const long long *wpart, const int *wcnt, int nb, int c0, int kc,
                                int l0, int dl, int d, const int *escale, double *C,
                                int *counts, hipStream_t s, int sph = 0);
// Fills n 64-bit words with zero (a kernel, so that it can be captured in a graph).
hipError_t launch_zero_words(long long *p, int64_t n, hipStream_t s);
// LDS bytes of one partial-sum workgroup for kc clusters and ppb points.
size_t kmeans_partial_lds(int kc, int ppb);
// out[0] = ||Cprev||_F^2, out[1] = ||Cprev - Cnew||_F^2 (single workgroup, fixed order).
hipError_t launch_kmeans_shift(const double *Cprev, const double *Cnew, int len, double *out,
                               hipStream_t s);
// Deterministic sum of v[0..n): partial sums per block into tmp[0..256), then out[0].
hipError_t launch_sum_double(const double *v, int n, double *tmp, double *out, hipStream_t s);
// Converts centres between double and T.
template <typename T>
hipError_t launch_convert(const T *src, int64_t n, double *dst, hipStream_t s);
template <typename T>
hipError_t launch_convert_back(const double *src, int64_t n, T *dst, hipStream_t s);
// Gathers rows idx[0..m) of X (n x d, type T) into C (m x d, double).
template <typename T>
hipError_t launch_gather_rows(const T *X, int d, const int *idx, int m, double *C,
                              hipStream_t s);

// k-means++ helpers. w[i] = squared distance of X_i to row `ci` of X (exact, double), with
// negative or non-finite values set to 0 and w[ci] = 0. sph: the cosine distance
// max(0, 1 - X_i.X_ci inv[i] inv[ci]) instead (inv null: rows taken as unit rows); the same
// choice for launch_candidate_cost and launch_min_update.
template <typename T>
hipError_t launch_dist_to_row(const T *X, int n, int d, int ci, double *w, hipStream_t s,
                              int sph = 0, const double *inv = nullptr);
// cp[i] = inclusive prefix sum of w[i]/sum (sum = *total), cp[n-1] = 1. tmp: >= 2*1024 doubles.
hipError_t launch_normalised_scan(const double *w, int n, const double *total, double *cp,
                                  double *tmp, hipStream_t s);
// out[t] = first i with cp[i] >= u[t] (lower_bound), t < nt.
hipError_t launch_lower_bound(const double *cp, int n, const double *u, int nt, int *out,
                              hipStream_t s);
// For each candidate row cand[t]: part[t*nb + b] = partial of sum_i max(0, min(w[i], d_i(t))).
template <typename T>
hipError_t launch_candidate_cost(const T *X, int n, int d, const double *w, const int *cand,
                                 int nt, double *part, int nb, hipStream_t s, int sph = 0,
                                 const double *inv = nullptr);
// w[i] = max(0, min(w[i], d_i(ci))), w[ci] = 0.
template <typename T>
hipError_t launch_min_update(const T *X, int n, int d, int ci, double *w, hipStream_t s,
                             int sph = 0, const double *inv = nullptr);
// Reduces part[t*nb + b] over b for each t into out[t] (fixed order).
hipError_t launch_reduce_rows(const double *part, int nt, int nb, double *out, hipStream_t s);
// Direct distances of rows of X to centres (double), sqrt applied: D[i*ldd + c]. sph: the
// cosine distance 1 - clamp(X_i.C_c inv[i], -1, 1) (inv null: 1 - clamp(X_i.C_c)).
template <typename T>
hipError_t launch_kmeans_transform(const T *X, int n, int d, const double *C, int k, T *D,
                                   int64_t ldd, hipStream_t s, int sph = 0,
                                   const double *inv = nullptr);
// Each row of C (k x d) scaled to unit length; a zero row is left as it is.
hipError_t launch_normalize_rows(double *C, int k, int d, hipStream_t s);
// inv[i] = 1/|X_i| in double, 0 for a zero row.
template <typename T>
hipError_t launch_row_inv_norms(const T *X, int n, int d, double *inv, hipStream_t s);
} // namespace dagpu

namespace dagpu {
// out[i] = ||X_i - C[labels[i]]||^2 in double (sph: 1 - X_i.C[labels[i]] inv[i], inv null:
// 1 - X_i.C[labels[i]]).
template <typename T>
hipError_t launch_label_sqdist(const T *X, int n, int d, const double *C, const int *labels,
                               double *out, hipStream_t s, int sph = 0,
                               const double *inv = nullptr);
} // namespace dagpu

namespace dagpu {
// Rounds doubles through a lower precision and back: mode 1 float, mode 2 half.
hipError_t launch_round_through(double *v, int64_t n, int mode, hipStream_t s);
// dst[i] = (T)(lower(src[i])) for a data matrix: mode 1 float, mode 2 half.
template <typename T>
hipError_t launch_round_data(const T *src, int64_t n, int mode, T *dst, hipStream_t s);
} // namespace dagpu

namespace dagpu {
// ---- fused kNN (fp32 MFMA) ----
// Rearranges packed operand rows (row_bytes each, ns halves) into the tiled training layout
// the fused tile reads: per 16-row tile, [half][16-byte chunk][lane]. rows_pad (a multiple
// of 32) rows are written; rows at or beyond `rows` are zero.
hipError_t launch_knn_tile_operand(const void *src, int rows, int row_bytes, int ns, void *dst,
                                   int rows_pad, hipStream_t s);
// Workgroups of the fused tile (operand form prec, padded width Kp, list length kc, variant)
// that the current device runs at once.
hipError_t knn_tile_slots(int prec, int Kp, int kc, int variant, int *slots);
// Candidates kept per query and split for a request of k neighbours (16, 32, 64 or 128).
int knn_tile_kc(int k);
// Queries per workgroup of the fused tile for a candidate-list length kc.
int knn_tile_qb(int kc);
// Largest nsplit * kc the merge accepts.
int knn_merge_max_candidates();
// Radius search on a rows x n tile of distance keys D (row-major, leading dimension ldd).
// Each row is cut into nseg = ceil(n / radius_segment_len(n)) segments of columns (a
// multiple of 64). launch_pairwise_direct_count writes the keys and, per row and 64-column
// tile, the number of keys <= thr other than column self0 + row (self0 >= 0);
// launch_radius_segsum sums those into segcnt[r * nseg + s]. The fill writes, from
// segoff[r * nseg + s] on, the indices j of the keys <= thr of segment s in ascending order
// (and their keys when key is not null); the sort orders
// each row's range [offs[r], offs[r+1]) by key, stably (ties stay in index order), with
// temporary storage of radius_sort_bytes bytes; the sqrt pass takes the square root of n
// values in place.
template <typename T, typename U>
hipError_t launch_pairwise_direct_count(int metric, const T *X, const T *Y, int m, int n, int k,
                                        U *D, int64_t ldd, const double *nx, const double *ny,
                                        double p, U thr, int self0, int *tilecnt,
                                        hipStream_t s);
hipError_t launch_radius_segsum(const int *tilecnt, int rows, int n, int *segcnt,
                                hipStream_t s);
template <typename U>
hipError_t launch_radius_fill(const U *D, int64_t ldd, int rows, int n, U thr, int self0,
                              const int *segoff, int *ind, U *key, hipStream_t s);
int radius_segment_len(int n);

// ---- k-NN classification and regression from neighbour lists ----
// Neighbours of query q: ind[b + t], t < cnt, with b = off[q], cnt = off[q+1] - off[q] (a
// radius search) or b = q*k, cnt = k when off is null; dist alongside (read only when
// weights = 1, i.e. 1/d weights). Classification writes proba (nq x ncls, row per query)
// from cls (class index of each training row) and, when pred is not null, the label
// classes[argmax] (fallback_label when the largest probability is 0). A query without
// neighbours lowers *first_empty to qbase + its index; outlier_mode 0 leaves its row zero, 1 and 2
// set proba[fill_idx] = 1 (none when fill_idx < 0). Regression writes the mean target (or
// the 1/d-weighted mean), and fallback for a query without neighbours.
template <typename T>
hipError_t launch_nn_classify(int nq, const int *off, int k, const int *ind, const T *dist,
                              int weights, const int *cls, const int *classes, int ncls,
                              int outlier_mode, int fill_idx, int fallback_label, T *proba,
                              int *pred, int qbase, int *first_empty, hipStream_t s);
template <typename T>
hipError_t launch_nn_regress(int nq, const int *off, int k, const int *ind, const T *dist,
                             int weights, const T *tgt, int outlier_mode, T fallback, T *out,
                             int qbase, int *first_empty, hipStream_t s);

// ---- IVF-Flat search ----
// Scans, for each (query q < nq, probe p < n_probe) pair, the inverted list probes[q*n_probe+p]
// (vectors X[perm[r]] for r in [list_off[c], list_off[c+1])) and writes the exact squared
// distances (ip: the negated dot products -q.x), rounded to T, and global ids to
// keys/idx[q*ld + seg[q*n_probe+p] + r - b].
template <typename T>
hipError_t launch_ivf_scan(const T *Q, int nq, int n_probe, int d, const T *X, const int *perm,
                           const int *list_off, const int *probes, const int *seg, int64_t ld,
                           T *keys, int *idx, hipStream_t s, int ip = 0);
// keys = +inf, idx = INT_MAX over n entries (candidate padding).
template <typename T> hipError_t launch_ivf_pad(T *keys, int *idx, int64_t n, hipStream_t s);
// Selected (key, column) pairs of n = rows x k -> (distance, index): the index is
// cand_idx[(t / k) * ld + column]; INT_MAX there becomes -1 with key +inf. Distance: post 0
// the key, 1 its square root, 2 1 + key (cosine), 3 -key (inner product). dist may be null.
template <typename T>
hipError_t launch_ivf_finish(const T *key, const int *sel, const int *cand_idx, int64_t ld, int k,
                             int64_t n, int post, T *dist, int *ind, hipStream_t s);
// n rows of d values scaled to unit length in place (1/|x| in double; zero rows unchanged).
template <typename T> hipError_t launch_ivf_normalize(T *X, int64_t n, int d, hipStream_t s);

// ---- DBSCAN over a neighbour graph in CSR form (offs: n + 1 64-bit offsets, ind) ----
// parent[i] = i.
hipError_t launch_dbscan_init(int *parent, int n, hipStream_t s);
// Joins the trees of every pair of core points that are neighbours (core[i] != 0), with
// lock-free union by smaller root: after the launch each connected set of core points forms
// one tree whose root is its smallest index.
hipError_t launch_dbscan_union(const long long *offs, const int *ind, const unsigned char *core,
                               int *parent, int n, hipStream_t s);
// For core points parent[i] becomes the root of i; is_root[i] = 1 for a core point that is
// its own root, 0 otherwise.
hipError_t launch_dbscan_flatten(const unsigned char *core, int *parent, int *is_root, int n,
                                 hipStream_t s);
// Exclusive prefix sum of n ints (hipCUB), with its temporary-storage query.
hipError_t dbscan_scan_bytes(int n, size_t *bytes);
hipError_t launch_dbscan_scan(void *temp, size_t bytes, const int *in, int *out, int n,
                              hipStream_t s);
// labels[i]: for a core point, the cluster number of its root (rid = scan of is_root); for a
// non-core point, the smallest cluster number among its core neighbours, or -1 (noise).
hipError_t launch_dbscan_label(const long long *offs, const int *ind, const unsigned char *core,
                               const int *parent, const int *rid, int *labels, int n,
                               hipStream_t s);
template <typename U> hipError_t radius_sort_bytes(int total, int rows, size_t *bytes);
template <typename U>
hipError_t launch_radius_sort(void *temp, size_t bytes, const U *key_in, U *key_out,
                              const int *ind_in, int *ind_out, int total, int rows,
                              const int *offs, hipStream_t s);
template <typename U> hipError_t launch_radius_sqrt(U *v, int64_t n, hipStream_t s);

// Fewest training splits the fused tile accepts for n rows at candidate-list length kc,
// operand precision prec and padded width Kp (sweeps that buffer 16-bit row offsets take
// splits of at most 65,536 rows).
int knn_tile_min_splits(int prec, int Kp, int n, int kc, int variant);
// Operand precision of the fused tile.
// KP_F16C: fp16 operands of training rows stored relative to the centre of their group
// (the queries are plain fp16); the tile adds each query's product with the group centre.
enum KnnPrec : int {
    KP_F32 = 0, KP_F64 = 1, KP_F16 = 2, KP_BF16 = 3, KP_F16X3 = 4, KP_BF16X3 = 5, KP_F16C = 6
};
// Bytes per padded operand row element group: split forms store [hi Kp][lo Kp] per row.
inline int knn_prec_row_factor(int prec) {
    return prec == KP_F16X3 || prec == KP_BF16X3 ? 2 : 1;
}
// Centred training rows of the KP_F16C form. The rows are stored in group order, each group
// padded to a whole number of 32-row steps (padded rows have -|y|^2/2 = -inf); step_group[t]
// is the group of rows 32t .. 32t+31, qc[g*nq + q] the product of query q with the centre of
// group g (float), perm[p] the original index of stored row p.
struct KnnCentring {
    const int *step_group = nullptr;
    const float *qc = nullptr;
    int G = 0;
    const int *perm = nullptr;
};
inline int knn_prec_elem_size(int prec) { return prec == KP_F64 ? 8 : prec == KP_F32 ? 4 : 2; }
// Padded feature count used by the fused tile for d features (a multiple of 32, <= 128).
int knn_tile_kp(int d);
// Q: nq x Kp row-major zero-padded rows in the operand type of prec (float, double,
// _Float16 or __bf16; split forms [hi Kp][lo Kp]); Y: the n training rows in the tiled
// layout of launch_knn_tile_operand, padded to a multiple of 32 rows; nyh: -|y|^2/2 per row
// (double for KP_F64, float otherwise).
// Writes part_key/part_idx[(q*nsplit + s)*kc + i] (keys |y|^2 - 2 q.y as float) and
// part_tau[q*nsplit + s]. kc in {16, 32, 64, 128}. variant 0: two 16-row training tiles
// per step, the next step's loads issued under the current step's candidate bookkeeping;
// 1: two tiles per step, each step loads, then computes; 2: as 0 with one tile per step.
// tau0 (may be null): per-query key bound that seeds the thresholds; rows with keys at or
// above it are not buffered and part_tau is at most tau0[q].
// nyh must hold a multiple of 32 entries (the padded rows at -inf).
hipError_t launch_knn_tile(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                           int n, int Kp, int nsplit, int kc, float *part_key, int *part_idx,
                           float *part_tau, const float *tau0, int variant, hipStream_t s,
                           const KnnCentring *cen = nullptr);
// Centred operand rows of the KP_F16C form: for stored row p with original row j = perm[p]
// (-1: a padded row) of group g = step_group[p / 32], dst row p = fp16(X_j - C_g) zero-padded
// to Kp features (row-major, rows_pad rows) and rnorm2[p] = |X_j - C_g|^2 (0 for padded rows).
template <typename T>
hipError_t launch_knn_centre_rows(const T *X, int d, const int *perm, const int *step_group,
                                  const double *C, int rows_pad, int Kp, void *dst,
                                  double *rnorm2, hipStream_t s);
// qc[g*nq + q] = (float) sum_l Q[q*d + l] C[g*d + l], accumulated in double (group-major, so
// that the queries of one sweep read one cache line per group).
template <typename T>
hipError_t launch_knn_query_centre(const T *Q, int nq, int d, const double *C, int G, float *qc,
                                   hipStream_t s);
// Error model of a stored key of the fused tile: relative and absolute representation
// error of the operands, the accumulation's unit roundoff, the largest training-row norm,
// the largest operand-row norm (the residual norm for the centred form) and the largest
// centre norm (0 when the rows are not centred). launch_knn_merge takes the same six values.
struct KnnKeyError {
    double u_in = 0, abs_in = 0, u_acc = 0, ymax = 0, rmax = 0, cmax = 0;
};
// tau0[q] = min over the nsplit thresholds part_tau[q*nsplit + s] of a sweep (a valid seed
// for a sweep of any superset of the swept rows). With Q (nq x d rows of type T), the seed is
// raised by 4 E_tile(|q|) + 2 E_seed(|q|), the error bounds of a key of the seeded sweep and
// of the sweep that made the thresholds, so that a seeded sweep's certification (threshold
// above the k-th stored key + 2 E_tile) does not fail for the seed alone.
template <typename T>
hipError_t launch_knn_seed(const float *part_tau, int nq, int nsplit, const T *Q, int d,
                           const KnnKeyError &tile_err, const KnnKeyError &seed_err,
                           float *tau0, hipStream_t s);
// Group-minimum seeds: Q and Y (n sample rows, n a multiple of 32, in the tile layout) and
// nyh as for launch_knn_tile in the forms KP_F32, KP_F16, KP_BF16, KP_F16X3, KP_BF16X3; the
// sample is cut into nsplit splits and each split's rows into 16 groups by offset mod 16.
// part_tau[q*nsplit + s] = the k-th smallest (1 <= k <= 16) of split s's 16 group-best keys
// of query q: an upper bound on the query's k-th key over any superset of the sample.
hipError_t launch_knn_seed_min(int prec, const void *Q, int nq, const void *Y, const void *nyh,
                               int n, int Kp, int nsplit, int k, float *part_tau,
                               hipStream_t s);
// Q: nq x d and Y: n x d unpadded rows (type T); u_in: relative representation error of the
// tile operands (0 when they are the data themselves); abs_in: absolute representation error
// per element (fp16 subnormals); u_acc: unit roundoff of the tile accumulation;
// ymax: max row norm of Y; rmax: max norm of the operand rows of the tile (ymax, or the
// largest residual of a centred form); cmax: max centre norm of a centred form (0 otherwise).
// Requires nsplit * kc <= knn_merge_max_candidates() and k <= kc.
// Certified queries get out_dist/out_idx (row-major nq x k); the others are appended to
// fb_list (fb_count must be zeroed before), and, with fb_bound, fb_bound[p] for list entry p
// receives an upper bound on that query's k-th key plus 3 E (the k-th stored key + 4 E,
// rounded up; +inf when fewer than k candidates were stored): a threshold seed for a second
// sweep of the same operand form.
template <typename T>
hipError_t launch_knn_merge(const float *part_key, const int *part_idx, const float *part_tau,
                            int nsplit, int kc, const T *Q, const T *Y, int nq, int d, int k,
                            double u_in, double abs_in, double u_acc, double ymax, double rmax,
                            double cmax, int post, T *out_dist, int *out_idx, int *fb_count,
                            int *fb_list, float *fb_bound, hipStream_t s);
} // namespace dagpu

namespace dagpu {
// ---- MFMA k-means assignment (fp32 or fp64 data, bf16 cross term) ----
// CP[c*Kp + l] = (E)C[c*d + l] (zero padded), chalf[c] = -||C_c||^2/2, *cmax2 =
// max ||C_c||^2 (must be zeroed before).
// sph: chalf = 0 (spherical k-means: the argmax of x.c, the reference's zero centre norms).
template <typename E>
hipError_t launch_centre_prep(const double *C, int k, int d, int Kp, E *CP, E *chalf,
                              double *cmax2, hipStream_t s, int sph = 0);
// X: n x Kp zero-padded rows of fp32 (fp32 MFMA) or fp64 (fp64 MFMA), Kp a multiple of 16
// up to 128; CP, chalf: launch_centre_prep in the same type. Any k: centres that do not fit
// a workgroup's LDS are staged in chunks. Writes labels; samples whose best two centres are
// within the error bound of the MFMA arithmetic are appended to list (count in *list_count,
// zeroed before) for launch_kmeans_recheck; the others update *changed. wg_per_cu: 0
// launches one workgroup per 256 samples (128 for fp64 rows wider than 64); n > 0 at most n
// workgroups per CU, each taking every gridDim.x-th block of samples.
template <typename E>
hipError_t launch_kmeans_assign_mfma(const E *X, int n, int Kp, const E *CP, const E *chalf,
                                     int k, double cmax, int *labels, const int *prev,
                                     int *changed, int *list_count, int *list, hipStream_t s,
                                     int wg_per_cu = 0);
// bf16 cross-term form of the same assignment: X and C16 are bf16 rows padded to Kp (a
// multiple of 32, at most 128) features; C16 has (k+15)/16*16 rows.
size_t kmeans_assign_bf16_lds(int k, int Kp);
hipError_t launch_kmeans_assign_bf16(const void *X, int n, int Kp, const void *C16,
                                     const float *chalf, int k, double cmax, int *labels,
                                     const int *prev, int *changed, int *list_count, int *list,
                                     hipStream_t s, int wg_per_cu = 0);
hipError_t launch_centre_prep_bf16(const double *C, int k, int d, int Kp, void *C16,
                                   float *chalf, double *cmax2, hipStream_t s, int sph = 0);
hipError_t launch_rows_to_bf16(const float *X, int n, int d, int ldx, void *Y, int Kp,
                               hipStream_t s);
// Exact double re-check of the listed samples (sph: argmax of x.c); CT: k*d doubles of
// scratch for the transposed centres.
template <typename T>
hipError_t launch_kmeans_recheck(const T *X, int ldx, int d, const double *C, int k,
                                 const int *list_count, const int *list, int *labels,
                                 const int *prev, int *changed, double *CT, hipStream_t s,
                                 int sph = 0);
} // namespace dagpu

namespace dagpu {
// Row gather/scatter for the kNN fallback: dst[i*ld + c] = src[idx[i]*ld + c] (gather) and
// dst[idx[i]*ld + c] = src[i*ld + c] (scatter), c < cols, i < m. elem: 4 or 8 bytes.
hipError_t launch_gather_scatter(const void *src, void *dst, const int *idx, int m, int cols,
                                 int elem, int scatter, hipStream_t s);
// ---- Batched k-means (kmeans_batch.hip) ----
// Instance states.
enum { KB_ACTIVE = 0, KB_CONV_LABELS = 1, KB_CONV_SHIFT = 2, KB_MAXIT = 3 };
// Per-instance device arrays: sizes, padded feature count of the operand copy (bf16: 32 or
// 64; fp32: 8, 16, 32 or 64) and element offsets of the instance's rows (X), operand rows
// (Xo), centres (k x d), operand centres (kpad x kp) and labels.
struct KBatchDesc {
    const int *n, *d, *k, *kp;
    const int64_t *xoff, *xboff, *coff, *cboff, *loff;
};
// An assignment tile or re-check entry (instance, sample) and a sample chunk.
struct KTile {
    int inst, start;
};
struct KChunk {
    int inst, start, count, pad;
};
// Device work buffers of a batch.
struct KBatchBuffers {
    const float *X;
    int prec;                         // cross term: 0 fp32, 1 bf16, 2 bf16 hi + lo
    void *Xo;                         // rows in the operand type, padded to kp (x2 for hi + lo)
    double *C0, *C1;                  // centres by parity
    int *L0, *L1;                     // labels by parity
    void *Cb;                         // centres in the operand type (kpad x kp, x2 for hi + lo)
    float *chalf;                     // -|c|^2/2, 64 per instance
    double *cmax2;                    // max |c|^2 per instance
    const KTile *tiles;               // assignment tiles
    int ntiles;
    const KChunk *chunks;             // sample chunks
    int nchunks;
    const int64_t *poff;              // partial-sum offset of each chunk
    const int *cbeg, *cend;           // chunk range of each instance
    int *list_count, *changed;
    KTile *list;                      // re-check list (instance, sample)
    int *escale;                      // 65 per instance
    long long *part;
    int *pcnt;                        // 64 per chunk
    int *state, *fpar, *n_iter;
    double *cin;                      // inertia per chunk
};
hipError_t kb_launch_prepare(const float *X, const float *Cinit, const KBatchDesc &b, int n_inst,
                             const KChunk *chunks, int nchunks, int64_t ctotal,
                             unsigned long long *amax, int *escale, int prec, void *Xo,
                             double *C0, int *state, int *fpar, int *n_iter, hipStream_t s);
// One assignment pass (mode 0: iteration with centre parity par; mode 1: final relabel) with
// pt 16-sample tiles per wave on `grid` persistent workgroups, plus the exact re-check.
hipError_t kb_launch_assign(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int mode,
                            int par, int pt, int grid, hipStream_t s);
// Centroid update, shift and convergence test of iteration `iter` (parity par).
hipError_t kb_launch_update(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int par,
                            int iter, double tol, size_t lds, hipStream_t s);
// Final relabel, inertia and outputs.
hipError_t kb_launch_finish(const KBatchBuffers &w, const KBatchDesc &b, int n_inst, int pt,
                            int grid, float *Cout, int *labels, double *inertia, hipStream_t s);
int kb_assign_blocks_per_cu(int prec, int pt);
// Allows the partial-sum kernel `lds` bytes of dynamic LDS (up to 160 KB).
hipError_t kb_set_partial_lds(size_t lds);

} // namespace dagpu

namespace dagpu {
// ---- Elkan's k-means bounds (kmeans_elkan.hip) ----
// cc[i*k+j] = |C_i - C_j| / 2, nxt[i] = min_{j != i} cc[i*k+j].
hipError_t launch_elkan_half(const double *C, int k, int d, double *cc, double *nxt,
                             hipStream_t s);
// Initial labels and bounds: lab (n), u (n), lb (n x k).
template <typename T>
hipError_t launch_elkan_init(const T *X, int n, int d, const double *C, int k, const double *cc,
                             int *lab, double *u, double *lb, hipStream_t s);
// One bounded assignment from the labels prev; counts changed labels into *changed.
template <typename T>
hipError_t launch_elkan_assign(const T *X, int n, int d, const double *C, int k, const double *cc,
                               const double *nxt, const int *prev, int *lab, double *u,
                               double *lb, int *changed, hipStream_t s);
// Centre shifts (shift, k doubles), then u += shift[lab], lb = max(lb - shift, 0).
hipError_t launch_elkan_bounds(const double *Cprev, const double *Cnew, int n, int k, int d,
                               const int *lab, double *shift, double *u, double *lb,
                               hipStream_t s);
// u = DBL_MAX, lb = 0.
hipError_t launch_elkan_reset(int n, int k, double *u, double *lb, hipStream_t s);
// ---- AFK-MC2 seeding ----
// Proposal q[i] = w[i]/(2 sum) + 1/(2n), or 1/n for a zero or non-finite sum.
hipError_t launch_afk_proposal(const double *w, int n, const double *sum, double *q,
                               hipStream_t s);
// One Markov chain of m steps over the canonical draws u (at least 2m + 1): out[0] = the
// chosen sample, out[1] = draws consumed.
hipError_t launch_afk_chain(const double *cp, const double *q, const double *dmin, int n, int m,
                            const double *u, int *out, hipStream_t s);
} // namespace dagpu
