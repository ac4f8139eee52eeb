#include "knn.hpp"
#include "../core/factories.hpp"
#include "../kernels/launch.hpp"
#include "../kmeans/kmeans.hpp"
#include "aoclda_metrics.h"
#include "aoclda_nearest_neighbors.h"
#include <algorithm>
#include <map>
#include <cmath>
#include <vector>
#include <limits>
#include <numeric>
#include <random>

namespace dagpu {

// Internal algorithm codes, nearest_neighbors_types.hpp:29 of the reference.
enum { NN_BRUTE = 0, NN_KD_TREE = 1, NN_BALL_TREE = 2, NN_AUTO = 3 };

template <typename T> NNHandle<T>::NNHandle(ErrorStack &e, GpuContext &g) : AlgHandle(e, g) {
    // Options of source/core/nearest_neighbors/nearest_neighbors_options.hpp:52-124.
    const da_int imax = std::numeric_limits<da_int>::max();
    const da_int imin = std::numeric_limits<da_int>::min();
    const double fmax = (double)std::numeric_limits<T>::max();
    const double fmin = (double)std::numeric_limits<T>::lowest();
    opts.add_int("number of neighbors", "Number of neighbors considered for k-nearest neighbors.",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 5);
    opts.add_int("leaf size", "Leaf size for k-d tree.", 1, LowerBound::greater_equal, imax,
                 UpperBound::p_inf, 30);
    opts.add_int("outlier label",
                 "Classification label for queries with no neighbors within the specified radius.",
                 imin, LowerBound::m_inf, imax, UpperBound::p_inf, 0);
    opts.add_real("minkowski parameter",
                  "Minkowski parameter for metric used for the computation of nearest neighbors.",
                  0.0, LowerBound::greater_than, fmax, UpperBound::p_inf, 2.0);
    opts.add_real("radius", "Maximum distance for the radius neighbors computation.", 0.0,
                  LowerBound::greater_equal, fmax, UpperBound::p_inf, 1.0);
    opts.add_real("outlier target",
                  "Regression target for queries with no neighbors within the specified radius.",
                  fmin, LowerBound::m_inf, fmax, UpperBound::p_inf, 0.0);
    opts.add_string("algorithm", "Algorithm used to compute the k-nearest neighbors.",
                    {{"auto", NN_AUTO}, {"brute", NN_BRUTE}, {"kd tree", NN_KD_TREE},
                     {"ball tree", NN_BALL_TREE}},
                    "auto");
    opts.add_string("metric", "Metric used to compute the pairwise distance matrix.",
                    {{"euclidean", da_euclidean},
                     {"l2", da_l2},
                     {"sqeuclidean", da_sqeuclidean},
                     {"manhattan", da_manhattan},
                     {"l1", da_l1},
                     {"cityblock", da_cityblock},
                     {"cosine", da_cosine},
                     {"minkowski", da_minkowski},
                     {"sqeuclidean_gemm", da_sqeuclidean_gemm},
                     {"euclidean_gemm", da_euclidean_gemm},
                     {"inner product", da_inner_product}},
                    "euclidean");
    opts.add_string("weights", "Weight function used to compute the k-nearest neighbors.",
                    {{"uniform", 0}, {"distance", 1}}, "uniform");
    opts.add_string("outlier handling",
                    "How to handle samples with no neighbors within the specified radius.",
                    {{"none", 0}, {"manual", 1}, {"most frequent", 2}}, "none");
    // MI355X extension: operand precision of the distance tiles of the squared-euclidean
    // search. The candidates are always re-ranked with exact double-precision distances and
    // the result is certified against the error bound of the chosen precision.
    opts.add_string("cross term precision",
                    "Operand precision of the MFMA distance tiles: working, fp32, fp16, bf16, "
                    "fp16x3 or bf16x3 (hi+lo split operands, three products), fp16c (fp16 "
                    "training rows stored about the centres of k-means groups).",
                    {{"working", 0}, {"fp32", 1}, {"fp16", 2}, {"bf16", 3}, {"fp16x3", 4},
                     {"bf16x3", 5}, {"fp16c", 6}},
                    "working");
    // MI355X extension: kernel variant of the fused tile (-1 = chosen by the tuning table,
    // 0 = two 16-row training tiles per step with the next step's loads under the current
    // step's bookkeeping, 1 = two tiles per step, load then compute, 2 = as 0 with one tile
    // per step, 3 = two tiles per step, each training fragment reloaded for the next step as
    // soon as the MFMAs that read it have issued).
    opts.add_int("tile variant", "Kernel variant of the fused kNN distance tile (-1 automatic).",
                 -1, LowerBound::greater_equal, 3, UpperBound::less_equal, -1);
    // MI355X extension: length of the candidate list each split of the fused tile keeps per
    // query (rounded up to 16, 32, 64 or 128; 0 = automatic). A longer list
    // lowers the split thresholds, so a low-precision form with a wider error bound still
    // certifies its queries instead of sending them to the exact fallback; a list shorter
    // than k, spread over enough splits, makes the sweep's bookkeeping cheaper.
    opts.add_int("candidates per split",
                 "Candidates kept per query and split by the fused kNN tile (0 = from k).", 0,
                 LowerBound::greater_equal, 128, UpperBound::less_equal, 0);
    // MI355X extension: for k <= 16 on training sets of 131,072 rows or more, the minima of
    // 16 row groups of an 8,192-row sample bound each query's k-th key before the sweep, so
    // that the sweep buffers far fewer rows. Speed only: the result is exact either way.
    // "auto" seeds the 16-bit operand forms; for the fp32 and fp64 forms the sample's
    // products cost more than the sweep saves (200,000 x 128, 20,000 queries: fp32 seed
    // 0.65 ms against 0.4 ms saved; fp16c 0.09 ms against about 1 ms).
    opts.add_string("seed thresholds",
                    "Seed the fused kNN tile's thresholds from group minima over a row sample.",
                    {{"auto", 2}, {"yes", 1}, {"no", 0}}, "auto");
}

// The candidate-list length: from k, or as set by "candidates per split". A list shorter
// than k is exact as well: every row a split leaves out has a key at or above that split's
// threshold, so the merge certifies a query only when all k nearest rows were kept.
// By default a split keeps 32 candidates for 32 < k <= 128 (max(32, ceil(k / 4)), rounded
// up to a list length). Against 64-long lists, measured with bf16x3 at 200,000 x 128
// (20,000 queries) and 2,000,000 x 96 (2,048 queries): k = 50 1.2x faster at both; k = 100
// 1.27x faster at 2,000,000 x 96, equal at 200,000 x 128 (10 of 20,000 queries then take the
// exact fallback); k = 128 slightly faster (91 fallbacks). 64-long lists had in turn been
// 1.7-2.3x faster than 128-long ones at k = 100. At k <= 32 a 16-long list leaves too many
// queries uncertified (k = 30: 455 of 20,000), so the list holds k.
int SearchParams::tile_kc(int k) const {
    return knn_tile_kc(kc_min > 0 ? kc_min : (k > 32 ? std::max(32, (k + 3) / 4) : k));
}

// Tuning table of the fused tile variant per operand precision and padded width: the
// rolling prefetch (variant 3) for every form. Measured at 200,000 x 128 with 20,000
// queries, k = 10: fp32 11.7 -> 10.2 ms, bf16x3 3.64 -> 3.48 ms per launch against the
// pipelined sweep (variant 0).
static int knn_tile_auto_variant(int prec, int kp) {
    (void)prec;
    (void)kp;
    return 3;
}

template <typename T> int NNHandle<T>::tile_prec() const {
    switch (opts.skey("cross term precision")) {
    case 1:
        return KP_F32;
    case 2:
        return KP_F16;
    case 3:
        return KP_BF16;
    case 4:
        return KP_F16X3;
    case 5:
        return KP_BF16X3;
    case 6:
        return KP_F16C;
    default:
        return std::is_same_v<T, double> ? KP_F64 : KP_F32;
    }
}

template <typename T> da_status NNHandle<T>::set_params() {
    n_neighbors_ = opts.ival("number of neighbors");
    algo_ = opts.skey("algorithm");
    metric_ = opts.skey("metric");
    p_ = opts.rval("minkowski parameter");
    leaf_size_ = opts.ival("leaf size");
    // Tree algorithms are incompatible with these metrics (nearest_neighbors.cpp:313-329).
    if (algo_ == NN_KD_TREE || algo_ == NN_BALL_TREE) {
        if (metric_ == da_cosine || metric_ == da_sqeuclidean || metric_ == da_sqeuclidean_gemm ||
            metric_ == da_inner_product)
            return err->error(da_status_incompatible_options,
                              "The tree algorithms are not compatible with the cosine, "
                              "squared Euclidean, or inner product distances.");
        if (metric_ == da_minkowski && p_ < 1.0)
            return err->error(da_status_incompatible_options,
                              "Tree algorithms are not compatible with the Minkowski metric "
                              "when 0 < p < 1.");
    }
    return da_status_success;
}

// nearest_neighbors.cpp:394-466: options that shaped set_data may not change afterwards.
template <typename T> da_status NNHandle<T>::check_options_update() {
    const da_int a = opts.skey("algorithm");
    if (a == NN_AUTO || a == NN_KD_TREE || a == NN_BALL_TREE) {
        if (a != algo_ || opts.ival("leaf size") != leaf_size_)
            return err->error(da_status_option_locked,
                              "Options need to be set before calling set_training_data().");
        if (a == NN_AUTO &&
            (opts.skey("metric") != metric_ || opts.rval("minkowski parameter") != p_))
            return err->error(da_status_option_locked,
                              "Options need to be set before calling set_training_data().");
    }
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::set_data(da_int n_samples, da_int n_features, const T *X, da_int ldx) {
    da_status s = check_2D_array<T>(order(), n_samples, n_features, X, ldx, "n_samples",
                                    "n_features", "X_train", "ldx_train");
    if (s != da_status_success)
        return s;
    s = gpu->init(err);
    if (s != da_status_success)
        return s;
    s = set_params();
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const size_t bytes = (size_t)n_samples * n_features * sizeof(T);
    DA_HIP(err, train_.reserve(bytes));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(train_.ptr(), n_features * sizeof(T), X,
                                     (size_t)ldx * sizeof(T), n_features * sizeof(T), n_samples,
                                     hipMemcpyHostToDevice, st));
    } else {
        DeviceBuffer stage;
        DA_HIP(err, stage.reserve(bytes));
        DA_HIP(err, copy_2d(stage.ptr(), n_samples * sizeof(T), X,
                                     (size_t)ldx * sizeof(T), n_samples * sizeof(T), n_features,
                                     hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(stage.as<T>(), n_samples, n_samples,
                                                   n_features, train_.as<T>(), st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    DA_HIP(err, train_sq_.reserve((size_t)n_samples * sizeof(double)));
    DA_HIP(err, launch_row_sqnorms<T>(train_.as<T>(), n_features, n_samples, n_features,
                                      train_sq_.as<double>(), st));
    DA_HIP(err, train_cos_.reserve((size_t)n_samples * sizeof(double)));
    DA_HIP(err, hipMemcpyAsync(train_cos_.ptr(), train_sq_.ptr(), n_samples * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
    DA_HIP(err, launch_cosine_norms(train_cos_.as<double>(), n_samples,
                                    std::numeric_limits<T>::epsilon(), st));
    DA_HIP(err, hipStreamSynchronize(st));
    n_samples_ = n_samples;
    n_features_ = n_features;
    trained_ = true;
    fused_ready_ = false;
    ypad_prec_ = -1;
    return prepare_fused(tile_prec());
}

// Squared-euclidean metrics are searched by the fused MFMA kernel (up to 128 features, k up
// to 128); everything else uses the key-tile + selection path.
// Most training splits of one fused search: the merge takes nsplit * kc candidates per
// query, and a split has at least 1,024 rows.
template <typename T> int NNHandle<T>::split_cap(int kc) const {
    return (int)std::max<int64_t>(1, std::min<int64_t>(knn_merge_max_candidates() / kc,
                                                       ((int64_t)n_samples_ + 1023) / 1024));
}

template <typename T> bool NNHandle<T>::fused_eligible(da_int k) const {
    if (!fused_ready_)
        return false;
    const bool sq = metric_ == da_euclidean || metric_ == da_sqeuclidean ||
                    metric_ == da_euclidean_gemm || metric_ == da_sqeuclidean_gemm ||
                    (metric_ == da_minkowski && p_ == 2.0);
    return sq && k <= 128 && n_features_ <= 128 && n_samples_ >= 16;
}

// Builds the zero-padded copy of the training rows in the operand precision of the fused
// tile, -|y|^2/2 per row and the largest row norm.
template <typename T> da_status NNHandle<T>::prepare_fused(int prec) {
    if (n_features_ > 128)
        return da_status_success;
    if (prec == KP_F16C)
        return prepare_centred();
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_;
    kp_ = knn_tile_kp(d);
    const size_t esz = (size_t)knn_prec_elem_size(prec) * knn_prec_row_factor(prec);
    // Pack the rows in the operand type, then rearrange them into 16-row tiles whose
    // per-lane operand slices are contiguous across the wave.
    const int rows_pad = (n + 31) / 32 * 32;
    {
        DeviceBuffer packed;
        DA_HIP(err, packed.reserve((size_t)n * kp_ * esz));
        DA_HIP(err, launch_pack_rows_prec<T>(train_.as<T>(), d, n, d, packed.ptr(), kp_, n, prec,
                                             st));
        DA_HIP(err, ypad_.reserve((size_t)rows_pad * kp_ * esz));
        DA_HIP(err, launch_knn_tile_operand(packed.ptr(), n, (int)(kp_ * esz),
                                            knn_prec_row_factor(prec), ypad_.ptr(), rows_pad, st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    std::vector<double> sq(n);
    DA_HIP(err, hipMemcpyAsync(sq.data(), train_sq_.ptr(), n * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    double mx = 0;
    for (int i = 0; i < n; i++)
        mx = std::max(mx, sq[i]);
    ymax_ = std::sqrt(mx);
    // -|y|^2/2 per row, padded to rows_pad entries with -inf (padded rows never score).
    if (prec == KP_F64) {
        std::vector<double> h(rows_pad, -std::numeric_limits<double>::infinity());
        for (int i = 0; i < n; i++)
            h[i] = -0.5 * sq[i];
        DA_HIP(err, nyh_.reserve(rows_pad * sizeof(double)));
        DA_HIP(err, hipMemcpyAsync(nyh_.ptr(), h.data(), rows_pad * sizeof(double),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, hipStreamSynchronize(st));
    } else {
        std::vector<float> h(rows_pad, -std::numeric_limits<float>::infinity());
        for (int i = 0; i < n; i++)
            h[i] = (float)(-0.5 * sq[i]);
        DA_HIP(err, nyh_.reserve(rows_pad * sizeof(float)));
        DA_HIP(err, hipMemcpyAsync(nyh_.ptr(), h.data(), rows_pad * sizeof(float),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    da_status ss = prepare_seed_sample(prec, sq);
    if (ss != da_status_success)
        return ss;
    ypad_prec_ = prec;
    fused_ready_ = true;
    return da_status_success;
}

// The centred fp16 form. The error of a product in the tile scales with the norms of its
// operands; storing each training row as its residual about the centre of its group (a
// k-means of the rows, one centre per 384 rows) shrinks the training operand from |y| to
// |y - c|, and the tile restores the key by starting each accumulator at -|y|^2/2 + q.c.
// The groups are laid out in a fixed random order, each padded to whole 32-row steps (a
// step then lies in one group): neighbours that share a group would otherwise sit in one
// training split, which would lower that split's threshold towards the k-th key and leave
// the query uncertified.
template <typename T> da_status NNHandle<T>::prepare_centred() {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_;
    kp_ = knn_tile_kp(d);
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(4096, n / 384));
    std::vector<double> C;
    std::vector<da_int> lab;
    {
        KMeansHandle<T> km(*err, *gpu);
        km.opts.set_string("initialization method", "random");
        km.opts.set_string("algorithm", "lloyd");
        km.opts.set_int("n_clusters", G);
        km.opts.set_int("n_init", 1);
        km.opts.set_int("max_iter", 8);
        km.opts.set_int("seed", 0);
        km.opts.set_real("convergence tolerance", 0.0);
        da_status s = km.set_data_device(n, d, train_.as<T>());
        if (s != da_status_success)
            return s;
        s = km.compute();
        if (s != da_status_success && s != da_status_maxit)
            return s;
        err->clear();
        C = km.centres();
        lab = km.labels();
    }
    // Stored order: the groups in a fixed random order, rows of a group in ascending index.
    std::vector<int> order(G), size(G, 0);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937 eng(12345);
    std::shuffle(order.begin(), order.end(), eng);
    for (int i = 0; i < n; i++)
        size[lab[i]]++;
    std::vector<int64_t> start(G);
    int64_t rows = 0;
    for (int g : order) {
        start[g] = rows;
        rows += (size[g] + 31) / 32 * 32;
    }
    if (rows > std::numeric_limits<int>::max() - 64)
        return err->error(da_status_memory_error, "Too many rows for the centred form.");
    const int rows_pad = std::max<int>(32, (int)rows);
    std::vector<int> perm(rows_pad, -1), step_group(rows_pad / 32, order[0]);
    std::vector<int64_t> fill(start);
    for (int i = 0; i < n; i++)
        perm[fill[lab[i]]++] = i;
    for (int g = 0; g < G; g++)
        for (int64_t p = start[g]; p < start[g] + (size[g] + 31) / 32 * 32; p += 32)
            step_group[p / 32] = g;
    double cm = 0;
    for (int g = 0; g < G; g++) {
        double c2 = 0;
        for (int l = 0; l < d; l++)
            c2 += C[(size_t)g * d + l] * C[(size_t)g * d + l];
        cm = std::max(cm, c2);
    }
    cmax_ = std::sqrt(cm) * (1 + 1e-12);
    DA_HIP(err, cen_C_.reserve(C.size() * sizeof(double)));
    DA_HIP(err, cen_perm_.reserve(perm.size() * sizeof(int)));
    DA_HIP(err, cen_step_.reserve(step_group.size() * sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(cen_C_.ptr(), C.data(), C.size() * sizeof(double),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipMemcpyAsync(cen_perm_.ptr(), perm.data(), perm.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipMemcpyAsync(cen_step_.ptr(), step_group.data(), step_group.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
    // Residual rows in fp16, rearranged into the tile's operand layout.
    std::vector<double> r2(rows_pad);
    {
        DeviceBuffer packed, rn;
        DA_HIP(err, packed.reserve((size_t)rows_pad * kp_ * 2));
        DA_HIP(err, rn.reserve((size_t)rows_pad * sizeof(double)));
        DA_HIP(err, launch_knn_centre_rows<T>(train_.as<T>(), d, cen_perm_.as<int>(),
                                              cen_step_.as<int>(), cen_C_.as<double>(), rows_pad,
                                              kp_, packed.ptr(), rn.as<double>(), st));
        DA_HIP(err, ypad_.reserve((size_t)rows_pad * kp_ * 2));
        DA_HIP(err, launch_knn_tile_operand(packed.ptr(), rows_pad, kp_ * 2, 1, ypad_.ptr(),
                                            rows_pad, st));
        DA_HIP(err, hipMemcpyAsync(r2.data(), rn.ptr(), rows_pad * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    double rm = 0;
    for (double v : r2)
        rm = std::max(rm, v);
    // The residuals are exact in double; their fp16 rounding is part of the operand error.
    rmax_ = std::sqrt(rm) * (1 + 1e-12);
    std::vector<double> sq(n);
    DA_HIP(err, hipMemcpyAsync(sq.data(), train_sq_.ptr(), n * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    double mx = 0;
    std::vector<float> h(rows_pad, -std::numeric_limits<float>::infinity());
    for (int p = 0; p < rows_pad; p++)
        if (perm[p] >= 0) {
            h[p] = (float)(-0.5 * sq[perm[p]]);
            mx = std::max(mx, sq[perm[p]]);
        }
    ymax_ = std::sqrt(mx);
    DA_HIP(err, nyh_.reserve(rows_pad * sizeof(float)));
    DA_HIP(err, hipMemcpyAsync(nyh_.ptr(), h.data(), rows_pad * sizeof(float),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    // The queries of this form are plain fp16 rows: its seeds come from a plain fp16 sample.
    da_status ss = prepare_seed_sample(KP_F16, sq);
    if (ss != da_status_success)
        return ss;
    cen_G_ = G;
    cen_rows_ = rows_pad;
    ypad_prec_ = KP_F16C;
    fused_ready_ = true;
    return da_status_success;
}

// Sample of the group-minimum threshold seeds: 8,192 rows at an even stride through the
// training rows, in the tile layout of the seed form. The whole training set is a superset
// of the sample, so a bound on a query's k-th key over the sample holds for every sweep.
template <typename T>
da_status NNHandle<T>::prepare_seed_sample(int prec, const std::vector<double> &sq) {
    seed_rows_ = 0;
    seed_prec_ = prec;
    const int n = n_samples_, d = n_features_;
    if (n < 131072 || prec == KP_F64 || prec == KP_F16C)
        return da_status_success;
    constexpr int S = 8192;
    hipStream_t st = gpu->stream();
    std::vector<int> idx(S);
    std::vector<float> h(S);
    for (int j = 0; j < S; j++) {
        idx[j] = (int)((int64_t)j * n / S);
        h[j] = (float)(-0.5 * sq[idx[j]]);
    }
    const size_t esz = (size_t)knn_prec_elem_size(prec) * knn_prec_row_factor(prec);
    {
        DeviceBuffer didx, rows, packed;
        DA_HIP(err, didx.reserve(S * sizeof(int)));
        DA_HIP(err, rows.reserve((size_t)S * d * sizeof(T)));
        DA_HIP(err, packed.reserve((size_t)S * kp_ * esz));
        DA_HIP(err, hipMemcpyAsync(didx.ptr(), idx.data(), S * sizeof(int), hipMemcpyHostToDevice,
                                   st));
        DA_HIP(err, launch_gather_scatter(train_.ptr(), rows.ptr(), didx.as<int>(), S, d,
                                          (int)sizeof(T), 0, st));
        DA_HIP(err, launch_pack_rows_prec<T>(rows.as<T>(), d, S, d, packed.ptr(), kp_, S, prec,
                                             st));
        DA_HIP(err, seed_y_.reserve((size_t)S * kp_ * esz));
        DA_HIP(err, launch_knn_tile_operand(packed.ptr(), S, (int)(kp_ * esz),
                                            knn_prec_row_factor(prec), seed_y_.ptr(), S, st));
        DA_HIP(err, seed_nyh_.reserve(S * sizeof(float)));
        DA_HIP(err, hipMemcpyAsync(seed_nyh_.ptr(), h.data(), S * sizeof(float),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    seed_rows_ = S;
    return da_status_success;
}

template <typename T> SearchParams NNHandle<T>::params() const {
    SearchParams p;
    p.prec = tile_prec();
    p.variant = opts.ival("tile variant");
    p.kc_min = opts.ival("candidates per split");
    const int seed = opts.skey("seed thresholds");
    p.seed = seed == 1 || (seed == 2 && p.prec != KP_F32 && p.prec != KP_F64);
    return p;
}

template <typename T> da_status NNHandle<T>::prepare_search(const SearchParams &p) {
    if (fused_ready_ && p.prec != ypad_prec_)
        return prepare_fused(p.prec);
    return da_status_success;
}

SearchContext::~SearchContext() {
    if (h_fb)
        (void)hipHostFree(h_fb);
}

// Fused path, launch phase: pack the queries, distance tile, merge/certify. The fallback
// count is copied to pinned host memory; search_finish reads it after the stream completes.
// Number of training splits for a batch of nq queries, picked by measurement. With qb
// queries per workgroup the grid has qblocks * nsplit workgroups and the device runs `slots`
// of them at once; s1 = the fewest splits that fill about 90% of the slots. More splits
// shorten each workgroup's sweep but restart more candidate lists (each start costs a burst
// of threshold updates), and a grid just over a whole number of rounds leaves the device
// idle in its last round, so the best count depends on the operand form and the data. The
// first batch of each size class (log2 of qblocks) times the candidates on its own queries
// and keeps the fastest; later batches reuse the choice. Splits are capped by the merge's
// capacity (nsplit * kc candidates per query) and at 1024 rows each, and kept at or above
// the minimum of sweeps with 16-bit row offsets.
template <typename T>
int NNHandle<T>::tile_splits(SearchContext &ctx, int prec, int kc, int variant, da_int nq,
                             const void *qpad, const float *seed, hipStream_t st,
                             const KnnCentring *cen) {
    int slots = 0;
    SearchContext &tc = ctx.tuning ? *ctx.tuning : ctx;
    const int skey = ((prec * 8 + variant) * 2 + (seed ? 1 : 0)) * 256 + kc;
    auto it = tc.slots.find(skey);
    if (it != tc.slots.end()) {
        slots = it->second;
    } else {
        if (knn_tile_slots(prec, kp_, kc, variant, &slots) != hipSuccess || slots <= 0)
            slots = 512;
        tc.slots[skey] = slots;
    }
    const int64_t n = tile_rows(prec);
    const int64_t qb = knn_tile_qb(kc);
    const int64_t qblocks = (nq + qb - 1) / qb;
    const int64_t hi = split_cap(kc);
    const int64_t lo = knn_tile_min_splits(prec, kp_, (int)n, kc, variant);
    const int64_t s1 = std::min<int64_t>(
        hi, std::max<int64_t>(lo, (slots * 9 / 10 + qblocks - 1) / qblocks));
    int bucket = 0;
    while (((int64_t)1 << (bucket + 1)) <= qblocks)
        bucket++;
    const int key = skey * 64 + bucket;
    auto clamp = [&](int64_t v) { return (int)std::max<int64_t>(lo, std::min<int64_t>(hi, v)); };
    auto ft = tc.split_factor.find(key);
    if (ft != tc.split_factor.end())
        return ft->second;
    // Candidates: the most splits whose workgroups all run in one round, and 1, 2 and 4
    // times the count that fills 90% of the slots. Time them (a first untimed launch warms
    // the caches) and keep the fastest.
    std::vector<int> cand;
    for (int64_t v : {(int64_t)slots / std::max<int64_t>(1, qblocks), s1, 2 * s1, 4 * s1}) {
        const int c = clamp(v);
        if (std::find(cand.begin(), cand.end(), c) == cand.end())
            cand.push_back(c);
    }
    int best = cand[0];
    float best_ms = -1.0f;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a)
            (void)hipEventDestroy(a);
        return clamp(s1);
    }
    bool warmed = false;
    for (int ns : cand) {
        const size_t np = (size_t)nq * ns;
        if (ctx.part_key.reserve(np * kc * sizeof(float)) != hipSuccess ||
            ctx.part_idx.reserve(np * kc * sizeof(int)) != hipSuccess ||
            ctx.part_tau.reserve(np * sizeof(float)) != hipSuccess)
            break;
        auto launch = [&]() {
            return launch_knn_tile(prec, qpad, nq, ypad_.ptr(), nyh_.ptr(), tile_rows(prec), kp_,
                                   ns, kc, ctx.part_key.as<float>(), ctx.part_idx.as<int>(),
                                   ctx.part_tau.as<float>(), seed, variant, st, cen);
        };
        if (!warmed) {
            if (launch() != hipSuccess)
                break;
            warmed = true;
        }
        (void)hipEventRecord(a, st);
        if (launch() != hipSuccess)
            break;
        (void)hipEventRecord(b, st);
        if (hipEventSynchronize(b) != hipSuccess)
            break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (best_ms < 0.0f || ms < best_ms) {
            best_ms = ms;
            best = ns;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    tc.split_factor[key] = best;
    return best;
}

// Error model of the keys the fused tile stores for operand form prec over this training set.
// Representation error of an operand element: relative u_in, absolute abs_in (fp16
// subnormals). Split forms keep hi + lo, so only the remainder of lo and the dropped lo*lo
// product (folded into u_in) remain.
template <typename T> KnnKeyError NNHandle<T>::key_error_model(int prec) const {
    const bool same = (prec == KP_F64 && std::is_same_v<T, double>) ||
                      (prec == KP_F32 && std::is_same_v<T, float>);
    KnnKeyError e;
    const double u16 = std::ldexp(1.0, -11), ub16 = std::ldexp(1.0, -8), u32 = std::ldexp(1.0, -24);
    if (!same) {
        switch (prec) {
        case KP_F32:
            e.u_in = u32;
            break;
        case KP_F16:
        case KP_F16C:
            e.u_in = u16 + u32;
            e.abs_in = std::ldexp(1.0, -25);
            break;
        case KP_BF16:
            e.u_in = ub16 + u32;
            break;
        case KP_F16X3:
            e.u_in = 1.5 * u16 * u16 * (1 + u16) + 2 * u32 * u16 + u32 * u32;
            e.abs_in = std::ldexp(1.0, -24);
            break;
        case KP_BF16X3:
            e.u_in = 1.5 * ub16 * ub16 * (1 + ub16) + 2 * u32 * ub16 + u32 * u32;
            break;
        default:
            break;
        }
    }
    e.u_acc = prec == KP_F64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
    e.ymax = ymax_;
    e.rmax = prec == KP_F16C ? rmax_ : ymax_;
    e.cmax = prec == KP_F16C ? cmax_ : 0.0;
    return e;
}

template <typename T>
da_status NNHandle<T>::search_launch(SearchContext &ctx, const SearchParams &sp, const T *dQ,
                                     da_int nq, da_int k, bool return_distance, da_int *d_ind,
                                     T *d_dist, bool &fused, const float *seed_in) {
    int kc = sp.tile_kc(k);
    // With the automatic length, a list shorter than k needs 2k candidates over the splits
    // the training set allows (at most one per 1,024 rows): on a small set the next length up
    // keeps the fused tile instead of the exact key tiles.
    if (sp.kc_min <= 0)
        while (kc < k && kc < 128 && (int64_t)split_cap(kc) * kc < 2 * (int64_t)k)
            kc *= 2;
    int variant = sp.variant;
    if (variant < 0 || sp.prec == KP_F16C)
        variant = sp.prec == KP_F16C ? 3 : knn_tile_auto_variant(sp.prec, kp_);
    fused = fused_eligible(k) && sp.prec == ypad_prec_ &&
            knn_tile_min_splits(sp.prec, kp_, tile_rows(sp.prec), kc, variant) <= split_cap(kc);
    ctx.last_fallback = 0;
    ctx.last_candidates = 0;
    ctx.last_params = sp;
    if (!fused)
        return search_tiles(ctx, dQ, nq, k, return_distance, d_ind, d_dist);
    hipStream_t st = ctx.gpu->stream();
    const int n = n_samples_, d = n_features_;
    const int prec = sp.prec;
    if (!ctx.h_fb)
        DA_HIP(err, hipHostMalloc((void **)&ctx.h_fb, sizeof(int)));
    const size_t esz = (size_t)knn_prec_elem_size(prec) * knn_prec_row_factor(prec);
    DA_HIP(err, ctx.qpad.reserve((size_t)nq * kp_ * esz));
    DA_HIP(err, launch_pack_rows_prec<T>(dQ, d, nq, d, ctx.qpad.ptr(), kp_, nq, prec, st));
    // Threshold seeds (k <= 16): for each query the k-th smallest of 16 group-best keys over
    // the seed sample (launch_knn_seed_min), in splits of at least 2,048 sample rows (groups
    // of 128 rows or more) and about 512 workgroups in all, then the smallest over the
    // splits. The bound lies far above the k-th key (near the (n / 2,048 / ns)-th for k = 10
    // with ns splits), so the seeded sweep still certifies while it buffers only the rows
    // below the seed. At 200,000 x 128 with 20,000 queries (4 splits) the seed takes 0.08 ms.
    const float *seed = seed_in;
    if (!seed && sp.seed && seed_rows_ > 0 && k <= 16) {
        KernelScope ks(*ctx.gpu, "knn_seed");
        const int qblocks = (int)((nq + 127) / 128);
        const int ns = std::max(1, std::min(seed_rows_ / 2048, (512 + qblocks - 1) / qblocks));
        DA_HIP(err, ctx.part_tau.reserve((size_t)nq * ns * sizeof(float)));
        DA_HIP(err, ctx.tau0.reserve((size_t)nq * sizeof(float)));
        DA_HIP(err, launch_knn_seed_min(seed_prec_, ctx.qpad.ptr(), nq, seed_y_.ptr(),
                                        seed_nyh_.ptr(), seed_rows_, kp_, ns, (int)k,
                                        ctx.part_tau.as<float>(), st));
        DA_HIP(err, launch_knn_seed<T>(ctx.part_tau.as<float>(), nq, ns, dQ, n_features_,
                                       key_error_model(prec), key_error_model(seed_prec_),
                                       ctx.tau0.as<float>(), st));
        seed = ctx.tau0.as<float>();
    }
    // The centred form: every query's product with every group centre.
    KnnCentring cen_args;
    const KnnCentring *cen = nullptr;
    if (prec == KP_F16C) {
        KernelScope ks(*ctx.gpu, "knn_query_centre");
        DA_HIP(err, ctx.qc.reserve((size_t)nq * cen_G_ * sizeof(float)));
        DA_HIP(err, launch_knn_query_centre<T>(dQ, nq, d, cen_C_.as<double>(), cen_G_,
                                               ctx.qc.as<float>(), st));
        cen_args.step_group = cen_step_.as<int>();
        cen_args.qc = ctx.qc.as<float>();
        cen_args.G = cen_G_;
        cen_args.perm = cen_perm_.as<int>();
        cen = &cen_args;
    }
    int nsplit = tile_splits(ctx, prec, kc, variant, nq, ctx.qpad.ptr(), seed, st, cen);
    if (kc < k) {
        // Lists shorter than k: enough splits that each holds a small share of the k nearest
        // rows (at least 2k candidates in all), else the exact key-tile search.
        nsplit = std::max(nsplit, std::min(split_cap(kc), (2 * (int)k + kc - 1) / kc));
        if ((int64_t)nsplit * kc < 2 * k || (int64_t)n < (int64_t)nsplit * 32) {
            fused = false;
            return search_tiles(ctx, dQ, nq, k, return_distance, d_ind, d_dist);
        }
    }
    const size_t np = (size_t)nq * nsplit;
    DA_HIP(err, ctx.part_key.reserve(np * kc * sizeof(float)));
    DA_HIP(err, ctx.part_idx.reserve(np * kc * sizeof(int)));
    DA_HIP(err, ctx.part_tau.reserve(np * sizeof(float)));
    DA_HIP(err, ctx.fb.reserve((size_t)(nq + 1) * sizeof(int)));
    DA_HIP(err, ctx.fb_bound.reserve((size_t)nq * sizeof(float)));
    {
        KernelScope ks(*ctx.gpu, ctx.second_pass ? "knn_retry_tile" : "knn_tile");
        DA_HIP(err, launch_knn_tile(prec, ctx.qpad.ptr(), nq, ypad_.ptr(), nyh_.ptr(),
                                    tile_rows(prec), kp_, nsplit, kc, ctx.part_key.as<float>(),
                                    ctx.part_idx.as<int>(), ctx.part_tau.as<float>(), seed,
                                    variant, st, cen));
    }
    int *fbc = ctx.fb.as<int>();
    DA_HIP(err, hipMemsetAsync(fbc, 0, sizeof(int), st));
    const int post = (metric_ == da_sqeuclidean || metric_ == da_sqeuclidean_gemm) ? 0 : 1;
    const KnnKeyError ke = key_error_model(prec);
    {
        KernelScope ks(*ctx.gpu, ctx.second_pass ? "knn_retry_merge" : "knn_merge");
        DA_HIP(err, launch_knn_merge<T>(ctx.part_key.as<float>(), ctx.part_idx.as<int>(),
                                        ctx.part_tau.as<float>(), nsplit, kc, dQ, train_.as<T>(),
                                        nq, d, k, ke.u_in, ke.abs_in, ke.u_acc, ke.ymax, ke.rmax,
                                        ke.cmax, post,
                                        return_distance ? d_dist : nullptr, d_ind, fbc, fbc + 1,
                                        ctx.fb_bound.as<float>(), st));
    }
    DA_HIP(err, hipMemcpyAsync(ctx.h_fb, fbc, sizeof(int), hipMemcpyDeviceToHost, st));
    ctx.last_candidates = nsplit * kc;
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::search_finish(SearchContext &ctx, const T *dQ, da_int nq, da_int k,
                                     bool return_distance, da_int *d_ind, T *d_dist,
                                     bool fused) {
    (void)nq;
    if (!fused)
        return da_status_success;
    hipStream_t st = ctx.gpu->stream();
    DA_HIP(err, hipStreamSynchronize(st));
    const int nfb = *ctx.h_fb;
    ctx.last_fallback = nfb;
    if (nfb > 0) {
        // The uncertified queries (listed by the merge kernel): gather their rows, search
        // them again, scatter the results back. The first pass is followed by a second
        // fused pass over them with lists of at least 32, seeded by the merge's bound on
        // each query's k-th key: it buffers only the rows below that bound (a few dozen), so
        // even a split that holds all of the rows near the k-th key (the centred form stores
        // neighbouring rows together) keeps them, and its threshold lies above the bound;
        // nearly all of them certify there. The rest take the exact fallback (key tiles +
        // selection).
        const int d = n_features_;
        const int *list = ctx.fb.as<int>() + 1;
        DA_HIP(err, ctx.fb_q.reserve((size_t)nfb * d * sizeof(T)));
        DA_HIP(err, launch_gather_scatter(dQ, ctx.fb_q.ptr(), list, nfb, d, sizeof(T), 0, st));
        DA_HIP(err, ctx.fb_ind.reserve((size_t)nfb * k * sizeof(da_int)));
        DA_HIP(err, ctx.fb_dist.reserve((size_t)nfb * k * sizeof(T)));
        da_status s;
        if (!ctx.second_pass) {
            if (!ctx.retry)
                ctx.retry = std::make_unique<SearchContext>();
            SearchContext &rc = *ctx.retry;
            rc.gpu = ctx.gpu;
            rc.tuning = ctx.tuning ? ctx.tuning : &ctx;
            rc.second_pass = true;
            SearchParams sp2 = ctx.last_params;
            sp2.seed = false;
            sp2.kc_min = std::max(sp2.kc_min, 32);
            bool fused2 = false;
            s = search_launch(rc, sp2, ctx.fb_q.as<T>(), nfb, k, true, ctx.fb_ind.as<da_int>(),
                              ctx.fb_dist.as<T>(), fused2, ctx.fb_bound.as<float>());
            if (s == da_status_success)
                s = search_finish(rc, ctx.fb_q.as<T>(), nfb, k, true, ctx.fb_ind.as<da_int>(),
                                  ctx.fb_dist.as<T>(), fused2);
        } else {
            s = search_tiles(ctx, ctx.fb_q.as<T>(), nfb, k, true, ctx.fb_ind.as<da_int>(),
                             ctx.fb_dist.as<T>());
        }
        if (s != da_status_success)
            return s;
        DA_HIP(err, launch_gather_scatter(ctx.fb_ind.ptr(), d_ind, list, nfb, k, sizeof(da_int), 1, st));
        if (return_distance)
            DA_HIP(err, launch_gather_scatter(ctx.fb_dist.ptr(), d_dist, list, nfb, k, sizeof(T), 1, st));
    }
    return da_status_success;
}

template <typename T> da_status NNHandle<T>::resolve_k(da_int &k) {
    if (k <= 0 && n_neighbors_ <= 0)
        return err->error(da_status_invalid_input,
                          "Number of requested neighbors must be positive.");
    if (k <= 0)
        k = opts.ival("number of neighbors");
    if (k > n_samples_)
        return err->error(da_status_invalid_input,
                          "Number of requested neighbors must be at least as big as the "
                          "number of samples.");
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::kneighbors(da_int n_queries, da_int n_features, const T *X_test,
                                  da_int ldx_test, da_int *n_ind, T *n_dist, da_int k,
                                  da_int return_distance) {
    // Checks in the order of nearest_neighbors.cpp:939-995 and :61-106.
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (n_ind == nullptr)
        return err->error(da_status_invalid_pointer, "n_ind is not a valid pointer.");
    if (X_test == nullptr)
        return err->error(da_status_invalid_pointer, "X_test is not a valid pointer.");
    if (n_features != n_features_)
        return err->error(da_status_invalid_array_dimension,
                          "n_features = " + std::to_string(n_features) +
                              " doesn't match the expected value " +
                              std::to_string(n_features_) + ".");
    s = check_2D_array<T>(order(), n_queries, n_features, X_test, ldx_test, "n_queries",
                          "n_features", "X_test", "ldx_test");
    if (s != da_status_success)
        return s;
    s = resolve_k(k);
    if (s != da_status_success)
        return s;
    if (return_distance && n_dist == nullptr)
        return err->error(da_status_invalid_pointer, "n_dist is not a valid pointer.");

    hipStream_t st = gpu->stream();
    const size_t qbytes = (size_t)n_queries * n_features * sizeof(T);
    DA_HIP(err, q_.reserve(qbytes));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(q_.ptr(), n_features * sizeof(T), X_test,
                                     (size_t)ldx_test * sizeof(T), n_features * sizeof(T),
                                     n_queries, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, q_stage_.reserve(qbytes));
        DA_HIP(err, copy_2d(q_stage_.ptr(), n_queries * sizeof(T), X_test,
                                     (size_t)ldx_test * sizeof(T), n_queries * sizeof(T),
                                     n_features, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(q_stage_.as<T>(), n_queries, n_queries,
                                                   n_features, q_.as<T>(), st));
    }
    const size_t ob = (size_t)n_queries * k;
    DA_HIP(err, out_ind_.reserve(ob * sizeof(da_int)));
    DA_HIP(err, out_dist_.reserve(ob * sizeof(T)));
    s = search(q_.as<T>(), n_queries, k, return_distance != 0, out_ind_.as<da_int>(),
               return_distance ? out_dist_.as<T>() : nullptr);
    if (s != da_status_success)
        return s;
    // Output is n_queries x k in the handle's storage order (nearest_neighbors.cpp:997-1009).
    if (order() == row_major) {
        DA_HIP(err, hipMemcpyAsync(n_ind, out_ind_.ptr(), ob * sizeof(da_int),
                                   hipMemcpyDeviceToHost, st));
        if (return_distance)
            DA_HIP(err, hipMemcpyAsync(n_dist, out_dist_.ptr(), ob * sizeof(T),
                                       hipMemcpyDeviceToHost, st));
    } else {
        // Transpose on the device (k x n_queries row-major == n_queries x k column-major).
        DA_HIP(err, q_stage_.reserve(ob * std::max(sizeof(T), sizeof(da_int))));
        DA_HIP(err, launch_colmajor_to_rowmajor<float>(
                        reinterpret_cast<const float *>(out_ind_.ptr()), k, k, n_queries,
                        q_stage_.as<float>(), st));
        DA_HIP(err, hipMemcpyAsync(n_ind, q_stage_.ptr(), ob * sizeof(da_int),
                                   hipMemcpyDeviceToHost, st));
        if (return_distance) {
            DA_HIP(err, hipStreamSynchronize(st));
            DA_HIP(err, launch_colmajor_to_rowmajor<T>(out_dist_.as<T>(), k, k, n_queries,
                                                       q_stage_.as<T>(), st));
            DA_HIP(err, hipMemcpyAsync(n_dist, q_stage_.ptr(), ob * sizeof(T),
                                       hipMemcpyDeviceToHost, st));
        }
    }
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::kneighbors_device(da_int n_queries, da_int n_features, const T *d_X_test,
                                         da_int ldx_test, da_int *d_ind, T *d_dist, da_int k,
                                         da_int return_distance) {
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (d_ind == nullptr || d_X_test == nullptr || (return_distance && d_dist == nullptr))
        return err->error(da_status_invalid_pointer, "A device pointer is null.");
    if (n_features != n_features_)
        return err->error(da_status_invalid_array_dimension,
                          "n_features doesn't match the training data.");
    if (n_queries < 1)
        return err->error(da_status_invalid_array_dimension, "n_queries must be >= 1.");
    if (ldx_test < n_features)
        return err->error(da_status_invalid_leading_dimension, "ldx_test < n_features.");
    s = resolve_k(k);
    if (s != da_status_success)
        return s;
    const T *dq = d_X_test;
    if (ldx_test != n_features) {
        DA_HIP(err, q_.reserve((size_t)n_queries * n_features * sizeof(T)));
        DA_HIP(err, copy_2d(q_.ptr(), n_features * sizeof(T), d_X_test,
                                     (size_t)ldx_test * sizeof(T), n_features * sizeof(T),
                                     n_queries, hipMemcpyDeviceToDevice, gpu->stream()));
        dq = q_.as<T>();
    }
    return search(dq, n_queries, k, return_distance != 0, d_ind,
                  return_distance ? d_dist : nullptr);
}

template <typename T>
da_status NNHandle<T>::search(const T *dQ, da_int nq, da_int k, bool return_distance,
                              da_int *d_ind, T *d_dist) {
    main_ctx.gpu = gpu;
    const SearchParams sp = params();
    da_status s = prepare_search(sp);
    if (s != da_status_success)
        return s;
    bool fused = false;
    s = search_launch(main_ctx, sp, dQ, nq, k, return_distance, d_ind, d_dist, fused);
    if (s != da_status_success)
        return s;
    s = search_finish(main_ctx, dQ, nq, k, return_distance, d_ind, d_dist, fused);
    last_fallback = main_ctx.last_fallback;
    last_candidates = main_ctx.last_candidates;
    return s;
}

// Generic path for every metric: key tiles (queries x training points, in double) followed
// by the per-row selection, the structure of kneighbors_compute_brute_force
// (nearest_neighbors.cpp:567-816) with the blocks sized for HBM instead of caches.
template <typename T>
da_status NNHandle<T>::search_tiles(SearchContext &ctx, const T *dQ, da_int nq, da_int k,
                                    bool return_distance, da_int *d_ind, T *d_dist) {
    hipStream_t st = ctx.gpu->stream();
    const da_int d = n_features_, n = n_samples_;
    if (k > 3840)
        return err->error(da_status_not_implemented,
                          "k > 3840 neighbours is not supported by the GPU selection kernel.");
    int km, post = 0, largest = 0;
    const double *nq_norm = nullptr, *nt_norm = nullptr;
    double p = p_;
    bool need_qcos = false, need_qsq = false;
    switch (metric_) {
    case da_euclidean:
        km = KM_SQEUCLIDEAN;
        post = 1;
        break;
    case da_sqeuclidean:
        km = KM_SQEUCLIDEAN;
        break;
    case da_euclidean_gemm:
        km = KM_SQEUCLIDEAN_GEMM;
        post = 1;
        need_qsq = true;
        break;
    case da_sqeuclidean_gemm:
        km = KM_SQEUCLIDEAN_GEMM;
        need_qsq = true;
        break;
    case da_manhattan:
        km = KM_MANHATTAN;
        break;
    case da_cosine:
        km = KM_COSINE;
        need_qcos = true;
        break;
    case da_inner_product:
        km = KM_INNER;
        largest = 1;
        break;
    case da_minkowski:
        if (p == 2.0) {
            km = KM_SQEUCLIDEAN;
            post = 1;
        } else if (p == 1.0) {
            km = KM_MANHATTAN;
        } else {
            km = KM_MINKOWSKI_KEY;
            post = 2;
        }
        break;
    default:
        return err->error(da_status_not_implemented, "Unknown metric.");
    }
    if (need_qsq || need_qcos) {
        DA_HIP(err, ctx.q_norm.reserve((size_t)nq * sizeof(double)));
        DA_HIP(err, launch_row_sqnorms<T>(dQ, d, nq, d, ctx.q_norm.as<double>(), st));
        if (need_qcos) {
            DA_HIP(err, launch_cosine_norms(ctx.q_norm.as<double>(), nq,
                                            std::numeric_limits<T>::epsilon(), st));
            nt_norm = train_cos_.as<double>();
        } else {
            nt_norm = train_sq_.as<double>();
        }
        nq_norm = ctx.q_norm.as<double>();
    }
    const size_t max_keys = (size_t)64 << 20; // 512 MB of double keys per chunk
    da_int chunk = (da_int)std::max<size_t>(1, std::min<size_t>(nq, max_keys / (size_t)n));
    DA_HIP(err, ctx.keys.reserve((size_t)chunk * n * sizeof(double)));
    DA_HIP(err, ctx.sel_key.reserve((size_t)chunk * k * sizeof(double)));
    DA_HIP(err, ctx.sel_idx.reserve((size_t)chunk * k * sizeof(int)));
    for (da_int q0 = 0; q0 < nq; q0 += chunk) {
        const da_int rows = std::min(chunk, nq - q0);
        {
            KernelScope ks(*ctx.gpu, "pairwise_direct");
            DA_HIP(err, (launch_pairwise_direct<T, double>(
                            km, dQ + (size_t)q0 * d, train_.as<T>(), rows, n, d,
                            ctx.keys.as<double>(), n, nq_norm ? nq_norm + q0 : nullptr, nt_norm, p,
                            0, 0, st)));
        }
        {
            KernelScope ks(*ctx.gpu, "row_topk");
            DA_HIP(err, launch_row_topk<double>(ctx.keys.as<double>(), n, rows, n, k, largest,
                                                ctx.sel_key.as<double>(), ctx.sel_idx.as<int>(), st));
        }
        DA_HIP(err, launch_knn_finalize<T>(ctx.sel_key.as<double>(), ctx.sel_idx.as<int>(), rows, k,
                                           post, p,
                                           return_distance ? d_dist + (size_t)q0 * k : nullptr,
                                           d_ind + (size_t)q0 * k, k, st));
    }
    return da_status_success;
}

// Copies n_queries x n_features queries (handle storage order, leading dimension ld) into
// q_ as compact row-major rows.
template <typename T>
da_status NNHandle<T>::upload_queries(da_int n_queries, const T *X_test, da_int ldx_test) {
    hipStream_t st = gpu->stream();
    const da_int d = n_features_;
    const size_t qbytes = (size_t)n_queries * d * sizeof(T);
    DA_HIP(err, q_.reserve(qbytes));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(q_.ptr(), d * sizeof(T), X_test, (size_t)ldx_test * sizeof(T),
                            d * sizeof(T), n_queries, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, q_stage_.reserve(qbytes));
        DA_HIP(err, copy_2d(q_stage_.ptr(), n_queries * sizeof(T), X_test,
                            (size_t)ldx_test * sizeof(T), n_queries * sizeof(T), d,
                            hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(q_stage_.as<T>(), n_queries, n_queries, d,
                                                   q_.as<T>(), st));
    }
    return da_status_success;
}

// Radius neighbours by brute force (nearest_neighbors.cpp:1918-2215): the squared-euclidean
// forms compare squared distances with radius^2 and return square roots; the others compare
// the distance itself. Keys are computed exactly (double direct differences) and rounded to
// T, then compared in T as the reference does. Each query's neighbours are in ascending
// training index, or ordered by (distance, index) with sort_results.
template <typename T>
da_status NNHandle<T>::radius_neighbors(da_int n_queries, da_int n_features, const T *X_test,
                                        da_int ldx_test, T radius, da_int return_distance,
                                        da_int sort_results) {
    if (!return_distance && sort_results)
        return err->error(da_status_invalid_input,
                          "Cannot sort results if distances are not returned.");
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (metric_ == da_inner_product)
        return err->error(da_status_incompatible_options,
                          "The inner product metric is not compatible with radius neighbors. "
                          "Use da_nn_kneighbors to find the k most similar training vectors "
                          "by inner product.");
    if (X_test == nullptr)
        return err->error(da_status_invalid_pointer, "X_test is not a valid pointer.");
    if (n_features != n_features_)
        return err->error(da_status_invalid_array_dimension,
                          "n_features = " + std::to_string(n_features) +
                              " doesn't match the expected value " +
                              std::to_string(n_features_) + ".");
    s = check_2D_array<T>(order(), n_queries, n_features, X_test, ldx_test, "n_queries",
                          "n_features", "X_test", "ldx_test");
    if (s != da_status_success)
        return s;
    const T opt_radius = (T)opts.rval("radius");
    if (radius < 0 && opt_radius < 0)
        return err->error(da_status_invalid_input,
                          "Radius of requested neighbors must be non-negative.");
    if (radius < 0)
        radius = opt_radius;
    rnn_done_ = false;
    model_trained = false;
    s = upload_queries(n_queries, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    s = radius_core(q_.as<T>(), n_queries, radius, return_distance != 0, sort_results != 0,
                    nullptr);
    if (s != da_status_success)
        return s;
    rnn_dist_ = return_distance != 0;
    rnn_done_ = true;
    model_trained = true;
    return da_status_success;
}

// The search itself, on row-major queries dQ in GPU memory. Without a sink the results go to
// the handle's host arrays (counts, indices, distances); with one, each query block's lists
// stay in GPU memory and are handed to the sink: offsets (rows + 1, from 0), indices and
// distances (when want_dist) of queries q0 .. q0 + rows - 1.
template <typename T>
da_status NNHandle<T>::radius_core(const T *dQ, da_int n_queries, T radius, bool want_dist,
                                   bool sort_results, const RadiusSink *sink) {
    if (!sink) {
        rnn_count_.assign(n_queries, 0);
        rnn_ind_.clear();
        rnn_d_.clear();
    }
    hipStream_t st = gpu->stream();
    const da_int d = n_features_, n = n_samples_;
    // Keys and thresholds per metric (set_params, nearest_neighbors.cpp:337-347).
    int km;
    bool squares = false, need_qsq = false, need_qcos = false;
    switch (metric_) {
    case da_euclidean:
        km = KM_SQEUCLIDEAN;
        squares = true;
        break;
    case da_euclidean_gemm:
        km = KM_SQEUCLIDEAN_GEMM;
        squares = true;
        need_qsq = true;
        break;
    case da_sqeuclidean:
        km = KM_SQEUCLIDEAN;
        break;
    case da_sqeuclidean_gemm:
        km = KM_SQEUCLIDEAN_GEMM;
        need_qsq = true;
        break;
    case da_manhattan:
        km = KM_MANHATTAN;
        break;
    case da_cosine:
        km = KM_COSINE;
        need_qcos = true;
        break;
    case da_minkowski:
        if (p_ == 2.0) {
            km = KM_SQEUCLIDEAN;
            squares = true;
        } else {
            km = p_ == 1.0 ? KM_MANHATTAN : KM_MINKOWSKI;
        }
        break;
    default:
        return err->error(da_status_not_implemented, "Unknown metric.");
    }
    const T thr = squares ? radius * radius : radius;
    const double *nq_norm = nullptr, *nt_norm = nullptr;
    if (need_qsq || need_qcos) {
        DA_HIP(err, main_ctx.q_norm.reserve((size_t)n_queries * sizeof(double)));
        DA_HIP(err, launch_row_sqnorms<T>(dQ, d, n_queries, d, main_ctx.q_norm.as<double>(), st));
        if (need_qcos) {
            DA_HIP(err, launch_cosine_norms(main_ctx.q_norm.as<double>(), n_queries,
                                            std::numeric_limits<T>::epsilon(), st));
            nt_norm = train_cos_.as<double>();
        } else {
            nt_norm = train_sq_.as<double>();
        }
        nq_norm = main_ctx.q_norm.as<double>();
    }
    // Query blocks whose key tile stays within 256 MB (and 2^28 entries).
    const size_t max_keys = ((size_t)256 << 20) / sizeof(T);
    const da_int chunk = (da_int)std::max<size_t>(
        1, std::min<size_t>(n_queries, std::min<size_t>(max_keys, (size_t)1 << 28) / (size_t)n));
    const int seg = radius_segment_len(n), nseg = (n + seg - 1) / seg;
    DA_HIP(err, rnn_keys_.reserve((size_t)chunk * n * sizeof(T)));
    DA_HIP(err, rnn_cnt_.reserve((size_t)chunk * nseg * sizeof(int)));
    DA_HIP(err, rnn_tile_.reserve((size_t)chunk * ((n + 63) / 64) * sizeof(int)));
    DA_HIP(err, rnn_segoff_.reserve((size_t)chunk * nseg * sizeof(int)));
    DA_HIP(err, rnn_off_.reserve((size_t)(chunk + 1) * sizeof(int)));
    std::vector<int> scnt((size_t)chunk * nseg), soff((size_t)chunk * nseg), off(chunk + 1);
    for (da_int q0 = 0; q0 < n_queries; q0 += chunk) {
        const da_int rows = std::min(chunk, n_queries - q0);
        {
            KernelScope ks(*gpu, "pairwise_direct");
            DA_HIP(err, (launch_pairwise_direct_count<T, T>(
                            km, dQ + (size_t)q0 * d, train_.as<T>(), rows, n, d, rnn_keys_.as<T>(),
                            n, nq_norm ? nq_norm + q0 : nullptr, nt_norm, p_, thr, -1,
                            rnn_tile_.as<int>(), st)));
        }
        KernelScope ks(*gpu, "radius_select");
        DA_HIP(err, launch_radius_segsum(rnn_tile_.as<int>(), rows, n, rnn_cnt_.as<int>(), st));
        DA_HIP(err, hipMemcpyAsync(scnt.data(), rnn_cnt_.ptr(), (size_t)rows * nseg * sizeof(int),
                                   hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
        // Offsets per row (for the sort) and per row segment (for the fill).
        off[0] = 0;
        for (da_int r = 0; r < rows; r++) {
            int c = 0;
            for (int g = 0; g < nseg; g++) {
                soff[(size_t)r * nseg + g] = off[r] + c;
                c += scnt[(size_t)r * nseg + g];
            }
            off[r + 1] = off[r] + c;
            if (!sink)
                rnn_count_[q0 + r] = c;
        }
        const int total = off[rows];
        if (total == 0 && !sink)
            continue;
        DA_HIP(err, hipMemcpyAsync(rnn_off_.ptr(), off.data(), (rows + 1) * sizeof(int),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, hipMemcpyAsync(rnn_segoff_.ptr(), soff.data(),
                                   (size_t)rows * nseg * sizeof(int), hipMemcpyHostToDevice, st));
        DA_HIP(err, rnn_ind_buf_.reserve((size_t)std::max(total, 1) * sizeof(int)));
        DA_HIP(err, rnn_key_buf_.reserve((size_t)std::max(total, 1) * sizeof(T)));
        DA_HIP(err, launch_radius_fill<T>(rnn_keys_.as<T>(), n, rows, n, thr, -1,
                                          rnn_segoff_.as<int>(),
                                          rnn_ind_buf_.as<int>(),
                                          want_dist ? rnn_key_buf_.as<T>() : nullptr, st));
        int *ind = rnn_ind_buf_.as<int>();
        T *key = rnn_key_buf_.as<T>();
        if (sort_results && total > 0) {
            size_t bytes = 0;
            DA_HIP(err, radius_sort_bytes<T>(total, rows, &bytes));
            DA_HIP(err, rnn_tmp_.reserve(std::max<size_t>(bytes, 1)));
            DA_HIP(err, rnn_ind2_.reserve((size_t)total * sizeof(int)));
            DA_HIP(err, rnn_key2_.reserve((size_t)total * sizeof(T)));
            DA_HIP(err, launch_radius_sort<T>(rnn_tmp_.ptr(), bytes, key, rnn_key2_.as<T>(), ind,
                                              rnn_ind2_.as<int>(), total, rows,
                                              rnn_off_.as<int>(), st));
            ind = rnn_ind2_.as<int>();
            key = rnn_key2_.as<T>();
        }
        if (want_dist && squares)
            DA_HIP(err, launch_radius_sqrt<T>(key, total, st));
        if (sink) {
            da_status ss = (*sink)(q0, rows, rnn_off_.as<int>(), ind, key);
            if (ss != da_status_success)
                return ss;
            continue;
        }
        const size_t base = rnn_ind_.size();
        rnn_ind_.resize(base + total);
        DA_HIP(err, hipMemcpyAsync(rnn_ind_.data() + base, ind, (size_t)total * sizeof(int),
                                   hipMemcpyDeviceToHost, st));
        if (want_dist) {
            rnn_d_.resize(base + total);
            DA_HIP(err, hipMemcpyAsync(rnn_d_.data() + base, key, (size_t)total * sizeof(T),
                                       hipMemcpyDeviceToHost, st));
        }
        DA_HIP(err, hipStreamSynchronize(st));
    }
    return da_status_success;
}

// ---- Classification and regression (nearest_neighbors.cpp:515-566, 1021-1910) ----

template <typename T>
da_status NNHandle<T>::set_labels(da_int n_samples, const da_int *y_train) {
    if (trained_ && n_samples != n_samples_)
        return err->error(da_status_invalid_array_dimension,
                          "n_samples = " + std::to_string(n_samples) +
                              " doesn't match the training data size " +
                              std::to_string(n_samples_) + ".");
    if (n_samples < 1)
        return err->error(da_status_invalid_array_dimension,
                          "The function was called with n_samples = " +
                              std::to_string(n_samples) + ". Constraint: n_samples >= 1.");
    if (y_train == nullptr)
        return err->error(da_status_invalid_pointer, "The array y_train_class is null.");
    y_class_.assign(y_train, y_train + n_samples);
    has_labels_ = true;
    classes_ready_ = false;
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::set_targets(da_int n_samples, const T *y_train) {
    if (trained_ && n_samples != n_samples_)
        return err->error(da_status_invalid_array_dimension,
                          "n_samples = " + std::to_string(n_samples) +
                              " doesn't match the training data size " +
                              std::to_string(n_samples_) + ".");
    if (n_samples < 1)
        return err->error(da_status_invalid_array_dimension,
                          "The function was called with n_samples = " +
                              std::to_string(n_samples) + ". Constraint: n_samples >= 1.");
    if (y_train == nullptr)
        return err->error(da_status_invalid_pointer, "The array y_train_reg is null.");
    y_reg_.assign(y_train, y_train + n_samples);
    has_targets_ = true;
    DA_HIP(err, tgt_.reserve((size_t)n_samples * sizeof(T)));
    DA_HIP(err, hipMemcpyAsync(tgt_.ptr(), y_reg_.data(), (size_t)n_samples * sizeof(T),
                               hipMemcpyHostToDevice, gpu->stream()));
    DA_HIP(err, hipStreamSynchronize(gpu->stream()));
    return da_status_success;
}

// Sorted distinct labels (available_classes, :1060-1086) and each training row's position
// among them, in GPU memory.
template <typename T> da_status NNHandle<T>::available_classes() {
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    if (!has_labels_)
        return err->error(da_status_no_data, "No classification labels have been set. Please "
                                             "call da_nn_set_labels_s or da_nn_set_labels_d.");
    if (classes_ready_)
        return da_status_success;
    if ((da_int)y_class_.size() != n_samples_)
        return err->error(da_status_invalid_array_dimension,
                          "The number of labels doesn't match the training data size " +
                              std::to_string(n_samples_) + ".");
    classes_ = y_class_;
    std::sort(classes_.begin(), classes_.end());
    classes_.erase(std::unique(classes_.begin(), classes_.end()), classes_.end());
    std::vector<int> idx(n_samples_);
    for (da_int i = 0; i < n_samples_; i++)
        idx[i] = (int)(std::lower_bound(classes_.begin(), classes_.end(), y_class_[i]) -
                       classes_.begin());
    hipStream_t st = gpu->stream();
    DA_HIP(err, cls_idx_.reserve((size_t)n_samples_ * sizeof(int)));
    DA_HIP(err, classes_d_.reserve(classes_.size() * sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(cls_idx_.ptr(), idx.data(), (size_t)n_samples_ * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipMemcpyAsync(classes_d_.ptr(), classes_.data(), classes_.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    classes_ready_ = true;
    return da_status_success;
}

template <typename T> da_status NNHandle<T>::classes(da_int *n_classes, da_int *classes) {
    // nearest_neighbors_public.hpp:79-98: n_classes <= 0 queries the count.
    if (n_classes == nullptr)
        return err->error(da_status_invalid_pointer, "n_classes is not a valid pointer.");
    da_status s = available_classes();
    if (s != da_status_success)
        return s;
    if (*n_classes <= 0) {
        *n_classes = (da_int)classes_.size();
        return da_status_success;
    }
    if (classes == nullptr)
        return err->error(da_status_invalid_pointer, "classes is not a valid pointer.");
    const da_int m = std::min<da_int>(*n_classes, (da_int)classes_.size());
    std::copy(classes_.begin(), classes_.begin() + m, classes);
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::check_queries(da_int n_queries, da_int n_features, const T *X_test,
                                     da_int ldx_test) {
    if (X_test == nullptr)
        return err->error(da_status_invalid_pointer, "X_test is not a valid pointer.");
    if (n_features != n_features_)
        return err->error(da_status_invalid_array_dimension,
                          "n_features = " + std::to_string(n_features) +
                              " doesn't match the expected value " +
                              std::to_string(n_features_) + ".");
    return check_2D_array<T>(order(), n_queries, n_features, X_test, ldx_test, "n_queries",
                             "n_features", "X_test", "ldx_test");
}

// The most frequent label, the smallest at equal counts (most_frequent_element, :1035-1058).
static da_int most_frequent_label(const std::vector<da_int> &y) {
    std::map<da_int, da_int> freq;
    for (da_int v : y)
        freq[v]++;
    da_int best = y[0], bc = 0;
    for (const auto &f : freq)
        if (f.second > bc) {
            best = f.first;
            bc = f.second;
        }
    return best;
}

template <typename T>
da_status NNHandle<T>::classify(da_int nq, da_int search_mode, T *d_proba, int *d_pred) {
    const int weights = (int)opts.skey("weights");
    if (metric_ == da_inner_product && weights == 1)
        return err->error(da_status_incompatible_options,
                          "The inner product metric is not compatible with weights=\"distance\". "
                          "Use weights=\"uniform\" instead.");
    da_status s = available_classes();
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const int ncls = (int)classes_.size();
    const int big = std::numeric_limits<int>::max();
    DA_HIP(err, vote_flag_.reserve(sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(vote_flag_.ptr(), &big, sizeof(int), hipMemcpyHostToDevice, st));
    int *flag = vote_flag_.as<int>();
    const int *cls = cls_idx_.as<int>(), *cl = classes_d_.as<int>();
    const int mode = (int)opts.skey("outlier handling");
    int fill_idx = -1;
    da_int fallback = 0;
    if (search_mode == knn_search_mode) {
        const da_int k = n_neighbors_;
        if (k > n_samples_)
            return err->error(da_status_invalid_input,
                              "The number of neighbors (" + std::to_string(k) +
                                  ") must not exceed the number of training samples.");
        DA_HIP(err, out_ind_.reserve((size_t)nq * k * sizeof(da_int)));
        DA_HIP(err, out_dist_.reserve((size_t)nq * k * sizeof(T)));
        s = search(q_.as<T>(), nq, k, weights == 1, out_ind_.as<da_int>(),
                   weights == 1 ? out_dist_.as<T>() : nullptr);
        if (s != da_status_success)
            return s;
        KernelScope ks(*gpu, "nn_vote");
        DA_HIP(err, launch_nn_classify<T>(nq, nullptr, k, out_ind_.as<int>(), out_dist_.as<T>(),
                                          weights, cls, cl, ncls, 0, -1, 0, d_proba, d_pred, 0,
                                          flag, st));
    } else if (search_mode == radius_search_mode) {
        if (metric_ == da_inner_product)
            return err->error(da_status_incompatible_options,
                              "The inner product metric is not compatible with radius "
                              "neighbors.");
        if (mode == 1) {
            fallback = opts.ival("outlier label");
            auto it = std::find(classes_.begin(), classes_.end(), fallback);
            fill_idx = it == classes_.end() ? -1 : (int)(it - classes_.begin());
        } else if (mode == 2) {
            fallback = most_frequent_label(y_class_);
            fill_idx = (int)(std::lower_bound(classes_.begin(), classes_.end(), fallback) -
                             classes_.begin());
        }
        RadiusSink sink = [&](da_int q0, int rows, const int *off, const int *ind,
                              const T *dist) -> da_status {
            KernelScope ks(*gpu, "nn_vote");
            DA_HIP(err, launch_nn_classify<T>(rows, off, 0, ind, dist, weights, cls, cl, ncls,
                                              mode, fill_idx, (int)fallback,
                                              d_proba + (size_t)q0 * ncls,
                                              d_pred ? d_pred + q0 : nullptr, (int)q0, flag, st));
            return da_status_success;
        };
        s = radius_core(q_.as<T>(), nq, (T)opts.rval("radius"), weights == 1, false, &sink);
        if (s != da_status_success)
            return s;
    } else {
        return err->error(da_status_invalid_input,
                          "Unknown search mode: " + std::to_string(search_mode) + ".");
    }
    int first = big;
    DA_HIP(err, hipMemcpyAsync(&first, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    if (first < nq && search_mode == radius_search_mode) {
        if (mode == 0)
            return err->error(da_status_operation_failed,
                              "Failed to compute probabilities for query " +
                                  std::to_string(first) +
                                  " since it does not have any neighbors within the specified "
                                  "radius.");
        if (mode == 1 && fill_idx == -1)
            return err->warn(da_status_outlier_warning,
                             "The manually set outlier label is not present in the training "
                             "labels. Setting probabilities to 0.0.");
    }
    return da_status_success;
}

template <typename T>
da_status NNHandle<T>::predict_proba(da_int n_queries, da_int n_features, const T *X_test,
                                     da_int ldx_test, T *proba, da_int search_mode) {
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    if (!has_labels_)
        return err->error(da_status_no_data, "No classification labels have been set. Please "
                                             "call da_nn_set_labels_s or da_nn_set_labels_d.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (proba == nullptr)
        return err->error(da_status_invalid_pointer, "proba is not a valid pointer.");
    if (n_queries < 1)
        return err->error(da_status_invalid_array_dimension,
                          "Number of queries must be greater than zero.");
    if (n_features < 1)
        return err->error(da_status_invalid_array_dimension,
                          "Number of features must be greater than zero.");
    s = check_queries(n_queries, n_features, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    s = available_classes();
    if (s != da_status_success)
        return s;
    s = upload_queries(n_queries, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    const size_t ncls = classes_.size(), total = (size_t)n_queries * ncls;
    DA_HIP(err, vote_out_.reserve(total * sizeof(T)));
    s = classify(n_queries, search_mode, vote_out_.as<T>(), nullptr);
    if (s != da_status_success && s != da_status_outlier_warning)
        return s;
    hipStream_t st = gpu->stream();
    // n_queries x n_classes in the handle's storage order (:1150-1156).
    if (order() == row_major) {
        DA_HIP(err, hipMemcpyAsync(proba, vote_out_.ptr(), total * sizeof(T),
                                   hipMemcpyDeviceToHost, st));
    } else {
        DA_HIP(err, q_stage_.reserve(total * sizeof(T)));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(vote_out_.as<T>(), ncls, ncls, n_queries,
                                                   q_stage_.as<T>(), st));
        DA_HIP(err, hipMemcpyAsync(proba, q_stage_.ptr(), total * sizeof(T),
                                   hipMemcpyDeviceToHost, st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    return s;
}

template <typename T>
da_status NNHandle<T>::predict(da_int n_queries, da_int n_features, const T *X_test,
                               da_int ldx_test, da_int *y_test, da_int search_mode) {
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    if (!has_labels_)
        return err->error(da_status_no_data, "No classification labels have been set. Please "
                                             "call da_nn_set_labels_s or da_nn_set_labels_d.");
    if (y_test == nullptr)
        return err->error(da_status_invalid_pointer, "y_test is not a valid pointer.");
    if (n_queries < 1)
        return err->error(da_status_invalid_array_dimension, "n_queries must be greater than 0.");
    s = check_queries(n_queries, n_features, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    s = available_classes();
    if (s != da_status_success)
        return s;
    s = upload_queries(n_queries, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    DA_HIP(err, vote_out_.reserve((size_t)n_queries * classes_.size() * sizeof(T)));
    DA_HIP(err, vote_pred_.reserve((size_t)n_queries * sizeof(int)));
    s = classify(n_queries, search_mode, vote_out_.as<T>(), vote_pred_.as<int>());
    if (s != da_status_success && s != da_status_outlier_warning)
        return s;
    DA_HIP(err, hipMemcpyAsync(y_test, vote_pred_.ptr(), (size_t)n_queries * sizeof(int),
                               hipMemcpyDeviceToHost, gpu->stream()));
    DA_HIP(err, hipStreamSynchronize(gpu->stream()));
    return s;
}

template <typename T>
da_status NNHandle<T>::predict(da_int n_queries, da_int n_features, const T *X_test,
                               da_int ldx_test, T *y_test, da_int search_mode) {
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    if (!has_targets_)
        return err->error(da_status_no_data, "No regression targets have been set. Please call "
                                             "da_nn_set_targets_s or da_nn_set_targets_d.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (y_test == nullptr)
        return err->error(da_status_invalid_pointer, "y_test is not a valid pointer.");
    if (n_queries < 1)
        return err->error(da_status_invalid_array_dimension, "n_queries must be greater than 0.");
    s = check_queries(n_queries, n_features, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    if ((da_int)y_reg_.size() != n_samples_)
        return err->error(da_status_invalid_array_dimension,
                          "The number of targets doesn't match the training data size " +
                              std::to_string(n_samples_) + ".");
    const int weights = (int)opts.skey("weights");
    if (metric_ == da_inner_product && weights == 1)
        return err->error(da_status_incompatible_options,
                          "The inner product metric is not compatible with weights=\"distance\". "
                          "Use weights=\"uniform\" instead.");
    s = upload_queries(n_queries, X_test, ldx_test);
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const int big = std::numeric_limits<int>::max();
    DA_HIP(err, vote_flag_.reserve(sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(vote_flag_.ptr(), &big, sizeof(int), hipMemcpyHostToDevice, st));
    DA_HIP(err, vote_out_.reserve((size_t)n_queries * sizeof(T)));
    int *flag = vote_flag_.as<int>();
    T *out = vote_out_.as<T>();
    const int mode = (int)opts.skey("outlier handling");
    T fallback = 0;
    if (search_mode == knn_search_mode) {
        const da_int k = n_neighbors_;
        if (k > n_samples_)
            return err->error(da_status_invalid_input,
                              "The number of neighbors (" + std::to_string(k) +
                                  ") must not exceed the number of training samples.");
        DA_HIP(err, out_ind_.reserve((size_t)n_queries * k * sizeof(da_int)));
        DA_HIP(err, out_dist_.reserve((size_t)n_queries * k * sizeof(T)));
        s = search(q_.as<T>(), n_queries, k, weights == 1, out_ind_.as<da_int>(),
                   weights == 1 ? out_dist_.as<T>() : nullptr);
        if (s != da_status_success)
            return s;
        KernelScope ks(*gpu, "nn_vote");
        DA_HIP(err, launch_nn_regress<T>(n_queries, nullptr, k, out_ind_.as<int>(),
                                         out_dist_.as<T>(), weights, tgt_.as<T>(), 0, (T)0, out,
                                         0, flag, st));
    } else if (search_mode == radius_search_mode) {
        if (metric_ == da_inner_product)
            return err->error(da_status_incompatible_options,
                              "The inner product metric is not compatible with radius "
                              "neighbors.");
        if (mode == 1) {
            fallback = (T)opts.rval("outlier target");
        } else if (mode == 2) {
            // Mean of the training targets (basic_statistics mean, summed in T).
            T sum = 0;
            for (T v : y_reg_)
                sum += v;
            fallback = sum / (T)n_samples_;
        }
        RadiusSink sink = [&](da_int q0, int rows, const int *off, const int *ind,
                              const T *dist) -> da_status {
            KernelScope ks(*gpu, "nn_vote");
            DA_HIP(err, launch_nn_regress<T>(rows, off, 0, ind, dist, weights, tgt_.as<T>(), mode,
                                             fallback, out + q0, (int)q0, flag, st));
            return da_status_success;
        };
        s = radius_core(q_.as<T>(), n_queries, (T)opts.rval("radius"), weights == 1, false,
                        &sink);
        if (s != da_status_success)
            return s;
    } else {
        return err->error(da_status_invalid_input,
                          "Unknown search mode: " + std::to_string(search_mode) + ".");
    }
    int first = big;
    DA_HIP(err, hipMemcpyAsync(&first, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    if (first < n_queries && search_mode == radius_search_mode && mode == 0)
        return err->error(da_status_operation_failed,
                          "Failed to predict targets for query " + std::to_string(first) +
                              " since it does not have any neighbors within the specified "
                              "radius.");
    DA_HIP(err, hipMemcpyAsync(y_test, out, (size_t)n_queries * sizeof(T),
                               hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    if (first < n_queries && search_mode == radius_search_mode && mode == 1)
        return err->warn(da_status_outlier_warning,
                         "One or more samples have no neighbors within the specified radius. "
                         "Setting the corresponding targets to " +
                             std::to_string((double)fallback) + ".");
    return da_status_success;
}

// Radius results (nearest_neighbors.cpp get_result, :180-300 and :2284-2330).
// Stored fields, in the order of neighbors<T>::serialize (source/core/nearest_neighbors/
// nearest_neighbors.cpp:2329-2460). The search is brute force for every algorithm option, so
// the working algorithm is stored as brute force and no tree follows the arrays; a stored
// model that carries a tree after them loads without it. The training rows are stored
// column-major, then the labels and targets (empty when not set). Fields the reference keeps
// as caches of its prediction (most frequent label, mean target) are stored as not computed.
template <typename T> da_status NNHandle<T>::serialize(ModelBuffer &b) {
    const bool load = b.loading();
    da_order ord = order();
    da_int n = n_samples_, d = n_features_;
    bool has_labels = has_labels_, has_targets = has_targets_, up_to_date = trained_,
         classes_done = classes_ready_, sort_results = false, trained = trained_,
         rnn_dist = rnn_dist_;
    std::vector<da_int> classes = classes_;
    da_int n_neighbors = n_neighbors_, algo = algo_, working_algo = NN_BRUTE, metric = metric_;
    const bool squares =
        metric_ == da_euclidean || metric_ == da_euclidean_gemm ||
        (metric_ == da_minkowski && p_ == 2.0);
    da_metric internal_metric = (da_metric)metric_;
    if (squares)
        internal_metric = metric_ == da_euclidean_gemm ? da_sqeuclidean_gemm : da_sqeuclidean;
    da_int leaf_size = leaf_size_;
    bool get_squares = squares;
    T p = (T)p_;
    da_int weights = opts.skey("weights"), outlier = opts.skey("outlier handling");
    da_int n_classes = classes_ready_ ? (da_int)classes_.size() : -1;
    T radius = (T)opts.rval("radius");
    std::vector<da_int> rcount;
    std::vector<std::vector<da_int>> rind;
    std::vector<std::vector<T>> rdist;
    if (!load && rnn_done_) {
        rcount = rnn_count_;
        size_t off = 0;
        for (da_int c : rnn_count_) {
            rind.emplace_back(rnn_ind_.begin() + off, rnn_ind_.begin() + off + c);
            if (rnn_dist_)
                rdist.emplace_back(rnn_d_.begin() + off, rnn_d_.begin() + off + c);
            off += (size_t)c;
        }
    }
    bool x_trained = trained_, mfl_done = false, manual_checked = false, mean_done = false;
    da_int mfl = -1, manual_index = -1, manual_label = opts.ival("outlier label");
    T mean_target = 0, manual_target = (T)opts.rval("outlier target");
    da_status s = da_status_success;
    auto io = [&](auto &v) {
        if (s == da_status_success)
            s = b.io(v);
    };
    io(ord);
    io(n);
    io(d);
    io(has_labels);
    io(has_targets);
    io(up_to_date);
    io(classes_done);
    io(classes);
    io(sort_results);
    io(trained);
    io(rnn_dist);
    io(n_neighbors);
    io(algo);
    io(working_algo);
    io(metric);
    io(internal_metric);
    io(leaf_size);
    io(get_squares);
    io(p);
    io(weights);
    io(outlier);
    io(n_classes);
    io(radius);
    io(rcount);
    io(rind);
    io(rdist);
    io(x_trained);
    io(mfl_done);
    io(mfl);
    io(manual_checked);
    io(manual_index);
    io(manual_label);
    io(mean_done);
    io(mean_target);
    io(manual_target);
    if (s != da_status_success)
        return s;
    std::vector<T> X, yreg;
    std::vector<da_int> ycls;
    if (!load) {
        if (trained_) {
            std::vector<T> rows((size_t)n * d);
            DA_HIP(err, hipMemcpyAsync(rows.data(), train_.ptr(), rows.size() * sizeof(T),
                                       hipMemcpyDeviceToHost, gpu->stream()));
            DA_HIP(err, hipStreamSynchronize(gpu->stream()));
            X.resize(rows.size());
            for (da_int i = 0; i < n; i++)
                for (da_int l = 0; l < d; l++)
                    X[(size_t)l * n + i] = rows[(size_t)i * d + l];
        }
        if (has_labels_)
            ycls = y_class_;
        if (has_targets_)
            yreg = y_reg_;
    }
    io(X);
    io(ycls);
    io(yreg);
    if (s != da_status_success || !load)
        return s;
    if (n < 0 || d < 0 || (int64_t)X.size() != (x_trained ? (int64_t)n * d : 0) ||
        (has_labels && (da_int)ycls.size() != n) || (has_targets && (da_int)yreg.size() != n) ||
        rind.size() != rcount.size() || (rnn_dist && rdist.size() != rcount.size()))
        return da_status_invalid_file_data;
    n_samples_ = n;
    n_features_ = d;
    trained_ = false;
    stored_X_ = std::move(X);
    has_labels_ = has_labels;
    has_targets_ = has_targets;
    y_class_ = std::move(ycls);
    y_reg_ = std::move(yreg);
    classes_ready_ = false;
    rnn_done_ = !rcount.empty();
    rnn_dist_ = rnn_dist;
    rnn_count_ = rcount;
    rnn_ind_.clear();
    rnn_d_.clear();
    for (size_t q = 0; q < rcount.size(); q++) {
        if ((da_int)rind[q].size() != rcount[q] || (rnn_dist && (da_int)rdist[q].size() != rcount[q]))
            return da_status_invalid_file_data;
        rnn_ind_.insert(rnn_ind_.end(), rind[q].begin(), rind[q].end());
        if (rnn_dist)
            rnn_d_.insert(rnn_d_.end(), rdist[q].begin(), rdist[q].end());
    }
    model_trained = x_trained;
    return da_status_success;
}

template <typename T> da_status NNHandle<T>::save_check() {
    if (!trained_)
        return err->error(da_status_no_data, "No training data have been set. Please call "
                                             "da_nn_set_data_s or da_nn_set_data_d.");
    return da_status_success;
}

// A loaded model searches again: the training rows go back to GPU memory through set_data
// (column-major, as stored), with the labels and targets after them.
template <typename T> da_status NNHandle<T>::restore() {
    if (stored_X_.empty())
        return da_status_success;
    std::string ord;
    (void)opts.get_string("storage order", ord);
    (void)opts.set_string("storage order", "column-major");
    std::vector<da_int> ycls = y_class_;
    std::vector<T> yreg = y_reg_;
    const bool lab = has_labels_, tgt = has_targets_;
    da_status s = set_data(n_samples_, n_features_, stored_X_.data(), n_samples_);
    (void)opts.set_string("storage order", ord);
    if (s != da_status_success)
        return s;
    stored_X_.clear();
    stored_X_.shrink_to_fit();
    if (lab)
        s = set_labels(n_samples_, ycls.data());
    if (s == da_status_success && tgt)
        s = set_targets(n_samples_, yreg.data());
    return s;
}

template <typename T>
da_status NNHandle<T>::get_result(da_result query, da_int *dim, da_int *result) {
    switch (query) {
    case da_nn_radius_neighbors_count:
    case da_nn_radius_neighbors_offsets:
    case da_nn_radius_neighbors_indices:
    case da_nn_radius_neighbors_indices_index:
        break;
    default:
        return AlgHandle::get_result(query, dim, result);
    }
    if (!rnn_done_)
        return err->warn(da_status_no_data, "Radius neighbors have not been computed. Please "
                                            "call da_nn_radius_neighbors first.");
    const da_int nq = (da_int)rnn_count_.size();
    auto too_small = [&](da_int need) {
        *dim = need;
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(need) + ".");
    };
    const da_int total = (da_int)rnn_ind_.size();
    switch (query) {
    case da_nn_radius_neighbors_count: {
        if (*dim < nq + 1)
            return too_small(nq + 1);
        for (da_int i = 0; i < nq; i++)
            result[i] = rnn_count_[i];
        result[nq] = total;
        return da_status_success;
    }
    case da_nn_radius_neighbors_offsets: {
        if (*dim < nq + 1)
            return too_small(nq + 1);
        da_int off = 0;
        for (da_int i = 0; i < nq; i++) {
            result[i] = rnn_count_[i] > 0 ? off : -1;
            off += rnn_count_[i];
        }
        result[nq] = off;
        return da_status_success;
    }
    case da_nn_radius_neighbors_indices: {
        if (*dim < total)
            return too_small(total);
        std::copy(rnn_ind_.begin(), rnn_ind_.end(), result);
        return da_status_success;
    }
    default: { // indices of one query, its index passed in result[0]
        const da_int q = result[0];
        if (q < 0 || q >= nq)
            return err->warn(da_status_invalid_input,
                             "The provided index is out of bounds. It should be in the range "
                             "[0, " + std::to_string(nq - 1) + "].");
        if (*dim < rnn_count_[q])
            return too_small(rnn_count_[q]);
        da_int off = 0;
        for (da_int i = 0; i < q; i++)
            off += rnn_count_[i];
        std::copy(rnn_ind_.begin() + off, rnn_ind_.begin() + off + rnn_count_[q], result);
        return da_status_success;
    }
    }
}

template <typename T>
da_status NNHandle<T>::get_result(da_result query, da_int *dim, T *result) {
    if (query != da_nn_radius_neighbors_distances &&
        query != da_nn_radius_neighbors_distances_index)
        return AlgHandle::get_result(query, dim, result);
    if (!rnn_done_)
        return err->warn(da_status_no_data, "Radius neighbors have not been computed. Please "
                                            "call da_nn_radius_neighbors_s or "
                                            "da_nn_radius_neighbors_d first.");
    if (!rnn_dist_)
        return err->warn(da_status_no_data,
                         "Distances were not requested during radius neighbors computation. "
                         "Please set return_distance to true in da_nn_radius_neighbors_s or "
                         "da_nn_radius_neighbors_d call.");
    const da_int nq = (da_int)rnn_count_.size();
    if (query == da_nn_radius_neighbors_distances) {
        const da_int total = (da_int)rnn_d_.size();
        if (*dim < total) {
            *dim = total;
            return err->warn(da_status_invalid_array_dimension,
                             "The array is too small. Please provide an array of at least "
                             "size: " + std::to_string(total) + ".");
        }
        std::copy(rnn_d_.begin(), rnn_d_.end(), result);
        return da_status_success;
    }
    const da_int q = (da_int)result[0];
    if (q < 0 || q >= nq)
        return err->warn(da_status_invalid_input,
                         "The provided index is out of bounds. It should be in the range [0, " +
                             std::to_string(nq - 1) + "].");
    if (*dim < rnn_count_[q]) {
        *dim = rnn_count_[q];
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(rnn_count_[q]) + ".");
    }
    da_int off = 0;
    for (da_int i = 0; i < q; i++)
        off += rnn_count_[i];
    std::copy(rnn_d_.begin() + off, rnn_d_.begin() + off + rnn_count_[q], result);
    return da_status_success;
}

template class NNHandle<float>;
template class NNHandle<double>;

std::unique_ptr<AlgHandle> make_nn_handle(da_precision p, ErrorStack &err, GpuContext &gpu) {
    if (p == da_double)
        return std::make_unique<NNHandle<double>>(err, gpu);
    return std::make_unique<NNHandle<float>>(err, gpu);
}

} // namespace dagpu
