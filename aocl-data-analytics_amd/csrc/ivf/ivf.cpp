#include "ivf.hpp"
#include "../core/factories.hpp"
#include "../kernels/launch.hpp"
#include "../kmeans/kmeans.hpp"
#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>

namespace dagpu {

// approx_nn_algorithm and approx_nn_metric (approximate_neighbors_types.hpp:29-30).
enum { ANN_IVFFLAT = 0, ANN_AUTO = 1 };
enum { ANN_EUCLIDEAN = 0, ANN_SQEUCLIDEAN = 1, ANN_INNER = 2, ANN_COSINE = 3 };

template <typename T>
ApproxNNHandle<T>::ApproxNNHandle(ErrorStack &e, GpuContext &g) : AlgHandle(e, g) {
    // Options of approximate_neighbors_options.hpp:38-103.
    const da_int imax = std::numeric_limits<da_int>::max();
    opts.add_int("number of neighbors", "Number of neighbors considered for k-nearest neighbors.",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 5);
    opts.add_int("n_list", "Number of lists to construct for inverted file indices", 1,
                 LowerBound::greater_equal, imax, UpperBound::p_inf, 1);
    opts.add_int("n_probe", "Number of lists to probe at search time for inverted file indices",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 1);
    opts.add_int("k-means_iter", "Maximum number of k-means iterations to perform at train time",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 10);
    opts.add_int("seed",
                 "Seed for random number generation; set to -1 for non-deterministic results.",
                 -1, LowerBound::greater_equal, imax, UpperBound::p_inf, 0);
    opts.add_real("train fraction", "Fraction of training data to use for k-means clustering.",
                  0.0, LowerBound::greater_than, 1.0, UpperBound::less_equal, 1.0);
    opts.add_string("algorithm", "Algorithm used to compute the approximate nearest neighbors.",
                    {{"auto", ANN_AUTO}, {"ivfflat", ANN_IVFFLAT}}, "ivfflat");
    opts.add_string("metric", "Metric used to compute distances.",
                    {{"euclidean", ANN_EUCLIDEAN},
                     {"sqeuclidean", ANN_SQEUCLIDEAN},
                     {"inner product", ANN_INNER},
                     {"cosine", ANN_COSINE}},
                    "sqeuclidean");
}

template <typename T>
da_status ApproxNNHandle<T>::set_training_data(da_int n_samples, da_int n_features, const T *X,
                                               da_int ldx) {
    // approximate_neighbors.cpp:187-231: the pointer is kept until training.
    data_set_ = false;
    model_trained = false;
    data_added_ = false;
    n_index_ = 0;
    order_ = order();
    da_status s = check_2D_array<T>(order_, n_samples, n_features, X, ldx, "n_samples",
                                    "n_features", "X_train", "ldx_train");
    if (s != da_status_success)
        return s;
    X_train_ = X;
    ldx_train_ = ldx;
    n_samples_ = n_samples;
    n_features_ = n_features;
    data_set_ = true;
    return da_status_success;
}

template <typename T> da_status ApproxNNHandle<T>::check_options_update() {
    // approximate_neighbors.cpp:152-185: n_list, algorithm and metric are fixed by train().
    if (opts.ival("n_list") != n_list_)
        return err->error(da_status_option_locked, "n_list cannot be changed after calling train().");
    if (opts.skey("algorithm") != algo_)
        return err->error(da_status_option_locked,
                          "algorithm cannot be changed after calling train().");
    if (opts.skey("metric") != metric_)
        return err->error(da_status_option_locked, "metric cannot be changed after calling train().");
    return da_status_success;
}

template <typename T> da_status ApproxNNHandle<T>::train() {
    if (!data_set_)
        return err->error(da_status_no_data,
                          "No data has been passed to the handle. Please call "
                          "da_approx_nn_set_training_data_s or da_approx_nn_set_training_data_d.");
    // read_training_options (approximate_neighbors.cpp:233-274).
    n_list_ = opts.ival("n_list");
    const da_int max_iter = opts.ival("k-means_iter");
    da_int seed = opts.ival("seed");
    const T frac = (T)opts.rval("train fraction");
    algo_ = opts.skey("algorithm");
    metric_ = opts.skey("metric");
    if (n_list_ > n_samples_)
        return err->error(da_status_invalid_array_dimension,
                          "n_samples = " + std::to_string(n_samples_) +
                              " must be at least as large as n_list = " + std::to_string(n_list_));
    da_status s = gpu->init(err);
    if (s != da_status_success)
        return s;
    // train() (:492-523): the engine is seeded once per training.
    if (seed == -1) {
        std::random_device r;
        seed = std::abs((da_int)r());
    }
    engine_.seed((std::mt19937::result_type)seed);
    model_trained = false;
    data_added_ = false;
    n_index_ = 0;
    // Training rows: all of them, or a selection sample (Algorithm S, da_std.hpp:102-118) of
    // max(n * fraction, n_list) rows in ascending order (subsample_training_data, :276-332).
    const T *rows = X_train_;
    da_int ld = ldx_train_, n_train = n_samples_;
    da_order ord = order_;
    std::vector<T> sub;
    if (frac < (T)1) {
        n_train = std::max((da_int)(n_samples_ * frac), n_list_);
        std::vector<da_int> pick;
        pick.reserve(n_train);
        da_int need = n_train, left = n_samples_;
        for (da_int i = 0; need > 0; i++, left--) {
            std::uniform_int_distribution<da_int> u(0, left - 1);
            if (u(engine_) < need) {
                pick.push_back(i);
                need--;
            }
        }
        sub.resize((size_t)n_train * n_features_);
        for (da_int r = 0; r < n_train; r++)
            for (da_int j = 0; j < n_features_; j++)
                sub[(size_t)r * n_features_ + j] =
                    order_ == row_major ? X_train_[(size_t)pick[r] * ldx_train_ + j]
                                        : X_train_[pick[r] + (size_t)j * ldx_train_];
        rows = sub.data();
        ld = n_features_;
        ord = row_major;
    }
    // The cosine metric trains on normalised rows (:419-426), as the GPU normalises the rows
    // it adds: 1/|x| in double, the product rounded to T, zero rows unchanged.
    std::vector<T> unit;
    if (metric_ == ANN_COSINE) {
        const size_t d = n_features_;
        unit.resize((size_t)n_train * d);
        for (da_int i = 0; i < n_train; i++) {
            double ss = 0.0;
            for (size_t j = 0; j < d; j++) {
                const double v = (double)(ord == row_major ? rows[(size_t)i * ld + j]
                                                           : rows[i + j * (size_t)ld]);
                unit[i * d + j] = (T)v;
                ss += v * v;
            }
            const double nrm = std::sqrt(ss);
            if (nrm > 0.0)
                for (size_t j = 0; j < d; j++)
                    unit[i * d + j] = (T)((double)unit[i * d + j] * (1.0 / nrm));
        }
        rows = unit.data();
        ld = n_features_;
        ord = row_major;
    }
    // Lloyd k-means on the GPU with centres drawn from the rows (train_ivfflat, :334-490):
    // one initialisation, tolerance sqrt(epsilon); spherical k-means without normalising the
    // data for the inner-product and cosine metrics (:458-461).
    KMeansHandle<T> km(*err, *gpu);
    if (metric_ == ANN_INNER || metric_ == ANN_COSINE) {
        km.opts.set_string("distance", "cosine");
        km.opts.set_string("normalize data", "no");
    }
    km.opts.set_string("storage order", ord == row_major ? "row-major" : "column-major");
    km.opts.set_string("initialization method", "random");
    km.opts.set_string("algorithm", "lloyd");
    km.opts.set_int("n_clusters", n_list_);
    km.opts.set_int("n_init", 1);
    km.opts.set_int("max_iter", max_iter);
    km.opts.set_int("seed", seed);
    km.opts.set_real("convergence tolerance", std::sqrt((double)std::numeric_limits<T>::epsilon()));
    s = km.set_data(n_train, n_features_, rows, ld);
    if (s != da_status_success)
        return s;
    s = km.compute();
    if (s != da_status_success && s != da_status_maxit)
        return s;
    // Centres come back in the k-means handle's storage order.
    std::vector<T> c((size_t)n_list_ * n_features_);
    da_int dim = (da_int)c.size();
    s = km.get_result(da_kmeans_cluster_centres, &dim, c.data());
    if (s != da_status_success)
        return s;
    centroids_.resize(c.size());
    for (da_int i = 0; i < n_list_; i++)
        for (da_int j = 0; j < n_features_; j++)
            centroids_[(size_t)i * n_features_ + j] =
                ord == row_major ? c[(size_t)i * n_features_ + j] : c[i + (size_t)j * n_list_];
    T info[6];
    dim = 6;
    s = km.get_result(da_rinfo, &dim, info);
    if (s != da_status_success)
        return s;
    kmeans_iter_ = (da_int)info[3];
    hipStream_t st = gpu->stream();
    DA_HIP(err, C_d_.reserve(centroids_.size() * sizeof(T)));
    DA_HIP(err, cdouble_d_.reserve(centroids_.size() * sizeof(double)));
    DA_HIP(err, hipMemcpyAsync(C_d_.ptr(), centroids_.data(), centroids_.size() * sizeof(T),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, launch_convert<T>(C_d_.as<T>(), (int64_t)centroids_.size(),
                                  cdouble_d_.as<double>(), st));
    DA_HIP(err, hipStreamSynchronize(st));
    list_sizes_.assign(n_list_, 0);
    lists_.assign(n_list_, {});
    err->clear();
    model_trained = true;
    return da_status_success;
}

// Appends n rows (handle storage order, leading dimension ld) to dst as row-major rows from
// row offset_rows on, keeping the rows already there.
template <typename T>
da_status ApproxNNHandle<T>::upload_rows(const T *X, da_int n, da_int ld, DeviceBuffer &dst,
                                         size_t offset_rows) {
    hipStream_t st = gpu->stream();
    const size_t d = n_features_, need = (offset_rows + n) * d * sizeof(T);
    if (dst.size() < need) {
        DA_HIP(err, X_spare_.reserve(std::max(need, dst.size() * 2)));
        if (offset_rows > 0)
            DA_HIP(err, hipMemcpyAsync(X_spare_.ptr(), dst.ptr(), offset_rows * d * sizeof(T),
                                       hipMemcpyDeviceToDevice, st));
        DA_HIP(err, hipStreamSynchronize(st));
        dst.swap(X_spare_);
        X_spare_.release();
    }
    T *out = dst.as<T>() + offset_rows * d;
    if (order_ == row_major) {
        DA_HIP(err, copy_2d(out, d * sizeof(T), X, (size_t)ld * sizeof(T), d * sizeof(T), n,
                            hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, q_stage_.reserve((size_t)n * d * sizeof(T)));
        DA_HIP(err, copy_2d(q_stage_.ptr(), n * sizeof(T), X, (size_t)ld * sizeof(T),
                            n * sizeof(T), d, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(q_stage_.as<T>(), n, n, (int)d, out, st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T>
da_status ApproxNNHandle<T>::add(da_int n_add, da_int n_features, const T *X, da_int ldx) {
    // add / add_ivfflat (approximate_neighbors.cpp:525-842): every vector joins the list of
    // its nearest centroid (exact distances, ties to the lower centroid), after the vectors
    // already there, in order of addition.
    if (!model_trained)
        return err->error(da_status_no_data, "No index has been trained. Please call "
                                             "da_approx_nn_train_s or da_approx_nn_train_d.");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    s = check_2D_array<T>(order_, n_add, n_features, X, ldx, "n_samples", "n_features", "X_add",
                          "ldx_add");
    if (s != da_status_success)
        return s;
    if (n_features != n_features_)
        return err->error(da_status_invalid_input,
                          "The function was called with n_features = " +
                              std::to_string(n_features) +
                              " but the index has been trained with " +
                              std::to_string(n_features_) + " features.");
    s = upload_rows(X, n_add, ldx, X_d_, (size_t)n_index_);
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    T *added = X_d_.as<T>() + (size_t)n_index_ * n_features_;
    // Cosine: the vectors are stored normalised (add_ivfflat, :577-598).
    if (metric_ == ANN_COSINE)
        DA_HIP(err, launch_ivf_normalize<T>(added, n_add, (int)n_features_, st));
    DA_HIP(err, lab_d_.reserve((size_t)n_add * sizeof(int)));
    {
        // Nearest centroid: squared distance, or the largest x.c for the inner-product and
        // cosine metrics (the GEMM with alpha = -1 of :715-731), ties to the lower centroid.
        // Rows of up to 128 features: the k-means assignment on the matrix cores in the data's
        // precision with the exact re-check of near ties (the same labels as the exact kernel).
        KernelScope ks(*gpu, "ivf_assign");
        const int sph = metric_ == ANN_INNER || metric_ == ANN_COSINE;
        const int d = (int)n_features_;
        if (d <= 128) {
            int kp = (d + 15) / 16 * 16;
            if (kp == 80 || kp == 112)
                kp = kp == 80 ? 96 : 128;
            DA_HIP(err, as_x_.reserve((size_t)n_add * kp * sizeof(T)));
            DA_HIP(err, as_c_.reserve((size_t)(n_list_ + 16) * kp * sizeof(T)));
            DA_HIP(err, as_h_.reserve((size_t)(n_list_ + 16) * sizeof(T)));
            DA_HIP(err, as_small_.reserve(4 * sizeof(double)));
            DA_HIP(err, as_list_.reserve((size_t)n_add * sizeof(int)));
            DA_HIP(err, as_ct_.reserve((size_t)n_list_ * d * sizeof(double)));
            double *cmax2 = as_small_.as<double>();
            int *cnt = reinterpret_cast<int *>(as_small_.as<double>() + 1); // changed, list count
            DA_HIP(err, hipMemsetAsync(as_small_.ptr(), 0, 4 * sizeof(double), st));
            DA_HIP(err, (launch_pack_rows<T, T>(added, d, n_add, d, as_x_.as<T>(), kp, n_add, st)));
            DA_HIP(err, launch_centre_prep<T>(cdouble_d_.as<double>(), n_list_, d, kp,
                                              as_c_.as<T>(), as_h_.as<T>(), cmax2, st, sph));
            double hc = 0;
            DA_HIP(err, hipMemcpyAsync(&hc, cmax2, sizeof(double), hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipStreamSynchronize(st));
            DA_HIP(err, launch_kmeans_assign_mfma<T>(as_x_.as<T>(), n_add, kp, as_c_.as<T>(),
                                                     as_h_.as<T>(), n_list_, std::sqrt(hc),
                                                     lab_d_.as<int>(), nullptr, cnt, cnt + 1,
                                                     as_list_.as<int>(), st));
            DA_HIP(err, launch_kmeans_recheck<T>(added, d, d, cdouble_d_.as<double>(), n_list_,
                                                 cnt + 1, as_list_.as<int>(), lab_d_.as<int>(),
                                                 nullptr, cnt, as_ct_.as<double>(), st, sph));
        } else {
            DA_HIP(err, launch_kmeans_assign<T>(added, n_add, d, cdouble_d_.as<double>(), n_list_,
                                                lab_d_.as<int>(), nullptr, nullptr, nullptr, st,
                                                sph));
        }
    }
    std::vector<int> lab(n_add);
    DA_HIP(err, hipMemcpyAsync(lab.data(), lab_d_.ptr(), (size_t)n_add * sizeof(int),
                               hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    for (da_int i = 0; i < n_add; i++) {
        lists_[lab[i]].push_back((int)(n_index_ + i));
        list_sizes_[lab[i]]++;
    }
    n_index_ += n_add;
    std::vector<int> perm, off(n_list_ + 1, 0);
    perm.reserve(n_index_);
    for (da_int c = 0; c < n_list_; c++) {
        perm.insert(perm.end(), lists_[c].begin(), lists_[c].end());
        off[c + 1] = (int)perm.size();
    }
    DA_HIP(err, perm_d_.reserve(perm.size() * sizeof(int)));
    DA_HIP(err, off_d_.reserve(off.size() * sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(perm_d_.ptr(), perm.data(), perm.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipMemcpyAsync(off_d_.ptr(), off.data(), off.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    data_added_ = true;
    return da_status_success;
}

template <typename T> da_status ApproxNNHandle<T>::train_and_add() {
    da_status s = train();
    if (s != da_status_success)
        return s;
    return add(n_samples_, n_features_, X_train_, ldx_train_);
}

template <typename T>
da_status ApproxNNHandle<T>::kneighbors(da_int n_queries, da_int n_features, const T *X_test,
                                        da_int ldx_test, da_int *n_ind, T *n_dist, da_int k,
                                        da_int return_distance) {
    // Checks in the order of approximate_neighbors.cpp:858-942.
    if (!model_trained)
        return err->error(da_status_no_data, "No index has been trained. Please call "
                                             "da_approx_nn_train_s or da_approx_nn_train_d.");
    if (!data_added_)
        return err->error(da_status_no_data, "No data has been added. Please call "
                                             "da_approx_nn_add_s or da_approx_nn_add_d");
    da_status s = check_options_update();
    if (s != da_status_success)
        return s;
    if (k <= 0)
        k = opts.ival("number of neighbors");
    if (k > n_index_)
        return err->error(da_status_invalid_input,
                          std::to_string(k) + " neighbors were requested but only " +
                              std::to_string(n_index_) + " samples have been added to the index");
    const da_int n_probe = opts.ival("n_probe");
    if (n_probe > n_list_)
        return err->error(da_status_invalid_input,
                          "n_probe=" + std::to_string(n_probe) +
                              " must be no larger than n_list=" + std::to_string(n_list_));
    if (n_ind == nullptr)
        return err->error(da_status_invalid_pointer, "n_ind is not a valid pointer.");
    if (return_distance && n_dist == nullptr)
        return err->error(da_status_invalid_pointer, "n_dist is not a valid pointer.");
    s = check_2D_array<T>(order_, n_queries, n_features, X_test, ldx_test, "n_samples",
                          "n_features", "X_test", "ldx_test");
    if (s != da_status_success)
        return s;
    if (n_features != n_features_)
        return err->error(da_status_invalid_input,
                          "The function was called with n_features = " +
                              std::to_string(n_features) +
                              " but the index has been trained with " +
                              std::to_string(n_features_) + " features.");
    if (k > 3840)
        return err->error(da_status_not_implemented,
                          "This build returns at most 3840 approximate neighbors per query.");
    hipStream_t st = gpu->stream();
    s = upload_rows(X_test, n_queries, ldx_test, q_d_, 0);
    if (s != da_status_success)
        return s;
    const int d = n_features_;
    const bool ip = metric_ == ANN_INNER || metric_ == ANN_COSINE;
    // Cosine: queries normalised like the stored vectors (:1068-1081).
    if (metric_ == ANN_COSINE)
        DA_HIP(err, launch_ivf_normalize<T>(q_d_.as<T>(), n_queries, d, st));
    const da_int max_list = *std::max_element(list_sizes_.begin(), list_sizes_.end());
    // Query blocks whose candidate rows (n_probe lists of at most max_list vectors) and
    // coarse keys stay within 2^26 entries.
    const int64_t row_cap = std::max<int64_t>({(int64_t)n_probe * max_list, (int64_t)k,
                                               (int64_t)n_list_});
    const da_int qb = (da_int)std::max<int64_t>(
        1, std::min<int64_t>(n_queries, ((int64_t)1 << 26) / row_cap));
    const size_t ob = (size_t)n_queries * k;
    DA_HIP(err, out_ind_.reserve(ob * sizeof(int)));
    DA_HIP(err, out_dist_.reserve(ob * sizeof(T)));
    DA_HIP(err, ckeys_.reserve((size_t)qb * n_list_ * sizeof(T)));
    DA_HIP(err, probe_key_.reserve((size_t)qb * n_probe * sizeof(T)));
    DA_HIP(err, probe_d_.reserve((size_t)qb * n_probe * sizeof(int)));
    DA_HIP(err, seg_d_.reserve((size_t)qb * n_probe * sizeof(int)));
    DA_HIP(err, sel_key_.reserve((size_t)qb * k * sizeof(T)));
    DA_HIP(err, sel_idx_.reserve((size_t)qb * k * sizeof(int)));
    std::vector<int> probes((size_t)qb * n_probe), seg((size_t)qb * n_probe);
    const T *Q = q_d_.as<T>();
    for (da_int q0 = 0; q0 < n_queries; q0 += qb) {
        const int rows = (int)std::min<da_int>(qb, n_queries - q0);
        // Coarse step: the n_probe nearest centroids of each query, by (distance, index);
        // for the inner-product and cosine metrics the largest q.c (:1093-1127).
        {
            KernelScope ks(*gpu, "ivf_coarse");
            DA_HIP(err, (launch_pairwise_direct<T, T>(ip ? KM_INNER : KM_SQEUCLIDEAN,
                                                      Q + (size_t)q0 * d, C_d_.as<T>(), rows,
                                                      n_list_, d, ckeys_.as<T>(), n_list_,
                                                      nullptr, nullptr, 2.0, 0, 0, st)));
            DA_HIP(err, launch_row_topk<T>(ckeys_.as<T>(), n_list_, rows, n_list_, n_probe,
                                           ip ? 1 : 0, probe_key_.as<T>(), probe_d_.as<int>(),
                                           st));
        }
        DA_HIP(err, hipMemcpyAsync(probes.data(), probe_d_.ptr(),
                                   (size_t)rows * n_probe * sizeof(int), hipMemcpyDeviceToHost,
                                   st));
        DA_HIP(err, hipStreamSynchronize(st));
        // Candidate columns of each probed list within its query's row, the lists in
        // ascending list number: selection by (key, column) then keeps, among equal keys,
        // the vector of the lower list and the earlier position in it, the vector the
        // reference's search (lists in number order, strict < on its heap) keeps.
        int64_t ld = k;
        std::vector<int> by_list(n_probe);
        for (int r = 0; r < rows; r++) {
            const int *pr = probes.data() + (size_t)r * n_probe;
            for (da_int p = 0; p < n_probe; p++)
                by_list[p] = p;
            std::sort(by_list.begin(), by_list.end(),
                      [&](int a, int b) { return pr[a] < pr[b]; });
            int64_t pos = 0;
            for (da_int p = 0; p < n_probe; p++) {
                seg[(size_t)r * n_probe + by_list[p]] = (int)pos;
                pos += list_sizes_[pr[by_list[p]]];
            }
            ld = std::max(ld, pos);
        }
        DA_HIP(err, hipMemcpyAsync(seg_d_.ptr(), seg.data(), (size_t)rows * n_probe * sizeof(int),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, cand_key_.reserve((size_t)rows * ld * sizeof(T)));
        DA_HIP(err, cand_idx_.reserve((size_t)rows * ld * sizeof(int)));
        {
            KernelScope ks(*gpu, "ivf_scan");
            DA_HIP(err, launch_ivf_pad<T>(cand_key_.as<T>(), cand_idx_.as<int>(),
                                          (int64_t)rows * ld, st));
            DA_HIP(err, launch_ivf_scan<T>(Q + (size_t)q0 * d, rows, n_probe, d, X_d_.as<T>(),
                                           perm_d_.as<int>(), off_d_.as<int>(), probe_d_.as<int>(),
                                           seg_d_.as<int>(), ld, cand_key_.as<T>(),
                                           cand_idx_.as<int>(), st, ip ? 1 : 0));
        }
        {
            KernelScope ks(*gpu, "ivf_select");
            DA_HIP(err, launch_row_topk<T>(cand_key_.as<T>(), ld, rows, (int)ld, k, 0,
                                           sel_key_.as<T>(), sel_idx_.as<int>(), st));
            const int post = metric_ == ANN_EUCLIDEAN ? 1
                             : metric_ == ANN_COSINE  ? 2
                             : metric_ == ANN_INNER   ? 3
                                                      : 0;
            DA_HIP(err, launch_ivf_finish<T>(sel_key_.as<T>(), sel_idx_.as<int>(),
                                             cand_idx_.as<int>(), ld, k, (int64_t)rows * k, post,
                                             out_dist_.as<T>() + (size_t)q0 * k,
                                             out_ind_.as<int>() + (size_t)q0 * k, st));
        }
    }
    // n_queries x k in the handle's storage order (kneighbors_compute_ivfflat, :1317-1344).
    if (order_ == row_major) {
        DA_HIP(err, hipMemcpyAsync(n_ind, out_ind_.ptr(), ob * sizeof(int), hipMemcpyDeviceToHost,
                                   st));
        if (return_distance)
            DA_HIP(err, hipMemcpyAsync(n_dist, out_dist_.ptr(), ob * sizeof(T),
                                       hipMemcpyDeviceToHost, st));
    } else {
        DA_HIP(err, q_stage_.reserve(ob * std::max(sizeof(T), sizeof(int))));
        DA_HIP(err, launch_colmajor_to_rowmajor<float>(reinterpret_cast<const float *>(
                                                           out_ind_.ptr()),
                                                       k, k, n_queries, q_stage_.as<float>(), st));
        DA_HIP(err, hipMemcpyAsync(n_ind, q_stage_.ptr(), ob * sizeof(int), hipMemcpyDeviceToHost,
                                   st));
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

// Results (approximate_neighbors.cpp:62-150).
// Stored fields, in the order of approximate_neighbors<T>::serialize
// (source/core/approximate_neighbors/approximate_neighbors.cpp:1346-1399): the centroids in
// the handle's storage order, each list's vectors (row-major, normalised for cosine), global
// indices and, for the squared euclidean search, squared norms; then the training rows.
template <typename T> da_status ApproxNNHandle<T>::serialize(ModelBuffer &b) {
    const bool load = b.loading();
    bool data_set = data_set_, trained = model_trained, added = data_added_;
    da_int n_neighbors = opts.ival("number of neighbors"), algo = algo_, internal_algo = ANN_IVFFLAT;
    // The search metric: squared distances for "euclidean" (:269-270).
    da_int internal_metric = metric_ == ANN_EUCLIDEAN ? (da_int)ANN_SQEUCLIDEAN : metric_;
    da_int metric = metric_, seed = opts.ival("seed"), internal_seed = seed;
    da_int n_train = n_samples_;
    da_order ord = order_;
    da_int n = n_samples_, d = n_features_;
    T frac = (T)opts.rval("train fraction");
    da_int n_list = n_list_, n_probe = opts.ival("n_probe"), max_iter = opts.ival("k-means_iter"),
           n_init = 1;
    T tol = (T)std::sqrt((double)std::numeric_limits<T>::epsilon());
    std::vector<T> centroids, cnorms;
    da_int ld_c = ord == column_major ? n_list : d;
    da_int n_index = n_index_;
    std::vector<std::vector<T>> vecs, norms;
    std::vector<std::vector<da_int>> gidx;
    std::vector<da_int> sizes = list_sizes_, old_sizes = list_sizes_;
    if (!load && model_trained) {
        centroids.resize(centroids_.size());
        for (da_int c = 0; c < n_list; c++)
            for (da_int j = 0; j < d; j++)
                centroids[ord == column_major ? (size_t)j * n_list + c : (size_t)c * d + j] =
                    centroids_[(size_t)c * d + j];
        if (internal_metric == ANN_SQEUCLIDEAN) {
            cnorms.assign(n_list, (T)0);
            for (da_int c = 0; c < n_list; c++)
                for (da_int j = 0; j < d; j++)
                    cnorms[c] += centroids_[(size_t)c * d + j] * centroids_[(size_t)c * d + j];
        }
        std::vector<T> rows((size_t)n_index_ * d);
        if (n_index_ > 0) {
            DA_HIP(err, hipMemcpyAsync(rows.data(), X_d_.ptr(), rows.size() * sizeof(T),
                                       hipMemcpyDeviceToHost, gpu->stream()));
            DA_HIP(err, hipStreamSynchronize(gpu->stream()));
        }
        vecs.resize(n_list);
        norms.resize(n_list);
        gidx.resize(n_list);
        for (da_int c = 0; c < n_list; c++)
            for (int g : lists_[c]) {
                gidx[c].push_back(g);
                T nrm = 0;
                for (da_int j = 0; j < d; j++) {
                    const T v = rows[(size_t)g * d + j];
                    vecs[c].push_back(v);
                    nrm += v * v;
                }
                if (internal_metric == ANN_SQEUCLIDEAN)
                    norms[c].push_back(nrm);
            }
    }
    da_status s = da_status_success;
    auto io = [&](auto &v) {
        if (s == da_status_success)
            s = b.io(v);
    };
    io(data_set);
    io(trained);
    io(added);
    io(n_neighbors);
    io(algo);
    io(internal_algo);
    io(internal_metric);
    io(metric);
    io(internal_seed);
    io(seed);
    io(n_train);
    io(ord);
    io(n);
    io(d);
    io(frac);
    io(n_list);
    io(n_probe);
    io(max_iter);
    io(n_init);
    io(tol);
    io(centroids);
    io(ld_c);
    io(n_index);
    io(vecs);
    io(gidx);
    io(sizes);
    io(old_sizes);
    io(norms);
    io(cnorms);
    if (s != da_status_success)
        return s;
    std::vector<T> X;
    if (!load && data_set_ && X_train_ != nullptr) {
        X.resize((size_t)n * d);
        for (da_int i = 0; i < n; i++)
            for (da_int j = 0; j < d; j++) {
                const size_t src = order_ == row_major ? (size_t)i * ldx_train_ + j
                                                       : (size_t)j * ldx_train_ + i;
                const size_t dst = order_ == row_major ? (size_t)i * d + j : (size_t)j * n + i;
                X[dst] = X_train_[src];
            }
    }
    io(X);
    if (s != da_status_success || !load)
        return s;
    if (n < 0 || d <= 0 || n_list <= 0 || n_index < 0 ||
        (trained && (int64_t)centroids.size() != (int64_t)n_list * d) ||
        (added && ((da_int)vecs.size() != n_list || (da_int)gidx.size() != n_list)) ||
        (!X.empty() && (int64_t)X.size() != (int64_t)n * d))
        return da_status_invalid_file_data;
    data_set_ = data_set && !X.empty();
    model_trained = trained;
    data_added_ = added;
    order_ = ord;
    n_samples_ = n;
    n_features_ = d;
    n_list_ = n_list;
    algo_ = algo;
    metric_ = metric;
    kmeans_iter_ = 0;
    stored_train_ = std::move(X);
    X_train_ = data_set_ ? stored_train_.data() : nullptr;
    ldx_train_ = ord == row_major ? d : n;
    centroids_.assign((size_t)n_list * d, (T)0);
    if (trained)
        for (da_int c = 0; c < n_list; c++)
            for (da_int j = 0; j < d; j++)
                centroids_[(size_t)c * d + j] =
                    centroids[ord == column_major ? (size_t)j * n_list + c : (size_t)c * d + j];
    list_sizes_.assign(n_list, 0);
    lists_.assign(n_list, {});
    n_index_ = added ? n_index : 0;
    stored_rows_.assign((size_t)n_index_ * d, (T)0);
    std::vector<char> seen((size_t)n_index_, 0);
    if (added) {
        for (da_int c = 0; c < n_list; c++) {
            if (vecs[c].size() != gidx[c].size() * (size_t)d)
                return da_status_invalid_file_data;
            for (size_t i = 0; i < gidx[c].size(); i++) {
                const da_int g = gidx[c][i];
                if (g < 0 || g >= n_index_ || seen[g])
                    return da_status_invalid_file_data;
                seen[g] = 1;
                lists_[c].push_back((int)g);
                std::copy(vecs[c].begin() + i * d, vecs[c].begin() + (i + 1) * d,
                          stored_rows_.begin() + (size_t)g * d);
            }
            list_sizes_[c] = (da_int)gidx[c].size();
        }
        for (char f : seen)
            if (!f)
                return da_status_invalid_file_data;
    }
    return da_status_success;
}

template <typename T> da_status ApproxNNHandle<T>::save_check() {
    if (!data_set_ || !model_trained)
        return err->error(da_status_no_data,
                          "Index has not yet been trained. Please call da_approx_nn_train_d or "
                          "da_approx_nn_train_s before saving the model.");
    return da_status_success;
}

// A loaded index searches again: centroids, vectors (in order of addition) and the lists go
// back to GPU memory as train() and add() leave them.
template <typename T> da_status ApproxNNHandle<T>::restore() {
    if (!model_trained)
        return da_status_success;
    da_status s = gpu->init(err);
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    DA_HIP(err, C_d_.reserve(centroids_.size() * sizeof(T)));
    DA_HIP(err, cdouble_d_.reserve(centroids_.size() * sizeof(double)));
    DA_HIP(err, hipMemcpyAsync(C_d_.ptr(), centroids_.data(), centroids_.size() * sizeof(T),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, launch_convert<T>(C_d_.as<T>(), (int64_t)centroids_.size(),
                                  cdouble_d_.as<double>(), st));
    if (data_added_ && n_index_ > 0) {
        DA_HIP(err, X_d_.reserve(stored_rows_.size() * sizeof(T)));
        DA_HIP(err, hipMemcpyAsync(X_d_.ptr(), stored_rows_.data(),
                                   stored_rows_.size() * sizeof(T), hipMemcpyHostToDevice, st));
        std::vector<int> perm, off(n_list_ + 1, 0);
        perm.reserve(n_index_);
        for (da_int c = 0; c < n_list_; c++) {
            perm.insert(perm.end(), lists_[c].begin(), lists_[c].end());
            off[c + 1] = (int)perm.size();
        }
        DA_HIP(err, perm_d_.reserve(perm.size() * sizeof(int)));
        DA_HIP(err, off_d_.reserve(off.size() * sizeof(int)));
        DA_HIP(err, hipMemcpyAsync(perm_d_.ptr(), perm.data(), perm.size() * sizeof(int),
                                   hipMemcpyHostToDevice, st));
        DA_HIP(err, hipMemcpyAsync(off_d_.ptr(), off.data(), off.size() * sizeof(int),
                                   hipMemcpyHostToDevice, st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    stored_rows_.clear();
    stored_rows_.shrink_to_fit();
    return da_status_success;
}

template <typename T>
da_status ApproxNNHandle<T>::get_result(da_result query, da_int *dim, T *result) {
    if (!model_trained)
        return err->warn(da_status_no_data,
                         "Index has not yet been trained. Please call da_approx_nn_train_d or "
                         "da_approx_nn_train_s before extracting results.");
    auto too_small = [&](da_int need) {
        *dim = need;
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(need) + ".");
    };
    switch (query) {
    case da_rinfo:
        if (*dim < 4)
            return too_small(4);
        result[0] = (T)n_list_;
        result[1] = (T)n_index_;
        result[2] = (T)n_features_;
        result[3] = (T)kmeans_iter_;
        return da_status_success;
    case da_approx_nn_cluster_centroids:
        if (*dim < n_list_ * n_features_)
            return too_small(n_list_ * n_features_);
        for (da_int i = 0; i < n_list_; i++)
            for (da_int j = 0; j < n_features_; j++) {
                const T v = centroids_[(size_t)i * n_features_ + j];
                if (order_ == column_major)
                    result[i + (size_t)j * n_list_] = v;
                else
                    result[(size_t)i * n_features_ + j] = v;
            }
        return da_status_success;
    default:
        return err->warn(da_status_unknown_query, "The requested result could not be found.");
    }
}

template <typename T>
da_status ApproxNNHandle<T>::get_result(da_result query, da_int *dim, da_int *result) {
    da_status s = AlgHandle::get_result(query, dim, result);
    if (s != da_status_unknown_query)
        return s;
    if (!model_trained)
        return err->warn(da_status_no_data,
                         "Index has not yet been trained. Please call da_approx_nn_train_d or "
                         "da_approx_nn_train_s before extracting results.");
    if (query != da_approx_nn_list_sizes)
        return err->warn(da_status_unknown_query, "The requested result could not be found.");
    if (*dim < n_list_) {
        *dim = n_list_;
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(n_list_) + ".");
    }
    std::copy(list_sizes_.begin(), list_sizes_.end(), result);
    return da_status_success;
}

template class ApproxNNHandle<float>;
template class ApproxNNHandle<double>;

std::unique_ptr<AlgHandle> make_approx_nn_handle(da_precision p, ErrorStack &err,
                                                 GpuContext &gpu) {
    if (p == da_double)
        return std::make_unique<ApproxNNHandle<double>>(err, gpu);
    return std::make_unique<ApproxNNHandle<float>>(err, gpu);
}

} // namespace dagpu
