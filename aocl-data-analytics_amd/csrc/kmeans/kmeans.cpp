#include "kmeans.hpp"
#include "../core/factories.hpp"
#include "../kernels/launch.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace dagpu {

// kmeans_types.hpp:39-41 of the reference.
enum { KM_LLOYD = 0, KM_ELKAN, KM_HARTIGAN_WONG, KM_MACQUEEN };
enum { INIT_RANDOM = 0, INIT_KMEANSPP, INIT_SUPPLIED, INIT_RANDOM_PARTITIONS, INIT_AFK_MCMC };
enum { EMPTY_IGNORE = 0, EMPTY_ERROR, EMPTY_SPLIT };

template <typename T> KMeansHandle<T>::KMeansHandle(ErrorStack &e, GpuContext &g) : AlgHandle(e, g) {
    // Options of source/core/clustering/kmeans/kmeans_options.hpp:46-137.
    const da_int imax = std::numeric_limits<da_int>::max();
    opts.add_int("n_clusters", "Number of clusters required.", 1, LowerBound::greater_equal, imax,
                 UpperBound::p_inf, 1);
    opts.add_int("n_init",
                 "Number of runs with different random seeds (ignored if you have specified "
                 "initial cluster centres).",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 10);
    opts.add_int("max_iter", "Maximum number of iterations.", 1, LowerBound::greater_equal, imax,
                 UpperBound::p_inf, 300);
    opts.add_int("low precision max_iter",
                 "If mixed precision iterative refinement is enabled, maximum number of "
                 "iterations for the low precision phase.",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 200);
    opts.add_int("seed", "Seed for random number generation; set to -1 for non-deterministic results.",
                 -1, LowerBound::greater_equal, imax, UpperBound::p_inf, 0);
    opts.add_int("afk-mc2 samples", "Number of samples to take for the AFK-MC2 initialization method.",
                 1, LowerBound::greater_equal, imax, UpperBound::p_inf, 50);
    opts.add_string("initialization method", "How to determine the initial cluster centres.",
                    {{"random", INIT_RANDOM},
                     {"k-means++", INIT_KMEANSPP},
                     {"supplied", INIT_SUPPLIED},
                     {"random partitions", INIT_RANDOM_PARTITIONS},
                     {"afk-mc2", INIT_AFK_MCMC}},
                    "k-means++");
    opts.add_string("algorithm", "Choice of underlying k-means algorithm.",
                    {{"lloyd", KM_LLOYD},
                     {"elkan", KM_ELKAN},
                     {"hartigan-wong", KM_HARTIGAN_WONG},
                     {"macqueen", KM_MACQUEEN}},
                    "lloyd");
    opts.add_string("mixed precision",
                    "Whether to use mixed precision iterative refinement, in which lower precision "
                    "arithmetic is used before switching to the working precision for the final "
                    "iterations.",
                    {{"yes", 1}, {"no", 0}}, "no");
    opts.add_string("empty clusters",
                    "How to deal with empty clusters at the end of a k-means iteration.",
                    {{"ignore", EMPTY_IGNORE}, {"error", EMPTY_ERROR}, {"split", EMPTY_SPLIT}},
                    "ignore");
    opts.add_string("distance",
                    "Distance metric used for clustering. Use 'euclidean' for standard k-means or "
                    "'cosine' for spherical k-means.",
                    {{"euclidean", 0}, {"cosine", 1}}, "euclidean");
    opts.add_string("normalize data",
                    "Whether to normalize the input data before clustering. This option is only "
                    "used if distance is set to cosine.",
                    {{"yes", 1}, {"no", 0}}, "yes");
    // MI355X extension: operand precision of the MFMA cross term of the assignment (fp32
    // data). "working": fp32 MFMA; "bf16": bf16 MFMA with fp32 accumulation. Both list the
    // samples whose two nearest centres are within the error bound and re-check them with
    // exact double distances, so the labels do not depend on the choice.
    opts.add_string("cross term precision",
                    "Operand precision of the MFMA cross term in the assignment step (fp32 data): "
                    "working or bf16.",
                    {{"working", 0}, {"bf16", 1}}, "working");
    // MI355X extension: workgroups per CU of the MFMA assignment (0: one per 256 samples).
    opts.add_int("assignment workgroups per cu",
                 "Workgroups per CU of the MFMA assignment step, each taking every n-th block "
                 "of 256 samples (0: one workgroup per block).",
                 0, LowerBound::greater_equal, 64, UpperBound::less_equal, 0);
    opts.add_real("convergence tolerance", "Convergence tolerance.", 0.0,
                  LowerBound::greater_equal, 0.0, UpperBound::p_inf, (double)(T)1.0e-4);
    opts.add_real("low precision convergence tolerance",
                  "If mixed precision iterative refinement is enabled, convergence tolerance for "
                  "the low precision phase.",
                  0.0, LowerBound::greater_equal, 0.0, UpperBound::p_inf, (double)(T)1.0e-2);
}

template <typename T> void KMeansHandle<T>::refresh() { model_trained = false; }

template <typename T>
da_status KMeansHandle<T>::upload_data(const T *src, da_int ld, DeviceBuffer &dst) {
    hipStream_t st = gpu->stream();
    const size_t bytes = (size_t)n_samples_ * n_features_ * sizeof(T);
    DA_HIP(err, dst.reserve(bytes));
    if (data_order_ == row_major) {
        DA_HIP(err, copy_2d(dst.ptr(), n_features_ * sizeof(T), src, (size_t)ld * sizeof(T),
                                     n_features_ * sizeof(T), n_samples_, hipMemcpyHostToDevice, st));
    } else {
        DeviceBuffer stage;
        DA_HIP(err, stage.reserve(bytes));
        DA_HIP(err, copy_2d(stage.ptr(), n_samples_ * sizeof(T), src, (size_t)ld * sizeof(T),
                                     n_samples_ * sizeof(T), n_features_, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(stage.as<T>(), n_samples_, n_samples_, n_features_,
                                                   dst.as<T>(), st));
        DA_HIP(err, hipStreamSynchronize(st));
    }
    return da_status_success;
}

// Chooses the fixed-point exponents of the centroid sums from the column maxima so that
// no partial or total of n values can exceed 2^60 (see kmeans.hip).
template <typename T>
template <typename U>
da_status KMeansHandle<T>::prepare_scales_for(const U *dX) {
    hipStream_t st = gpu->stream();
    const int d = n_features_;
    DA_HIP(err, tmp_.reserve((size_t)std::max(d, 2048) * sizeof(double)));
    DA_HIP(err, hipMemsetAsync(tmp_.ptr(), 0, (size_t)d * sizeof(double), st));
    DA_HIP(err, launch_col_absmax<U>(dX, n_samples_, d, tmp_.as<double>(), st));
    std::vector<double> amax(d);
    DA_HIP(err, hipMemcpyAsync(amax.data(), tmp_.ptr(), d * sizeof(double), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    da_status s = upload_scales(amax, escale_);
    if (s != da_status_success)
        return s;
    // Sums of normalised samples (spherical k-means): every value is at most 1 in magnitude.
    return upload_scales(std::vector<double>(d, 1.0), escale_unit_);
}

template <typename T>
da_status KMeansHandle<T>::upload_scales(const std::vector<double> &amax, DeviceBuffer &dst) {
    hipStream_t st = gpu->stream();
    const int d = n_features_;
    int l2n = 0;
    while (((int64_t)1 << l2n) < (int64_t)n_samples_ + 1)
        l2n++;
    escale_host_.assign(d + 1, 0);
    for (int l = 0; l < d; l++) {
        if (amax[l] > 0 && std::isfinite(amax[l])) {
            int E = 0;
            std::frexp(amax[l], &E);
            escale_host_[l] = 60 - l2n - E;
        }
    }
    escale_host_[d] = 61 - l2n;
    DA_HIP(err, dst.reserve((d + 1) * sizeof(int)));
    DA_HIP(err, hipMemcpyAsync(dst.ptr(), escale_host_.data(), (d + 1) * sizeof(int),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T>
da_status KMeansHandle<T>::set_data(da_int n_samples, da_int n_features, const T *A, da_int lda) {
    refresh();
    initdone_ = false;
    data_order_ = order();
    da_status s = check_2D_array<T>(data_order_, n_samples, n_features, A, lda, "n_samples",
                                    "n_features", "A", "lda");
    if (s != da_status_success)
        return s;
    s = gpu->init(err);
    if (s != da_status_success)
        return s;
    A_usr_ = A;
    lda_usr_ = lda;
    n_samples_ = n_samples;
    n_features_ = n_features;
    s = upload_data(A, lda, X_);
    if (s != da_status_success)
        return s;
    return finish_set_data();
}

template <typename T>
da_status KMeansHandle<T>::set_data_device(da_int n_samples, da_int n_features, const T *dX) {
    refresh();
    initdone_ = false;
    data_order_ = row_major;
    if (n_samples < 1 || n_features < 1)
        return err->error(da_status_invalid_array_dimension, "The data has no rows or columns.");
    if (dX == nullptr)
        return err->error(da_status_invalid_pointer, "The data pointer is null.");
    da_status s = gpu->init(err);
    if (s != da_status_success)
        return s;
    (void)opts.set_string("storage order", "row-major");
    n_samples_ = n_samples;
    n_features_ = n_features;
    const size_t bytes = (size_t)n_samples * n_features * sizeof(T);
    DA_HIP(err, X_.reserve(bytes));
    DA_HIP(err, hipMemcpyAsync(X_.ptr(), dX, bytes, hipMemcpyDeviceToDevice, gpu->stream()));
    return finish_set_data();
}

template <typename T> da_status KMeansHandle<T>::finish_set_data() {
    const da_int n_samples = n_samples_, n_features = n_features_;
    da_status s = prepare_scales_for<T>(X_.as<T>());
    if (s != da_status_success)
        return s;
    mfma_ok_ = false;
    bf16_ok_ = false;
    xb_ready_ = false;
    // Padded copy for the MFMA assignment in the data's precision (feature count rounded up
    // to 16, at most 128).
    if (n_features <= 128) {
        kp_ = (n_features + 15) / 16 * 16;
        if (kp_ == 80 || kp_ == 112)
            kp_ = kp_ == 80 ? 96 : 128;
        DA_HIP(err, xpad_.reserve((size_t)n_samples * kp_ * sizeof(T)));
        DA_HIP(err, (launch_pack_rows<T, T>(X_.as<T>(), n_features, n_samples, n_features,
                                            xpad_.as<T>(), kp_, n_samples, gpu->stream())));
        DA_HIP(err, hipStreamSynchronize(gpu->stream()));
        mfma_ok_ = true;
        if constexpr (std::is_same_v<T, float>) {
            kp16_ = (n_features + 31) / 32 * 32;
            bf16_ok_ = true;
            xb_ready_ = false;
        }
    }
    initdone_ = true;
    model_trained = false;
    // n_clusters is bounded by n_samples once the data is known (kmeans.cpp:283-297).
    da_int temp = opts.ival("n_clusters");
    opts.rebound_int("n_clusters", n_samples, UpperBound::less_equal);
    opts.set_int("n_clusters", std::min(temp, n_samples));
    if (temp > n_samples)
        return err->warn(da_status_incompatible_options,
                         "The requested number of clusters has been decreased from " +
                             std::to_string(temp) + " to " + std::to_string(n_samples) +
                             " due to the size of the data array.");
    return da_status_success;
}

template <typename T> da_status KMeansHandle<T>::set_init_centres(const T *C, da_int ldc) {
    if (!initdone_)
        return err->error(da_status_no_data, "No data has been passed to the handle. Please call "
                                             "da_kmeans_set_data_s or da_kmeans_set_data_d.");
    if (order() != data_order_)
        return err->error(da_status_incompatible_options,
                          "The storage order option was changed after data was set.");
    da_int k = opts.ival("n_clusters");
    da_status s = check_2D_array<T>(data_order_, k, n_features_, C, ldc, "n_clusters", "n_features",
                                    "C", "ldc");
    if (s != da_status_success)
        return s;
    // Row-major centres are copied now, column-major ones read at compute time, as the
    // reference does (kmeans.cpp:300-348).
    if (data_order_ == row_major) {
        C_rowmajor_copy_.resize((size_t)k * n_features_);
        for (da_int i = 0; i < k; i++)
            for (da_int j = 0; j < n_features_; j++)
                C_rowmajor_copy_[(size_t)i * n_features_ + j] = C[(size_t)i * ldc + j];
        C_usr_ = C_rowmajor_copy_.data();
        ldc_usr_ = n_features_;
    } else {
        C_usr_ = C;
        ldc_usr_ = ldc;
    }
    n_clusters_ = k;
    centres_supplied_ = true;
    return da_status_success;
}

template <typename T>
da_status KMeansHandle<T>::upload_rows_double(const T *src, da_int rows, da_int ld, da_order ord,
                                              DeviceBuffer &dst) {
    std::vector<double> h((size_t)rows * n_features_);
    for (da_int i = 0; i < rows; i++)
        for (da_int j = 0; j < n_features_; j++)
            h[(size_t)i * n_features_ + j] =
                (double)(ord == row_major ? src[(size_t)i * ld + j] : src[i + (size_t)j * ld]);
    DA_HIP(err, dst.reserve(h.size() * sizeof(double)));
    DA_HIP(err, hipMemcpyAsync(dst.ptr(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                               gpu->stream()));
    DA_HIP(err, hipStreamSynchronize(gpu->stream()));
    return da_status_success;
}

template <typename T>
template <typename U>
da_status KMeansHandle<T>::lloyd_step(const U *dX, const double *Cin, double *Cout, int *lab_new,
                                      const int *lab_old, bool update, int &changed,
                                      double &normc2, double &shift2) {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    int *dchanged = reinterpret_cast<int *>(small_.as<double>() + 4);
    DA_HIP(err, hipMemsetAsync(dchanged, 0, 2 * sizeof(int), st));
    const bool on_data = std::is_same_v<U, T> && (const void *)dX == (const void *)X_.ptr();
    const bool use_bf16 = on_data && std::is_same_v<U, float> && bf16_ok_ &&
                          opts.skey("cross term precision") == 1 &&
                          kmeans_assign_bf16_lds(k, kp16_) <= 160 * 1024;
    const bool use_mfma = on_data && mfma_ok_;
    if (use_bf16) {
        // bf16 cross term on MFMA, exact re-check of near ties (kmeans_mfma.hip).
        int *dlist_count = dchanged + 1;
        double *cmax2 = small_.as<double>() + 6;
        DA_HIP(err, xb_.reserve((size_t)n * kp16_ * 2));
        if (!xb_ready_) {
            DA_HIP(err, launch_rows_to_bf16(reinterpret_cast<const float *>(X_.ptr()), n, d, d,
                                            xb_.ptr(), kp16_, st));
            xb_ready_ = true;
        }
        DA_HIP(err, c16_.reserve((size_t)(k + 16) * kp16_ * 2));
        DA_HIP(err, chalf_.reserve((size_t)(k + 16) * sizeof(float)));
        DA_HIP(err, list_.reserve((size_t)n * sizeof(int)));
        DA_HIP(err, hipMemsetAsync(cmax2, 0, sizeof(double), st));
        DA_HIP(err, launch_centre_prep_bf16(Cin, k, d, kp16_, c16_.ptr(), chalf_.as<float>(),
                                            cmax2, st, sph_));
        double hc = 0;
        DA_HIP(err, hipMemcpyAsync(&hc, cmax2, sizeof(double), hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
        {
            KernelScope ks(*gpu, "kmeans_assign");
            DA_HIP(err, launch_kmeans_assign_bf16(xb_.ptr(), n, kp16_, c16_.ptr(),
                                                  chalf_.as<float>(), k, std::sqrt(hc), lab_new,
                                                  lab_old, dchanged, dlist_count, list_.as<int>(),
                                                  st, (int)opts.ival("assignment workgroups per cu")));
        }
        {
            KernelScope ks(*gpu, "kmeans_recheck");
            DA_HIP(err, ct_.reserve((size_t)k * d * sizeof(double)));
            DA_HIP(err, launch_kmeans_recheck<U>(dX, d, d, Cin, k, dlist_count, list_.as<int>(),
                                                 lab_new, lab_old, dchanged, ct_.as<double>(),
                                                 st, sph_));
        }
    } else if (use_mfma) {
        // Cross term on the fp32 or fp64 matrix cores, exact re-check of near ties
        // (kmeans_mfma.hip).
        int *dlist_count = dchanged + 1;
        double *cmax2 = small_.as<double>() + 6;
        DA_HIP(err, hipMemsetAsync(cmax2, 0, sizeof(double), st));
        DA_HIP(err, launch_centre_prep<T>(Cin, k, d, kp_, c32_.as<T>(), chalf_.as<T>(), cmax2,
                                          st, sph_));
        double hc = 0;
        DA_HIP(err, hipMemcpyAsync(&hc, cmax2, sizeof(double), hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
        {
            KernelScope ks(*gpu, "kmeans_assign");
            DA_HIP(err, launch_kmeans_assign_mfma<T>(xpad_.as<T>(), n, kp_, c32_.as<T>(),
                                                     chalf_.as<T>(), k, std::sqrt(hc), lab_new,
                                                  lab_old, dchanged, dlist_count, list_.as<int>(), st,
                                                  (int)opts.ival("assignment workgroups per cu")));
        }
        {
            KernelScope ks(*gpu, "kmeans_recheck");
            DA_HIP(err, ct_.reserve((size_t)k * d * sizeof(double)));
            DA_HIP(err, launch_kmeans_recheck<U>(dX, d, d, Cin, k, dlist_count, list_.as<int>(),
                                                 lab_new, lab_old, dchanged, ct_.as<double>(),
                                                 st, sph_));
        }
    } else {
        KernelScope ks(*gpu, "kmeans_assign");
        DA_HIP(err, launch_kmeans_assign<U>(dX, n, d, Cin, k, lab_new, lab_old, dchanged,
                                            mind_.as<double>(), st, sph_));
    }
    if (update) {
        da_status s = update_centres<U>(dX, lab_new, Cin, Cout);
        if (s != da_status_success)
            return s;
    }
    return read_step(changed, normc2, shift2);
}

template <typename T>
template <typename U>
da_status KMeansHandle<T>::update_centres(const U *dX, const int *lab_new, const double *Cin,
                                          double *Cout) {
    hipStream_t st = gpu->stream();
    const int d = n_features_, k = n_clusters_;
    KernelScope ks(*gpu, "kmeans_update");
    // Spherical k-means: sums of x (or of x/|x|), normalised (kmeans_lloyd.hpp:279-300,
    // 363-405).
    da_status s = centre_sums<U>(dX, lab_new, Cout, winv() ? escale_unit_.as<int>() : escale_.as<int>(),
                                 winv(), sph_);
    if (s != da_status_success)
        return s;
    // The low-precision phase stores its centres in the lower precision (kmeans<float> or
    // kmeans<_Float16> in kmeans.cpp:1824-1876); the move is measured after the rounding.
    if (lp_mode_ != 0)
        DA_HIP(err, launch_round_through(Cout, (int64_t)k * d, lp_mode_, st));
    DA_HIP(err, launch_kmeans_shift(Cin, Cout, k * d, small_.as<double>(), st));
    return da_status_success;
}

template <typename T>
da_status KMeansHandle<T>::read_step(int &changed, double &normc2, double &shift2) {
    hipStream_t st = gpu->stream();
    double h[5];
    DA_HIP(err, hipMemcpyAsync(h, small_.ptr(), 5 * sizeof(double), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    std::memcpy(&changed, &h[4], sizeof(int));
    int listed = 0;
    std::memcpy(&listed, reinterpret_cast<const char *>(&h[4]) + sizeof(int), sizeof(int));
    rechecked_ += listed;
    normc2 = h[0];
    shift2 = h[1];
    return da_status_success;
}

// One iteration of Elkan's algorithm (elkan_iteration, kmeans_elkan.hpp:372-513): the
// bounded assignment from lab_old, then (with update) the centre update, the bound update
// with the centre shifts and the half distances of the new centres.
template <typename T>
template <typename U>
da_status KMeansHandle<T>::elkan_step(const U *dX, const double *Cin, double *Cout, int *lab_new,
                                      const int *lab_old, bool update, int &changed,
                                      double &normc2, double &shift2) {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    int *dchanged = reinterpret_cast<int *>(small_.as<double>() + 4);
    DA_HIP(err, hipMemsetAsync(dchanged, 0, 2 * sizeof(int), st));
    {
        KernelScope ks(*gpu, "kmeans_assign");
        DA_HIP(err, launch_elkan_assign<U>(dX, n, d, Cin, k, ecc_.as<double>(), enxt_.as<double>(),
                                           lab_old, lab_new, eu_.as<double>(), el_.as<double>(),
                                           dchanged, st));
    }
    if (update) {
        da_status s = update_centres<U>(dX, lab_new, Cin, Cout);
        if (s != da_status_success)
            return s;
        DA_HIP(err, launch_elkan_bounds(Cin, Cout, n, k, d, lab_new, eshift_.as<double>(),
                                        eu_.as<double>(), el_.as<double>(), st));
        DA_HIP(err, launch_elkan_half(Cout, k, d, ecc_.as<double>(), enxt_.as<double>(), st));
    }
    return read_step(changed, normc2, shift2);
}

// Centres from labels: the exact fixed-point sums in passes of at most KMEANS_PASS_PAIRS
// (cluster, feature) pairs, each summed by nblocks_ workgroups and reduced per cluster;
// divided by the cluster sizes, or (sph) normalised once every feature is summed.
template <typename T>
template <typename U>
da_status KMeansHandle<T>::centre_sums(const U *dX, const int *labels, double *C, const int *esc,
                                       const double *w, int sph) {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    for (int c0 = 0; c0 < k; c0 += kc_) {
        const int kc = std::min<int>(kc_, k - c0);
        for (int l0 = 0; l0 < d; l0 += dl_) {
            const int dl = std::min<int>(dl_, d - l0);
            DA_HIP(err, launch_kmeans_partial_sums<U>(dX, n, d, labels, c0, kc, l0, dl, esc, ppb_,
                                                      nblocks_, part_.as<long long>(),
                                                      pcnt_.as<int>(), st, w));
            DA_HIP(err, launch_kmeans_reduce(part_.as<long long>(), pcnt_.as<int>(), nblocks_, c0,
                                             kc, l0, dl, d, esc, C, counts_.as<int>(), st, sph));
        }
    }
    if (sph)
        DA_HIP(err, launch_normalize_rows(C, k, d, st));
    return da_status_success;
}

template <typename T>
template <typename U>
da_status KMeansHandle<T>::centres_from_labels(const U *dX, const int *labels, double *C) {
    return centre_sums<U>(dX, labels, C, escale_.as<int>(), nullptr, 0);
}

// Empty-cluster handling after an iteration (kmeans.cpp:1127-1330): "error" flags the run,
// "split" moves the furthest point of a non-singleton cluster into each empty cluster. The
// split bookkeeping is sequential and only runs when a cluster has emptied, so it is done
// on the host on copies of the labels, centres and data.
template <typename T>
template <typename U>
da_status KMeansHandle<T>::handle_empty_clusters(const U *dX, double *C, int *labels, bool &split,
                                                 bool &empty_error) {
    split = false;
    empty_error = false;
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    // The cluster sizes of the update just made (k integers) decide whether anything is to
    // be done; the labels and the data only come to the host when a cluster is empty.
    std::vector<int> count(k, 0);
    DA_HIP(err, hipMemcpyAsync(count.data(), counts_.ptr(), k * sizeof(int), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    int n_empty = 0;
    for (int c = 0; c < k; c++)
        n_empty += count[c] == 0;
    if (n_empty == 0)
        return da_status_success;
    if (empty_mode_ == EMPTY_ERROR) {
        empty_error = true;
        return da_status_success;
    }
    std::vector<int> lab(n);
    DA_HIP(err, hipMemcpyAsync(lab.data(), labels, n * sizeof(int), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    std::vector<U> A((size_t)n * d);
    std::vector<double> cen((size_t)k * d), inv;
    DA_HIP(err, hipMemcpyAsync(A.data(), dX, A.size() * sizeof(U), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipMemcpyAsync(cen.data(), C, cen.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (winv()) {
        inv.resize(n);
        DA_HIP(err, hipMemcpyAsync(inv.data(), winv(), n * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    // Distance of a sample to a centre: squared Euclidean, or for spherical k-means
    // 1 - x.c/|x| (1 - x.c for data taken as unit rows), kmeans.cpp:1171-1199.
    auto sqd = [&](int i, int c) {
        double s = 0;
        if (sph_) {
            for (int j = 0; j < d; j++)
                s += (double)A[(size_t)i * d + j] * cen[(size_t)c * d + j];
            return 1.0 - (inv.empty() ? s : s * inv[i]);
        }
        for (int j = 0; j < d; j++) {
            double t = (double)A[(size_t)i * d + j] - cen[(size_t)c * d + j];
            s += t * t;
        }
        return s;
    };
    // Spherical k-means relocates x/|x| (x for data taken as unit rows) and normalises both
    // centres it changes (kmeans.cpp:1224-1265).
    auto normalise = [&](int c) {
        double ss = 0;
        for (int j = 0; j < d; j++)
            ss += cen[(size_t)c * d + j] * cen[(size_t)c * d + j];
        if (ss > 0) {
            const double r = 1.0 / std::sqrt(ss);
            for (int j = 0; j < d; j++)
                cen[(size_t)c * d + j] *= r;
        }
    };
    std::vector<double> dist(n);
    for (int i = 0; i < n; i++)
        dist[i] = sqd(i, lab[i]);
    for (int c = 0; c < k; c++) {
        if (count[c] != 0)
            continue;
        int best = -1;
        double bd = -1.0;
        for (int i = 0; i < n; i++)
            if (count[lab[i]] > 1 && std::isfinite(dist[i]) && dist[i] > bd) {
                bd = dist[i];
                best = i;
            }
        if (best < 0)
            break;
        const int old = lab[best];
        const int n_old = count[old];
        const double wb = inv.empty() ? 1.0 : inv[best];
        for (int j = 0; j < d; j++) {
            double &cv = cen[(size_t)old * d + j];
            cv = (cv * n_old - (double)A[(size_t)best * d + j] * wb) / (n_old - 1);
        }
        for (int j = 0; j < d; j++)
            cen[(size_t)c * d + j] = (double)A[(size_t)best * d + j] * wb;
        if (sph_) {
            normalise(old);
            normalise(c);
        }
        count[old] -= 1;
        count[c] = 1;
        lab[best] = c;
        dist[best] = 0.0;
        if (n_empty > 1)
            for (int i = 0; i < n; i++)
                if (lab[i] == old)
                    dist[i] = sqd(i, old);
    }
    DA_HIP(err, hipMemcpyAsync(C, cen.data(), cen.size() * sizeof(double), hipMemcpyHostToDevice, st));
    DA_HIP(err, hipMemcpyAsync(labels, lab.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    split = true;
    return da_status_success;
}

// One k-means run from the centres in Ccur_ (perform_kmeans, kmeans.cpp:996-1058).
template <typename T>
template <typename U>
da_status KMeansHandle<T>::run_lloyd(const U *dX, double tol, da_int max_iter, RunResult &res) {
    hipStream_t st = gpu->stream();
    int conv = 0;
    da_int it = 0;
    int changed = 0;
    double normc2 = 0, shift2 = 0;
    res = RunResult{};
    auto step = [&](const double *Cin, double *Cout, int *lab_new, const int *lab_old,
                    bool update) {
        return elkan_ ? elkan_step<U>(dX, Cin, Cout, lab_new, lab_old, update, changed, normc2,
                                      shift2)
                      : lloyd_step<U>(dX, Cin, Cout, lab_new, lab_old, update, changed, normc2,
                                      shift2);
    };
    if (elkan_) {
        // init_elkan (kmeans_elkan.hpp:143-233): labels and bounds of the initial centres.
        const int n = n_samples_, d = n_features_, k = n_clusters_;
        DA_HIP(err, eu_.reserve((size_t)n * sizeof(double)));
        DA_HIP(err, el_.reserve((size_t)n * k * sizeof(double)));
        DA_HIP(err, ecc_.reserve((size_t)k * k * sizeof(double)));
        DA_HIP(err, enxt_.reserve((size_t)k * sizeof(double)));
        DA_HIP(err, eshift_.reserve((size_t)k * sizeof(double)));
        DA_HIP(err, launch_elkan_half(Ccur_, k, d, ecc_.as<double>(), enxt_.as<double>(), st));
        DA_HIP(err, launch_elkan_init<U>(dX, n, d, Ccur_, k, ecc_.as<double>(), lab_cur_,
                                         eu_.as<double>(), el_.as<double>(), st));
    }
    for (it = 0; it < max_iter; it++) {
        std::swap(Cprev_, Ccur_);
        std::swap(lab_prev_, lab_cur_);
        da_status s = step(Cprev_, Ccur_, lab_cur_, lab_prev_, true);
        if (s != da_status_success)
            return s;
        if (empty_mode_ != EMPTY_IGNORE) {
            bool split = false, eerr = false;
            s = handle_empty_clusters<U>(dX, Ccur_, lab_cur_, split, eerr);
            if (s != da_status_success)
                return s;
            if (eerr) {
                res.empty_found = true;
                return da_status_success;
            }
            if (split) {
                // The bounds do not hold after a split (kmeans.cpp:1312-1317).
                if (elkan_)
                    DA_HIP(err, launch_elkan_reset(n_samples_, n_clusters_, eu_.as<double>(),
                                                   el_.as<double>(), st));
                continue;
            }
        }
        // convergence_test, kmeans.cpp:1359-1397.
        conv = 0;
        if (it > 1 && changed == 0)
            conv = 2;
        else if (std::sqrt(shift2) < tol * std::sqrt(normc2))
            conv = 1;
        if (conv > 0)
            break;
    }
    res.n_iter = it;
    res.converged = conv;
    if (conv == 1 || it == max_iter) {
        std::swap(lab_prev_, lab_cur_);
        std::swap(Cprev_, Ccur_);
        da_status s = elkan_ ? step(Cprev_, nullptr, lab_cur_, lab_prev_, false)
                             : step(Cprev_, nullptr, lab_cur_, nullptr, false);
        if (s != da_status_success)
            return s;
        std::swap(Cprev_, Ccur_);
    }
    // compute_current_inertia (kmeans.cpp:1060-1124), exact squared distances summed in a
    // fixed order.
    // Spherical k-means: sum of 1 - x.c/|x| (1 - x.c when the data are taken as unit rows).
    DA_HIP(err, launch_label_sqdist<U>(dX, n_samples_, n_features_, Ccur_, lab_cur_,
                                       mind_.as<double>(), st, sph_, winv()));
    DA_HIP(err, launch_sum_double(mind_.as<double>(), n_samples_, tmp_.as<double>(), small_.as<double>() + 2, st));
    double inertia = 0;
    DA_HIP(err, hipMemcpyAsync(&inertia, small_.as<double>() + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    res.inertia = inertia;
    return da_status_success;
}

// k-means++ (kmeans.cpp:1524-1664): the random draws are made on the host with the same
// libstdc++ engine and distributions in the same order; the distance passes, the weight
// normalisation and the inverse-CDF lookup of std::discrete_distribution run on the GPU.
template <typename T> da_status KMeansHandle<T>::kmeans_plusplus() {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    const T *dX = X_.as<T>();
    std::vector<int> chosen(k, 0);
    const da_int n_trials = 2 + (da_int)std::log((double)k);
    std::uniform_int_distribution<> dis_int(0, n - 1);
    const int first = dis_int(mt_);
    chosen[0] = first;
    double *w = w_.as<double>();
    DA_HIP(err, launch_dist_to_row<T>(dX, n, d, first, w, st, sph_, winv()));
    double *total = small_.as<double>() + 3;
    DA_HIP(err, launch_sum_double(w, n, tmp_.as<double>(), total, st));
    double htot = 0;
    DA_HIP(err, hipMemcpyAsync(&htot, total, sizeof(double), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    if (!(htot > 0.0)) {
        // All points coincide with the first centre: take the first k points.
        for (int c = 0; c < k; c++)
            chosen[c] = c;
    } else {
        // Partial sums of the candidate costs: one per workgroup, about 1,024 samples each, at
        // most 2,048 (their n_trials * nb partials fit the n doubles of cp_, n_trials <= 32).
        const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(2048, (int64_t)n / 1024));
        std::vector<double> u(n_trials), cost(n_trials);
        std::vector<int> cand(n_trials);
        double *du = tmp_.as<double>() + 1024;
        int *dcand = idx_.as<int>();
        double *dcost = tmp_.as<double>() + 1024 + 64;
        double *part = cp_.as<double>(); // reused after the lookup
        for (int c = 1; c < k; c++) {
            if (n >= 2) {
                DA_HIP(err, launch_sum_double(w, n, tmp_.as<double>(), total, st));
                DA_HIP(err, launch_normalised_scan(w, n, total, cp_.as<double>(), tmp_.as<double>(), st));
                for (int t = 0; t < n_trials; t++)
                    u[t] = std::generate_canonical<double, std::numeric_limits<double>::digits>(mt_);
                DA_HIP(err, hipMemcpyAsync(du, u.data(), n_trials * sizeof(double), hipMemcpyHostToDevice, st));
                DA_HIP(err, launch_lower_bound(cp_.as<double>(), n, du, n_trials, dcand, st));
                DA_HIP(err, hipMemcpyAsync(cand.data(), dcand, n_trials * sizeof(int), hipMemcpyDeviceToHost, st));
                DA_HIP(err, hipStreamSynchronize(st));
            } else {
                std::fill(cand.begin(), cand.end(), 0);
            }
            DA_HIP(err, launch_candidate_cost<T>(dX, n, d, w, dcand, n_trials, part, nb, st, sph_,
                                                 winv()));
            DA_HIP(err, launch_reduce_rows(part, n_trials, nb, dcost, st));
            DA_HIP(err, hipMemcpyAsync(cost.data(), dcost, n_trials * sizeof(double), hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipStreamSynchronize(st));
            int best = 0;
            double best_cost = std::numeric_limits<double>::infinity();
            for (int t = 0; t < n_trials; t++) {
                bool repeat = false;
                for (int j = 0; j < t; j++)
                    repeat |= cand[j] == cand[t];
                if (repeat)
                    continue;
                if (std::isfinite(cost[t]) && cost[t] < best_cost) {
                    best_cost = cost[t];
                    best = cand[t];
                }
            }
            chosen[c] = best;
            DA_HIP(err, launch_min_update<T>(dX, n, d, best, w, st, sph_, winv()));
        }
    }
    DA_HIP(err, hipMemcpyAsync(idx_.ptr(), chosen.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
    DA_HIP(err, launch_gather_rows<T>(dX, d, idx_.as<int>(), k, Ccur_, st));
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

// initialize_centres, kmeans.cpp:1400-1486.
template <typename T> da_status KMeansHandle<T>::initialise_centres(da_int method) {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    switch (method) {
    case INIT_SUPPLIED: {
        std::vector<double> h((size_t)k * d);
        for (int i = 0; i < k; i++)
            for (int j = 0; j < d; j++)
                h[(size_t)i * d + j] = (double)(data_order_ == row_major
                                                    ? C_usr_[(size_t)i * ldc_usr_ + j]
                                                    : C_usr_[i + (size_t)j * ldc_usr_]);
        DA_HIP(err, hipMemcpyAsync(Ccur_, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
        DA_HIP(err, hipStreamSynchronize(st));
        return da_status_success;
    }
    case INIT_RANDOM: {
        // da_std::sample (da_std.hpp:102-118), selection sampling: sample i of the remaining
        // `left` is taken when a draw of uniform_int_distribution(0, left - 1) is below the
        // number still needed (one draw per sample visited, unlike std::sample).
        std::vector<int> p32;
        p32.reserve(k);
        da_int need = k, left = n;
        for (da_int i = 0; need > 0; i++, left--) {
            std::uniform_int_distribution<da_int> u(0, left - 1);
            if (u(mt_) < need) {
                p32.push_back((int)i);
                need--;
            }
        }
        DA_HIP(err, hipMemcpyAsync(idx_.ptr(), p32.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_gather_rows<T>(X_.as<T>(), d, idx_.as<int>(), k, Ccur_, st));
        DA_HIP(err, hipStreamSynchronize(st));
        return da_status_success;
    }
    case INIT_RANDOM_PARTITIONS: {
        std::uniform_int_distribution<> dis(0, k - 1);
        std::vector<int> lab(n);
        for (int i = 0; i < n; i++)
            lab[i] = dis(mt_);
        DA_HIP(err, hipMemcpyAsync(lab_cur_, lab.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
        da_status s = centres_from_labels<T>(X_.as<T>(), lab_cur_, Ccur_);
        if (s != da_status_success)
            return s;
        // The reference divides by the count without a guard, so an empty partition gives
        // a NaN centre.
        std::vector<int> cnt(k);
        DA_HIP(err, hipMemcpyAsync(cnt.data(), counts_.ptr(), k * sizeof(int), hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
        std::vector<double> nanrow(d, std::numeric_limits<double>::quiet_NaN());
        for (int c = 0; c < k; c++)
            if (cnt[c] == 0)
                DA_HIP(err, hipMemcpy(Ccur_ + (size_t)c * d, nanrow.data(), d * sizeof(double),
                                      hipMemcpyHostToDevice));
        return da_status_success;
    }
    case INIT_KMEANSPP:
        return kmeans_plusplus();
    case INIT_AFK_MCMC:
        return afk_mc2();
    default:
        return err->error(da_status_invalid_option, "Unknown initialization method.");
    }
}

// AFK-MC2 (afk_mcmc_init, kmeans.cpp:1667-1811): the first centre uniform; the proposal
// q(x) = d(x, c1)^2 / (2 sum) + 1/(2n) sampled through std::discrete_distribution; each further
// centre is the end of a Markov chain of "afk-mc2 samples" steps whose acceptance compares
// the squared distances to the nearest chosen centre. The engine draws are made on the host
// with the same libstdc++ engine and distributions; the distance passes, the proposal, its
// cumulative distribution and the chain itself run on the GPU. A chain takes one canonical
// draw per proposal and one per acceptance test that needs a uniform, so the host prepares
// the most a chain can use (2m + 1) and then advances its engine by the number consumed.
template <typename T> da_status KMeansHandle<T>::afk_mc2() {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_, k = n_clusters_;
    const T *dX = X_.as<T>();
    const int m = (int)opts.ival("afk-mc2 samples");
    std::vector<int> chosen(k, 0);
    std::uniform_int_distribution<> dis_int(0, n - 1);
    chosen[0] = dis_int(mt_);
    // w: squared distance of each sample to its nearest chosen centre.
    double *w = w_.as<double>();
    DA_HIP(err, launch_dist_to_row<T>(dX, n, d, chosen[0], w, st, sph_, winv()));
    if (k > 1) {
        double *total = small_.as<double>() + 3;
        int *dout = reinterpret_cast<int *>(small_.as<double>() + 10);
        DA_HIP(err, aq_.reserve((size_t)n * sizeof(double)));
        DA_HIP(err, au_.reserve((size_t)(2 * m + 1) * sizeof(double)));
        double *q = aq_.as<double>(), *cp = cp_.as<double>();
        DA_HIP(err, launch_sum_double(w, n, tmp_.as<double>(), total, st));
        DA_HIP(err, launch_afk_proposal(w, n, total, q, st));
        // std::discrete_distribution: the weights normalised by their sum, accumulated, the
        // last entry set to 1.
        DA_HIP(err, launch_sum_double(q, n, tmp_.as<double>(), total, st));
        DA_HIP(err, launch_normalised_scan(q, n, total, cp, tmp_.as<double>(), st));
        std::vector<double> u(2 * (size_t)m + 1);
        for (int c = 1; c < k; c++) {
            const std::mt19937_64 saved = mt_;
            for (auto &v : u)
                v = std::generate_canonical<double, std::numeric_limits<double>::digits>(mt_);
            DA_HIP(err, hipMemcpyAsync(au_.ptr(), u.data(), u.size() * sizeof(double),
                                       hipMemcpyHostToDevice, st));
            DA_HIP(err, launch_afk_chain(cp, q, w, n, m, au_.as<double>(), dout, st));
            int out[2] = {0, 0};
            DA_HIP(err, hipMemcpyAsync(out, dout, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipStreamSynchronize(st));
            mt_ = saved;
            mt_.discard((unsigned long long)out[1]);
            chosen[c] = out[0];
            DA_HIP(err, launch_min_update<T>(dX, n, d, out[0], w, st, sph_, winv()));
        }
    }
    DA_HIP(err, hipMemcpyAsync(idx_.ptr(), chosen.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
    DA_HIP(err, launch_gather_rows<T>(dX, d, idx_.as<int>(), k, Ccur_, st));
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T> da_status KMeansHandle<T>::compute() {
    rechecked_ = 0;
    if (!initdone_)
        return err->error(da_status_no_data, "No data has been passed to the handle. Please call "
                                             "da_kmeans_set_data_s or da_kmeans_set_data_d.");
    if (order() != data_order_)
        return err->error(da_status_incompatible_options,
                          "The storage order option was changed after data was set.");
    da_int k = opts.ival("n_clusters");
    const da_int init = opts.skey("initialization method");
    da_int n_init = opts.ival("n_init");
    const da_int max_iter = opts.ival("max_iter");
    const double tol = opts.rval("convergence tolerance");
    da_int seed = opts.ival("seed");
    const da_int algorithm = opts.skey("algorithm");
    const bool mixed = opts.skey("mixed precision") == 1;
    empty_mode_ = opts.skey("empty clusters");
    const bool spherical = opts.skey("distance") == 1;
    // Spherical k-means with Elkan's algorithm runs Lloyd's iteration: Elkan's angular bounds
    // (kmeans_elkan.hpp:143-233) only skip distance evaluations, the labels are the argmax of
    // x.c either way (an exact tie goes to the lowest centre here).
    elkan_ = algorithm == KM_ELKAN && !spherical;
    sph_ = spherical;
    sph_norm_ = spherical && opts.skey("normalize data") == 1;
    lp_mode_ = 0;
    opts.rebound_int("n_clusters", std::numeric_limits<da_int>::max(), UpperBound::p_inf);

    da_status status = da_status_success;
    if (n_init > 1 && init == INIT_SUPPLIED) {
        n_init = 1;
        err->warn(da_status_incompatible_options,
                  "n_init was set to " + std::to_string(opts.ival("n_init")) +
                      " but the initialization method was set to 'supplied'. The k-means "
                      "algorithm will only be run once.");
    }
    if (algorithm == KM_HARTIGAN_WONG && (k == 1 || k >= n_samples_))
        return err->error(da_status_incompatible_options,
                          "The Hartigan-Wong algorithm requires 1 < k < n_samples.");
    if (spherical && algorithm == KM_HARTIGAN_WONG)
        return err->error(da_status_incompatible_options,
                          "Cosine distance is not compatible with the Hartigan-Wong algorithm. "
                          "Please use Lloyd, Elkan, or MacQueen.");
    if (algorithm == KM_HARTIGAN_WONG || algorithm == KM_MACQUEEN)
        return err->error(da_status_not_implemented,
                          "Only the Lloyd and Elkan k-means algorithms are implemented by the "
                          "MI355X build.");
    if (spherical && mixed)
        return err->error(da_status_not_implemented,
                          "Mixed-precision refinement of spherical (cosine) k-means is not "
                          "implemented by the MI355X build.");
    if (k > n_samples_)
        return err->error(da_status_incompatible_options,
                          "n_clusters = " + std::to_string(k) + ", and n_samples = " +
                              std::to_string(n_samples_) + ". Constraint: n_clusters <= n_samples.");
    if (init == INIT_SUPPLIED && !centres_supplied_)
        return err->error(da_status_no_data, "The initialization method was set to 'supplied' but "
                                             "no initial centres have been provided.");
    if (init == INIT_SUPPLIED && k != n_clusters_)
        return err->error(da_status_incompatible_options,
                          "n_clusters was changed after the initial centres were supplied.");
    n_clusters_ = k;
    const int n = n_samples_, d = n_features_;
    hipStream_t st = gpu->stream();

    // Work buffers.
    const size_t cb = (size_t)k * d * sizeof(double);
    DA_HIP(err, Ca_.reserve(cb));
    DA_HIP(err, Cb_.reserve(cb));
    DA_HIP(err, lab_a_.reserve((size_t)n * sizeof(int)));
    DA_HIP(err, lab_b_.reserve((size_t)n * sizeof(int)));
    DA_HIP(err, hipMemsetAsync(lab_a_.ptr(), 0, (size_t)n * sizeof(int), st));
    DA_HIP(err, hipMemsetAsync(lab_b_.ptr(), 0, (size_t)n * sizeof(int), st));
    DA_HIP(err, counts_.reserve((size_t)k * sizeof(int)));
    DA_HIP(err, mind_.reserve((size_t)n * sizeof(double)));
    DA_HIP(err, small_.reserve(16 * sizeof(double)));
    DA_HIP(err, tmp_.reserve(4096 * sizeof(double)));
    DA_HIP(err, idx_.reserve((size_t)std::max(k, 64) * sizeof(int)));
    DA_HIP(err, w_.reserve((size_t)n * sizeof(double)));
    DA_HIP(err, cp_.reserve((size_t)std::max(n, 256 * 64) * sizeof(double)));
    // Centroid sums: chunks of up to 8,192 points (sorted by label in LDS), about two per
    // workgroup, one workgroup per CU; passes of at most KMEANS_PASS_PAIRS (cluster, feature)
    // pairs summed in registers (all of them for k = 256 at d = 32), at most 12,288 clusters
    // per pass (the label counters share the 160 KB of LDS with the point list).
    {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus < 1)
            cus = 256;
        ppb_ = std::max<da_int>(256, std::min<da_int>(8192, (n + 2 * cus - 1) / (2 * cus)));
        const da_int nchunks = (n + ppb_ - 1) / ppb_;
        nblocks_ = std::max<da_int>(1, std::min<da_int>(nchunks, cus));
    }
    dl_ = std::min<da_int>(d, KMEANS_PASS_PAIRS);
    kc_ = std::max<da_int>(1, std::min<da_int>({(da_int)k, (da_int)12288,
                                                (da_int)(KMEANS_PASS_PAIRS / dl_)}));
    DA_HIP(err, part_.reserve((size_t)nblocks_ * kc_ * dl_ * 2 * sizeof(long long)));
    if (mfma_ok_) {
        DA_HIP(err, c32_.reserve((size_t)(k + 16) * kp_ * sizeof(T)));
        DA_HIP(err, chalf_.reserve((size_t)(k + 16) * sizeof(T)));
        DA_HIP(err, list_.reserve((size_t)n * sizeof(int)));
    }
    DA_HIP(err, pcnt_.reserve((size_t)nblocks_ * kc_ * sizeof(int)));
    Ccur_ = Ca_.as<double>();
    Cprev_ = Cb_.as<double>();
    lab_cur_ = lab_a_.as<int>();
    lab_prev_ = lab_b_.as<int>();
    if (sph_norm_) {
        // data_inv_norms, kmeans.cpp:681-703: safe_inv(|x_i|).
        DA_HIP(err, inv_.reserve((size_t)n * sizeof(double)));
        DA_HIP(err, launch_row_inv_norms<T>(X_.as<T>(), n, d, inv_.as<double>(), st));
    }

    // initialize_rng, kmeans.cpp:1814-1822.
    if (init != INIT_SUPPLIED) {
        if (seed == -1) {
            std::random_device r;
            seed = std::abs((da_int)r());
        }
        mt_.seed(seed);
    }

    bool valid = false, warn_maxit = false;
    best_inertia_ = std::numeric_limits<double>::infinity();
    best_centres_.clear();
    for (da_int run = 0; run < n_init; run++) {
        da_status s = initialise_centres(init);
        if (s != da_status_success)
            return s;
        // Spherical k-means starts from unit centres (kmeans.cpp:1478-1486).
        if (sph_)
            DA_HIP(err, launch_normalize_rows(Ccur_, k, d, st));
        da_int lp_iter = 0;
        if (mixed) {
            // lower_precision_init, kmeans.cpp:1824-1876: a full k-means on the data rounded
            // to the lower precision (float for double, half for float) from the current
            // centres rounded likewise; its centres, rounded, seed the working-precision run.
            const double lp_tol = opts.rval("low precision convergence tolerance");
            const da_int lp_max = opts.ival("low precision max_iter");
            if (lp_tol <= tol)
                return err->error(da_status_incompatible_options,
                                  "Low precision convergence tolerance must be greater than "
                                  "convergence tolerance.");
            const int mode = std::is_same_v<T, double> ? 1 : 2;
            DA_HIP(err, Xlp_.reserve((size_t)n * d * sizeof(T)));
            DA_HIP(err, launch_round_data<T>(X_.as<T>(), (int64_t)n * d, mode, Xlp_.as<T>(), st));
            DA_HIP(err, launch_round_through(Ccur_, (int64_t)k * d, mode, st));
            RunResult lp;
            lp_mode_ = mode;
            s = run_lloyd<T>(Xlp_.as<T>(), lp_tol, lp_max, lp);
            lp_mode_ = 0;
            if (s != da_status_success)
                return s;
            if (lp.empty_found)
                return err->error(da_status_empty_clusters,
                                  "All 1 run(s) of the k-means algorithm resulted in at least one "
                                  "empty cluster during the k-means iterations. A better set of "
                                  "initial cluster centers is needed.");
            lp_iter = lp.n_iter;
            DA_HIP(err, launch_round_through(Ccur_, (int64_t)k * d, mode, st));
        }
        RunResult r;
        s = run_lloyd<T>(X_.as<T>(), tol, max_iter, r);
        if (s != da_status_success)
            return s;
        if (r.empty_found)
            continue;
        valid = true;
        if (r.inertia < best_inertia_ || best_centres_.empty()) {
            best_inertia_ = r.inertia;
            best_n_iter_ = r.n_iter;
            best_lp_n_iter_ = lp_iter;
            warn_maxit = (r.converged == 0);
            best_centres_.resize((size_t)k * d);
            best_labels_.resize(n);
            std::vector<int> lab(n);
            DA_HIP(err, hipMemcpyAsync(best_centres_.data(), Ccur_, cb, hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipMemcpyAsync(lab.data(), lab_cur_, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipStreamSynchronize(st));
            std::copy(lab.begin(), lab.end(), best_labels_.begin());
        }
    }
    if (!valid)
        return err->error(da_status_empty_clusters,
                          "All " + std::to_string(n_init) +
                              " run(s) of the k-means algorithm resulted in at least one empty "
                              "cluster during the k-means iterations. A better set of initial "
                              "cluster centers is needed.");
    // The reported centres are in the working precision.
    for (auto &v : best_centres_)
        v = (double)(T)v;
    DA_HIP(err, Cbest_.reserve(cb));
    DA_HIP(err, hipMemcpyAsync(Cbest_.ptr(), best_centres_.data(), cb, hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    model_trained = true;
    if (warn_maxit)
        return err->warn(da_status_maxit, "The maximum number of iterations was reached.");
    return status;
}

template <typename T>
da_status KMeansHandle<T>::predict(da_int k_samples, da_int k_features, const T *Y, da_int ldy,
                                   da_int *Y_labels) {
    if (!model_trained)
        return err->warn(da_status_no_data, "The k-means has not been computed. Please call "
                                            "da_kmeans_compute_s or da_kmeans_compute_d.");
    if (k_features != n_features_)
        return err->error(da_status_invalid_input,
                          "The function was called with k_features = " + std::to_string(k_features) +
                              " but the k-means has been computed with " +
                              std::to_string(n_features_) + " features.");
    da_status s = check_2D_array<T>(order(), k_samples, k_features, Y, ldy, "k_samples",
                                    "k_features", "Y", "ldy");
    if (s != da_status_success)
        return s;
    if (Y_labels == nullptr)
        return err->error(da_status_invalid_pointer, "The array Y_labels is null.");
    hipStream_t st = gpu->stream();
    const size_t bytes = (size_t)k_samples * k_features * sizeof(T);
    DA_HIP(err, q_.reserve(bytes));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(q_.ptr(), k_features * sizeof(T), Y, (size_t)ldy * sizeof(T),
                                     k_features * sizeof(T), k_samples, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, q_stage_.reserve(bytes));
        DA_HIP(err, copy_2d(q_stage_.ptr(), k_samples * sizeof(T), Y, (size_t)ldy * sizeof(T),
                                     k_samples * sizeof(T), k_features, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(q_stage_.as<T>(), k_samples, k_samples,
                                                   k_features, q_.as<T>(), st));
    }
    DA_HIP(err, out_.reserve((size_t)k_samples * sizeof(int)));
    // Spherical k-means: the largest y.c (the reference's -C Y^T with zero centre norms,
    // kmeans.cpp:975-985).
    DA_HIP(err, launch_kmeans_assign<T>(q_.as<T>(), k_samples, k_features, Cbest_.as<double>(),
                                        n_clusters_, out_.as<int>(), nullptr, nullptr, nullptr, st,
                                        sph_));
    std::vector<int> lab(k_samples);
    DA_HIP(err, hipMemcpyAsync(lab.data(), out_.ptr(), k_samples * sizeof(int), hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    for (da_int i = 0; i < k_samples; i++)
        Y_labels[i] = lab[i];
    return da_status_success;
}

template <typename T>
da_status KMeansHandle<T>::transform(da_int m_samples, da_int m_features, const T *X, da_int ldx,
                                     T *X_transform, da_int ldx_transform) {
    if (!model_trained)
        return err->warn(da_status_no_data, "The k-means has not been computed. Please call "
                                            "da_kmeans_compute_s or da_kmeans_compute_d.");
    if (m_features != n_features_)
        return err->error(da_status_invalid_input,
                          "The function was called with m_features = " + std::to_string(m_features) +
                              " but the k-means has been computed with " +
                              std::to_string(n_features_) + " features.");
    da_status s = check_2D_array<T>(order(), m_samples, m_features, X, ldx, "m_samples",
                                    "m_features", "X", "ldx");
    if (s != da_status_success)
        return s;
    s = check_2D_array<T>(order(), m_samples, n_clusters_, X_transform, ldx_transform, "m_samples",
                          "n_clusters", "X_transform", "ldx_transform");
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const size_t bytes = (size_t)m_samples * m_features * sizeof(T);
    DA_HIP(err, q_.reserve(bytes));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(q_.ptr(), m_features * sizeof(T), X, (size_t)ldx * sizeof(T),
                                     m_features * sizeof(T), m_samples, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, q_stage_.reserve(bytes));
        DA_HIP(err, copy_2d(q_stage_.ptr(), m_samples * sizeof(T), X, (size_t)ldx * sizeof(T),
                                     m_samples * sizeof(T), m_features, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(q_stage_.as<T>(), m_samples, m_samples,
                                                   m_features, q_.as<T>(), st));
    }
    const int k = n_clusters_;
    const size_t ob = (size_t)m_samples * k * sizeof(T);
    DA_HIP(err, out_.reserve(2 * ob));
    // Spherical k-means: cosine distances 1 - x.c/|x| (1 - x.c when the data are taken as
    // unit rows), kmeans.cpp:821-863.
    const double *qinv = nullptr;
    if (sph_ && sph_norm_) {
        DA_HIP(err, qinv_.reserve((size_t)m_samples * sizeof(double)));
        DA_HIP(err, launch_row_inv_norms<T>(q_.as<T>(), m_samples, m_features, qinv_.as<double>(),
                                            st));
        qinv = qinv_.as<double>();
    }
    DA_HIP(err, launch_kmeans_transform<T>(q_.as<T>(), m_samples, m_features, Cbest_.as<double>(),
                                           k, out_.as<T>(), k, st, sph_, qinv));
    if (order() == row_major) {
        DA_HIP(err, copy_2d(X_transform, (size_t)ldx_transform * sizeof(T), out_.ptr(),
                                     k * sizeof(T), k * sizeof(T), m_samples, hipMemcpyDeviceToHost, st));
    } else {
        T *tr = out_.as<T>() + (size_t)m_samples * k;
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(out_.as<T>(), k, k, m_samples, tr, st));
        DA_HIP(err, copy_2d(X_transform, (size_t)ldx_transform * sizeof(T), tr,
                                     m_samples * sizeof(T), m_samples * sizeof(T), k,
                                     hipMemcpyDeviceToHost, st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

// Stored fields of a computed k-means, in the order of kmeans<T>::serialize
// (source/core/clustering/kmeans/kmeans.cpp:1882-1912): the centres column-major k x d in
// the working precision, and their squared norms (the reference's workc1; zero for
// spherical k-means, as its predict step keeps them).
template <typename T> da_status KMeansHandle<T>::serialize(ModelBuffer &b) {
    const bool load = b.loading();
    bool trained = model_trained;
    da_order ord = order();
    da_int algorithm = opts.skey("algorithm"), n_samples = n_samples_, n_features = n_features_,
           n_iter = best_n_iter_, seed = opts.ival("seed"), k = n_clusters_,
           max_iter = opts.ival("max_iter"), n_init = opts.ival("n_init");
    T inertia = (T)best_inertia_, tol = (T)opts.rval("convergence tolerance");
    std::vector<da_int> labels = best_labels_;
    std::vector<T> centres, norms;
    bool sph = sph_, sph_norm = sph_norm_;
    if (!load) {
        centres.resize((size_t)k * n_features);
        norms.assign((size_t)k, (T)0);
        for (da_int c = 0; c < k; c++)
            for (da_int l = 0; l < n_features; l++) {
                const double v = best_centres_[(size_t)c * n_features + l];
                centres[(size_t)l * k + c] = (T)v;
                if (!sph)
                    norms[c] += (T)(v * v);
            }
    }
    da_status s = da_status_success;
    auto io = [&](auto &v) {
        if (s == da_status_success)
            s = b.io(v);
    };
    io(trained);
    io(ord);
    io(algorithm);
    io(n_samples);
    io(n_features);
    io(n_iter);
    io(inertia);
    io(seed);
    io(k);
    io(tol);
    io(max_iter);
    io(n_init);
    io(labels);
    io(centres);
    io(norms);
    io(sph);
    io(sph_norm);
    if (s != da_status_success || !load)
        return s;
    if (n_samples < 0 || n_features <= 0 || k <= 0 || (da_int)labels.size() != n_samples ||
        (int64_t)centres.size() != (int64_t)k * n_features)
        return da_status_invalid_file_data;
    model_trained = trained;
    n_samples_ = n_samples;
    n_features_ = n_features;
    n_clusters_ = k;
    best_n_iter_ = n_iter;
    best_lp_n_iter_ = 0;
    best_inertia_ = (double)inertia;
    best_labels_ = labels;
    best_centres_.assign((size_t)k * n_features, 0.0);
    for (da_int c = 0; c < k; c++)
        for (da_int l = 0; l < n_features; l++)
            best_centres_[(size_t)c * n_features + l] = (double)centres[(size_t)l * k + c];
    sph_ = sph;
    sph_norm_ = sph && sph_norm;
    return da_status_success;
}

template <typename T> da_status KMeansHandle<T>::save_check() {
    if (!model_trained)
        return err->error(da_status_no_data,
                          "k-means clustering has not yet been computed. Please call "
                          "da_kmeans_compute_s or da_kmeans_compute_d before saving the model.");
    return da_status_success;
}

// A loaded k-means predicts and transforms: its centres go to GPU memory.
template <typename T> da_status KMeansHandle<T>::restore() {
    if (!model_trained)
        return da_status_success;
    da_status s = gpu->init(err);
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const size_t cb = best_centres_.size() * sizeof(double);
    DA_HIP(err, Cbest_.reserve(cb));
    DA_HIP(err, hipMemcpyAsync(Cbest_.ptr(), best_centres_.data(), cb, hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T>
da_status KMeansHandle<T>::get_result_typed(da_result query, da_int *dim, T *result) {
    if (!model_trained)
        return err->warn(da_status_no_data, "k-means clustering has not yet been computed. Please "
                                            "call da_kmeans_compute_s or da_kmeans_compute_d "
                                            "before extracting results.");
    const da_int k = n_clusters_, d = n_features_;
    switch (query) {
    case da_rinfo:
        if (*dim < 6) {
            *dim = 6;
            return err->warn(da_status_invalid_array_dimension,
                             "The array is too small. Please provide an array of at least size: 6.");
        }
        result[0] = (T)n_samples_;
        result[1] = (T)n_features_;
        result[2] = (T)k;
        result[3] = (T)best_n_iter_;
        result[4] = (T)best_inertia_;
        result[5] = (T)best_lp_n_iter_;
        return da_status_success;
    case da_kmeans_cluster_centres:
        if (*dim < k * d) {
            *dim = k * d;
            return err->warn(da_status_invalid_array_dimension,
                             "The array is too small. Please provide an array of at least size: " +
                                 std::to_string(k * d) + ".");
        }
        for (da_int i = 0; i < k; i++)
            for (da_int j = 0; j < d; j++) {
                const T v = (T)best_centres_[(size_t)i * d + j];
                if (order() == column_major)
                    result[i + (size_t)j * k] = v;
                else
                    result[(size_t)i * d + j] = v;
            }
        return da_status_success;
    default:
        return err->warn(da_status_unknown_query, "The requested result could not be found.");
    }
}

template <typename T>
da_status KMeansHandle<T>::get_result(da_result query, da_int *dim, double *result) {
    if constexpr (std::is_same_v<T, double>)
        return get_result_typed(query, dim, reinterpret_cast<T *>(result));
    else
        return err->error(da_status_wrong_type, "Wrong precision type.");
}
template <typename T>
da_status KMeansHandle<T>::get_result(da_result query, da_int *dim, float *result) {
    if constexpr (std::is_same_v<T, float>)
        return get_result_typed(query, dim, reinterpret_cast<T *>(result));
    else
        return err->error(da_status_wrong_type, "Wrong precision type.");
}

template <typename T>
da_status KMeansHandle<T>::get_result(da_result query, da_int *dim, da_int *result) {
    if (query == da_trained)
        return AlgHandle::get_result(query, dim, result);
    if (!model_trained)
        return err->warn(da_status_no_data, "k-means clustering has not yet been computed. Please "
                                            "call da_kmeans_compute_s or da_kmeans_compute_d "
                                            "before extracting results.");
    if (query == da_kmeans_labels) {
        if (*dim < n_samples_) {
            *dim = n_samples_;
            return err->warn(da_status_invalid_array_dimension,
                             "The array is too small. Please provide an array of at least size: " +
                                 std::to_string(n_samples_) + ".");
        }
        for (da_int i = 0; i < n_samples_; i++)
            result[i] = best_labels_[i];
        return da_status_success;
    }
    return err->warn(da_status_unknown_query, "The requested result could not be found.");
}

template class KMeansHandle<float>;
template class KMeansHandle<double>;

std::unique_ptr<AlgHandle> make_kmeans_handle(da_precision p, ErrorStack &err, GpuContext &gpu) {
    if (p == da_double)
        return std::make_unique<KMeansHandle<double>>(err, gpu);
    return std::make_unique<KMeansHandle<float>>(err, gpu);
}

} // namespace dagpu
