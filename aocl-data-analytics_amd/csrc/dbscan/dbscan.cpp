#include "dbscan.hpp"
#include "../core/factories.hpp"
#include "../kernels/launch.hpp"
#include "aoclda_metrics.h"
#include <algorithm>
#include <cmath>
#include <limits>

namespace dagpu {

// Algorithm codes of the "algorithm" option (da_neighbors_types::nn_algorithm).
enum { DB_BRUTE = 0, DB_KD_TREE = 1, DB_BALL_TREE = 2, DB_AUTO = 3 };

template <typename T>
DBSCANHandle<T>::DBSCANHandle(ErrorStack &e, GpuContext &g) : AlgHandle(e, g) {
    // Options of source/core/clustering/dbscan/dbscan_options.hpp:41-96.
    const da_int imax = std::numeric_limits<da_int>::max();
    const double fmax = (double)std::numeric_limits<T>::max();
    opts.add_int("min samples", "Minimum number of neighborhood samples for a core point.", 1,
                 LowerBound::greater_equal, imax, UpperBound::p_inf, 5);
    opts.add_int("leaf size", "Leaf size for k-d tree or ball tree.", 1,
                 LowerBound::greater_equal, imax, UpperBound::p_inf, 30);
    opts.add_string("algorithm", "Choice of algorithm.",
                    {{"brute", DB_BRUTE},
                     {"kd tree", DB_KD_TREE},
                     {"ball tree", DB_BALL_TREE},
                     {"auto", DB_AUTO}},
                    "auto");
    opts.add_string("metric", "Choice of metric used to compute pairwise distances.",
                    {{"euclidean", da_euclidean_gemm},
                     {"l2", da_l2},
                     {"sqeuclidean", da_sqeuclidean_gemm},
                     {"manhattan", da_manhattan},
                     {"l1", da_l1},
                     {"cityblock", da_cityblock},
                     {"cosine", da_cosine},
                     {"minkowski", da_minkowski}},
                    "euclidean");
    opts.add_real("eps",
                  "Maximum distance for two samples to be considered in each other's "
                  "neighborhood.",
                  0.0, LowerBound::greater_equal, fmax, UpperBound::p_inf, 1.0e-4);
    opts.add_real("power", "The power of the Minkowski metric used.", 0.0,
                  LowerBound::greater_equal, fmax, UpperBound::p_inf, 2.0);
}

template <typename T> void DBSCANHandle<T>::refresh() {
    // dbscan.cpp:177-188: results are dropped when an option changes; the data stays.
    model_trained = false;
    labels_.clear();
    core_.clear();
    n_core_ = 0;
    n_clusters_ = 0;
}

template <typename T>
da_status DBSCANHandle<T>::set_data(da_int n_samples, da_int n_features, const T *A,
                                    da_int lda) {
    // dbscan.cpp:190-213, with the samples copied into GPU memory (row-major).
    refresh();
    initdone_ = false;
    const da_order ord = order();
    da_status s = check_2D_array<T>(ord, n_samples, n_features, A, lda, "n_samples",
                                    "n_features", "A", "lda");
    if (s != da_status_success)
        return s;
    s = gpu->init(err);
    if (s != da_status_success)
        return s;
    hipStream_t st = gpu->stream();
    const size_t bytes = (size_t)n_samples * n_features * sizeof(T);
    DA_HIP(err, X_.reserve(bytes));
    if (ord == row_major) {
        DA_HIP(err, copy_2d(X_.ptr(), n_features * sizeof(T), A, (size_t)lda * sizeof(T),
                            n_features * sizeof(T), n_samples, hipMemcpyHostToDevice, st));
    } else {
        DA_HIP(err, tmp_.reserve(bytes));
        DA_HIP(err, copy_2d(tmp_.ptr(), n_samples * sizeof(T), A, (size_t)lda * sizeof(T),
                            n_samples * sizeof(T), n_features, hipMemcpyHostToDevice, st));
        DA_HIP(err, launch_colmajor_to_rowmajor<T>(tmp_.as<T>(), n_samples, n_samples,
                                                   n_features, X_.as<T>(), st));
    }
    DA_HIP(err, hipStreamSynchronize(st));
    n_samples_ = n_samples;
    n_features_ = n_features;
    lda_in_ = lda;
    initdone_ = true;
    model_trained = false;
    return da_status_success;
}

// The eps-neighbour graph in CSR form (radius_neighbors.cpp:47-200 builds the same lists):
// per block of rows, the exact key tile against all samples (double accumulation, rounded
// to T) with the count of the keys <= thr other than the point itself taken as the tile is
// written, then a fill of the neighbour indices in ascending order. Leaves count_ (host), offs_ (n + 1 int64) and ind_.
template <typename T> da_status DBSCANHandle<T>::neighbour_graph(int km, T thr) {
    hipStream_t st = gpu->stream();
    const int n = n_samples_, d = n_features_;
    const double *nrm = nullptr;
    if (km == KM_COSINE) {
        DA_HIP(err, norms_.reserve((size_t)n * sizeof(double)));
        DA_HIP(err, launch_row_sqnorms<T>(X_.as<T>(), d, n, d, norms_.as<double>(), st));
        DA_HIP(err, launch_cosine_norms(norms_.as<double>(), n, std::numeric_limits<T>::epsilon(),
                                        st));
        nrm = norms_.as<double>();
    }
    // Row blocks whose key tile stays within 1 GiB (and 2^28 entries): each block costs a
    // host round trip for its counts, so blocks are made large.
    const size_t max_keys = std::min<size_t>(((size_t)1 << 30) / sizeof(T), (size_t)1 << 28);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n, max_keys / (size_t)n));
    const int seg = radius_segment_len(n), nseg = (n + seg - 1) / seg;
    DA_HIP(err, keys_.reserve((size_t)chunk * n * sizeof(T)));
    DA_HIP(err, cnt_.reserve((size_t)chunk * nseg * sizeof(int)));
    DA_HIP(err, tile_.reserve((size_t)chunk * ((n + 63) / 64) * sizeof(int)));
    DA_HIP(err, boff_.reserve((size_t)chunk * nseg * sizeof(int)));
    count_.assign(n, 0);
    std::vector<int> scnt((size_t)chunk * nseg), boff((size_t)chunk * nseg);
    long long E = 0;
    size_t cap = ind_.size() / sizeof(int);
    const T *X = X_.as<T>();
    for (int r0 = 0; r0 < n; r0 += chunk) {
        const int rows = std::min(chunk, n - r0);
        {
            KernelScope ks(*gpu, "pairwise_direct");
            DA_HIP(err, (launch_pairwise_direct_count<T, T>(
                            km, X + (size_t)r0 * d, X, rows, n, d, keys_.as<T>(), n,
                            nrm ? nrm + r0 : nullptr, nrm, (double)p_, thr, r0, tile_.as<int>(),
                            st)));
        }
        {
            KernelScope ks(*gpu, "radius_segsum");
            DA_HIP(err, launch_radius_segsum(tile_.as<int>(), rows, n, cnt_.as<int>(), st));
        }
        DA_HIP(err, hipMemcpyAsync(scnt.data(), cnt_.ptr(), (size_t)rows * nseg * sizeof(int),
                                   hipMemcpyDeviceToHost, st));
        DA_HIP(err, hipStreamSynchronize(st));
        // Offsets of each row segment within the block's part of the neighbour list.
        int tb = 0;
        for (int r = 0; r < rows; r++) {
            int c = 0;
            for (int g = 0; g < nseg; g++) {
                boff[(size_t)r * nseg + g] = tb + c;
                c += scnt[(size_t)r * nseg + g];
            }
            count_[r0 + r] = c;
            tb += c;
        }
        if (tb == 0)
            continue;
        if ((size_t)(E + tb) > cap) {
            // Grow the neighbour list, keeping what is filled.
            const size_t ncap = std::max<size_t>({(size_t)(E + tb), 2 * cap, (size_t)1 << 20});
            DA_HIP(err, ind_spare_.reserve(ncap * sizeof(int)));
            if (E > 0)
                DA_HIP(err, hipMemcpyAsync(ind_spare_.ptr(), ind_.ptr(), (size_t)E * sizeof(int),
                                           hipMemcpyDeviceToDevice, st));
            DA_HIP(err, hipStreamSynchronize(st));
            ind_.swap(ind_spare_);
            ind_spare_.release();
            cap = ncap;
        }
        DA_HIP(err, hipMemcpyAsync(boff_.ptr(), boff.data(), (size_t)rows * nseg * sizeof(int),
                                   hipMemcpyHostToDevice, st));
        KernelScope ks(*gpu, "radius_fill");
        DA_HIP(err, launch_radius_fill<T>(keys_.as<T>(), n, rows, n, thr, r0, boff_.as<int>(),
                                          ind_.as<int>() + E, nullptr, st));
        E += tb;
    }
    edges_ = E;
    std::vector<long long> offs(n + 1);
    offs[0] = 0;
    for (int i = 0; i < n; i++)
        offs[i + 1] = offs[i] + count_[i];
    DA_HIP(err, offs_.reserve((size_t)(n + 1) * sizeof(long long)));
    DA_HIP(err, hipMemcpyAsync(offs_.ptr(), offs.data(), (n + 1) * sizeof(long long),
                               hipMemcpyHostToDevice, st));
    DA_HIP(err, hipStreamSynchronize(st));
    return da_status_success;
}

template <typename T> da_status DBSCANHandle<T>::compute() {
    if (!initdone_)
        return err->error(da_status_no_data, "No data has been passed to the handle. Please call "
                                             "da_dbscan_set_data_s or da_dbscan_set_data_d.");
    // Options (dbscan.cpp:221-233).
    eps_ = (T)opts.rval("eps");
    min_samples_ = opts.ival("min samples");
    leaf_size_ = opts.ival("leaf size");
    p_ = (T)opts.rval("power");
    const da_int algorithm = opts.skey("algorithm");
    const da_int metric = opts.skey("metric");
    model_trained = false;
    labels_.clear();
    core_.clear();
    // Incompatible options (dbscan.cpp:244-253).
    if (algorithm == DB_KD_TREE || algorithm == DB_BALL_TREE) {
        if (metric == da_cosine || metric == da_sqeuclidean_gemm)
            return err->error(da_status_incompatible_options,
                              "Tree algorithms are not compatible with the cosine or squared "
                              "Euclidean distances.");
        if (metric == da_minkowski && p_ < (T)1.0)
            return err->error(da_status_incompatible_options,
                              "Tree algorithms are not compatible with the Minkowski metric "
                              "when p < 1.");
    }
    // Every algorithm ('auto' picks a k-d tree for up to 15 features) is served by the
    // brute-force search: trees return the same neighbourhoods.
    //
    // Key and threshold per metric (radius_neighbors.cpp:63-71): "euclidean" compares squared
    // distances with eps^2; "l2" (da_euclidean) the distance with eps; "sqeuclidean" the
    // squared distance with eps; minkowski with p = 2 is evaluated as the squared distance
    // and, as in the reference, compared with eps itself; p = 1 is manhattan
    // (pairwise_distances.cpp:36-75).
    int km = KM_SQEUCLIDEAN;
    T thr = eps_;
    switch (metric) {
    case da_euclidean_gemm:
        thr = eps_ * eps_;
        break;
    case da_euclidean:
        km = KM_EUCLIDEAN;
        break;
    case da_sqeuclidean_gemm:
        break;
    case da_manhattan:
        km = KM_MANHATTAN;
        break;
    case da_cosine:
        km = KM_COSINE;
        break;
    case da_minkowski:
        if (!(p_ > (T)0))
            return err->error(da_status_invalid_input,
                              "The power of the Minkowski metric must be positive.");
        km = p_ == (T)2.0 ? KM_SQEUCLIDEAN : (p_ == (T)1.0 ? KM_MANHATTAN : KM_MINKOWSKI);
        break;
    default:
        return err->error(da_status_not_implemented, "Unknown metric.");
    }
    da_status s = neighbour_graph(km, thr);
    if (s != da_status_success)
        return s;

    // Core points: at least min_samples - 1 neighbours besides the point itself
    // (dbscan.cpp:459-468).
    const int n = n_samples_;
    hipStream_t st = gpu->stream();
    core_.assign(n, 0);
    n_core_ = 0;
    for (int i = 0; i < n; i++)
        if (count_[i] >= min_samples_ - 1) {
            core_[i] = 1;
            n_core_++;
        }
    DA_HIP(err, core_d_.reserve((size_t)n));
    DA_HIP(err, parent_.reserve((size_t)n * sizeof(int)));
    DA_HIP(err, is_root_.reserve((size_t)n * sizeof(int)));
    DA_HIP(err, rid_.reserve((size_t)n * sizeof(int)));
    DA_HIP(err, labels_d_.reserve((size_t)n * sizeof(int)));
    size_t scan_bytes = 0;
    DA_HIP(err, dbscan_scan_bytes(n, &scan_bytes));
    DA_HIP(err, tmp_.reserve(std::max<size_t>(scan_bytes, 1)));
    DA_HIP(err, hipMemcpyAsync(core_d_.ptr(), core_.data(), n, hipMemcpyHostToDevice, st));
    const long long *offs = offs_.as<long long>();
    const int *ind = ind_.as<int>();
    const unsigned char *core = core_d_.as<unsigned char>();
    {
        KernelScope ks(*gpu, "dbscan_union");
        DA_HIP(err, launch_dbscan_init(parent_.as<int>(), n, st));
        DA_HIP(err, launch_dbscan_union(offs, ind, core, parent_.as<int>(), n, st));
    }
    {
        KernelScope ks(*gpu, "dbscan_label");
        DA_HIP(err, launch_dbscan_flatten(core, parent_.as<int>(), is_root_.as<int>(), n, st));
        DA_HIP(err, launch_dbscan_scan(tmp_.ptr(), scan_bytes, is_root_.as<int>(), rid_.as<int>(),
                                       n, st));
        DA_HIP(err, launch_dbscan_label(offs, ind, core, parent_.as<int>(), rid_.as<int>(),
                                        labels_d_.as<int>(), n, st));
    }
    labels_.assign(n, -1);
    DA_HIP(err, hipMemcpyAsync(labels_.data(), labels_d_.ptr(), (size_t)n * sizeof(int),
                               hipMemcpyDeviceToHost, st));
    DA_HIP(err, hipStreamSynchronize(st));
    da_int mx = -1;
    for (int i = 0; i < n; i++)
        mx = std::max(mx, labels_[i]);
    n_clusters_ = mx + 1;
    model_trained = true;
    return da_status_success;
}

// Results (dbscan.cpp:57-175).
template <typename T>
da_status DBSCANHandle<T>::get_result(da_result query, da_int *dim, T *result) {
    if (!model_trained)
        return err->warn(da_status_no_data,
                         "DBSCAN clustering has not yet been computed. Please call "
                         "da_dbscan_compute_s or da_dbscan_compute_d before extracting results.");
    const da_int rinfo_size = 9;
    if (query != da_rinfo)
        return err->warn(da_status_unknown_query, "The requested result could not be found.");
    if (*dim < rinfo_size) {
        *dim = rinfo_size;
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(rinfo_size) + ".");
    }
    result[0] = (T)n_samples_;
    result[1] = (T)n_features_;
    result[2] = (T)lda_in_;
    result[3] = eps_;
    result[4] = (T)min_samples_;
    result[5] = (T)leaf_size_;
    result[6] = p_;
    result[7] = (T)n_core_;
    result[8] = (T)n_clusters_;
    return da_status_success;
}

template <typename T>
da_status DBSCANHandle<T>::get_result(da_result query, da_int *dim, da_int *result) {
    da_status s = AlgHandle::get_result(query, dim, result);
    if (s != da_status_unknown_query)
        return s;
    if (!model_trained)
        return err->warn(da_status_no_data,
                         "DBSCAN clustering has not yet been computed. Please call "
                         "da_dbscan_compute_s or da_dbscan_compute_d before extracting results.");
    auto too_small = [&](da_int need) {
        *dim = need;
        return err->warn(da_status_invalid_array_dimension,
                         "The array is too small. Please provide an array of at least size: " +
                             std::to_string(need) + ".");
    };
    switch (query) {
    case da_dbscan_labels:
        if (*dim < n_samples_)
            return too_small(n_samples_);
        std::copy(labels_.begin(), labels_.end(), result);
        return da_status_success;
    case da_dbscan_core_sample_indices: {
        if (*dim < n_core_)
            return too_small(n_core_);
        da_int j = 0;
        for (da_int i = 0; i < n_samples_; i++)
            if (core_[i])
                result[j++] = i;
        return da_status_success;
    }
    case da_dbscan_n_clusters:
        *result = n_clusters_;
        return da_status_success;
    case da_dbscan_n_core_samples:
        *result = n_core_;
        return da_status_success;
    default:
        return err->warn(da_status_unknown_query, "The requested result could not be found.");
    }
}

template class DBSCANHandle<float>;
template class DBSCANHandle<double>;

std::unique_ptr<AlgHandle> make_dbscan_handle(da_precision p, ErrorStack &err, GpuContext &gpu) {
    if (p == da_double)
        return std::make_unique<DBSCANHandle<double>>(err, gpu);
    return std::make_unique<DBSCANHandle<float>>(err, gpu);
}

} // namespace dagpu
