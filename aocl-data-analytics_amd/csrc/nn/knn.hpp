// Brute-force k-nearest-neighbour handle (the role of da_neighbors::neighbors<T>,
// source/core/nearest_neighbors/nearest_neighbors.hpp:43-243 of the reference), with the
// training set resident in GPU memory.
#pragma once
#include "../core/handle.hpp"
#include "../kernels/launch.hpp"
#include <functional>
#include <map>
#include <memory>
#include <vector>

namespace dagpu {

// Parameters of a search fixed by the caller (read from the options once).
struct SearchParams {
    int prec = 0;     // operand precision of the fused tile (KP_*)
    int variant = -1; // fused tile kernel variant (-1: tuning table)
    int kc_min = 0;   // minimum candidates kept per (query, split) on the fused tile
    bool seed = false; // seed the fused tile's thresholds (group minima over a row sample)
    // Candidate-list length of the fused tile for k neighbours.
    int tile_kc(int k) const;
};

// Scratch state of one in-flight search: its stream (through a GpuContext) and the device
// buffers of the fused and tile paths. A handle owns one; a serving loop owns several so
// that batches overlap on the device.
struct SearchContext {
    GpuContext *gpu = nullptr;
    std::map<int, int> slots;        // concurrent workgroups per fused-tile configuration
    std::map<int, int> split_factor; // tuned split count per (form, list length, size class)
    // When set, the tuned maps above are read and written in this context instead: contexts
    // that search the same handle from one thread share their tuning.
    SearchContext *tuning = nullptr;
    DeviceBuffer q_norm, keys, sel_key, sel_idx, qpad, part_key, part_idx, part_tau, fb, fb_q,
        fb_ind, fb_dist, fb_bound, tau0, qc;
    // Pinned host word receiving the fallback count of the last fused launch.
    int *h_fb = nullptr;
    da_int last_fallback = 0, last_candidates = 0;
    // Parameters of the last fused launch, and the context of the second fused pass over
    // the queries it left uncertified (created on first use; second_pass marks it).
    SearchParams last_params;
    std::unique_ptr<SearchContext> retry;
    bool second_pass = false;
    ~SearchContext();
};

template <typename T> class NNHandle : public AlgHandle {
  public:
    NNHandle(ErrorStack &err, GpuContext &gpu);
    da_status set_data(da_int n_samples, da_int n_features, const T *X, da_int ldx);
    da_status kneighbors(da_int n_queries, da_int n_features, const T *X_test, da_int ldx_test,
                         da_int *n_ind, T *n_dist, da_int k, da_int return_distance);
    // Device-resident variant: X_test row-major (ld), outputs row-major n_queries x k.
    da_status kneighbors_device(da_int n_queries, da_int n_features, const T *d_X_test,
                                da_int ldx_test, da_int *d_ind, T *d_dist, da_int k,
                                da_int return_distance);
    // Radius neighbours (nearest_neighbors.cpp:1918-2215): results are kept in the handle
    // for da_handle_get_result_? (counts, offsets, indices, distances).
    da_status radius_neighbors(da_int n_queries, da_int n_features, const T *X_test,
                               da_int ldx_test, T radius, da_int return_distance,
                               da_int sort_results);
    // Classification labels and regression targets (nearest_neighbors.cpp:515-566), copied
    // into GPU memory; k-NN or radius-neighbour prediction (:1094-1910).
    da_status set_labels(da_int n_samples, const da_int *y_train);
    da_status set_targets(da_int n_samples, const T *y_train);
    da_status classes(da_int *n_classes, da_int *classes);
    da_status predict_proba(da_int n_queries, da_int n_features, const T *X_test,
                            da_int ldx_test, T *proba, da_int search_mode);
    da_status predict(da_int n_queries, da_int n_features, const T *X_test, da_int ldx_test,
                      da_int *y_test, da_int search_mode);
    da_status predict(da_int n_queries, da_int n_features, const T *X_test, da_int ldx_test,
                      T *y_test, da_int search_mode);
    using AlgHandle::get_result;
    da_status get_result(da_result query, da_int *dim, da_int *result) override;
    da_status get_result(da_result query, da_int *dim, T *result) override;
    da_status serialize(ModelBuffer &b) override;
    da_status save_check() override;
    da_status restore() override;
    void refresh() override {}

    // Search entry points shared with the serving loop. dQ: row-major n_features columns;
    // outputs row-major nq x k. search_launch enqueues everything that needs no host
    // decision; search_finish waits, searches again the queries the certification listed,
    // and leaves the results in d_ind/d_dist (enqueued on ctx's stream). seed_in: per-query
    // threshold seeds of the fused tile (in place of the sample seeds).
    SearchParams params() const;
    da_status prepare_search(const SearchParams &p);
    da_status search_launch(SearchContext &ctx, const SearchParams &p, const T *dQ, da_int nq,
                            da_int k, bool return_distance, da_int *d_ind, T *d_dist,
                            bool &fused, const float *seed_in = nullptr);
    da_status search_finish(SearchContext &ctx, const T *dQ, da_int nq, da_int k,
                            bool return_distance, da_int *d_ind, T *d_dist, bool fused);
    bool trained() const { return trained_; }
    da_int n_samples() const { return n_samples_; }
    da_int n_features() const { return n_features_; }
    ErrorStack *errors() { return err; }

    SearchContext main_ctx;
    da_int last_fallback = 0, last_candidates = 0;

  private:
    using RadiusSink =
        std::function<da_status(da_int q0, int rows, const int *off, const int *ind, const T *dist)>;
    da_status upload_queries(da_int n_queries, const T *X_test, da_int ldx_test);
    da_status radius_core(const T *dQ, da_int n_queries, T radius, bool want_dist,
                          bool sort_results, const RadiusSink *sink);
    da_status check_queries(da_int n_queries, da_int n_features, const T *X_test,
                            da_int ldx_test);
    da_status available_classes();
    // Class probabilities (row per query) and, when d_pred is set, labels of the queries in
    // q_, by k-NN or radius search; reports outliers as the reference does.
    da_status classify(da_int n_queries, da_int search_mode, T *d_proba, int *d_pred);
    da_status set_params();
    da_status check_options_update();
    da_status resolve_k(da_int &k);
    da_status search(const T *dQ, da_int nq, da_int k, bool return_distance, da_int *d_ind,
                     T *d_dist);
    da_status search_tiles(SearchContext &ctx, const T *dQ, da_int nq, da_int k,
                           bool return_distance, da_int *d_ind, T *d_dist);
    bool fused_eligible(da_int k) const;
    int split_cap(int kc) const;
    int tile_splits(SearchContext &ctx, int prec, int kc, int variant, da_int nq,
                    const void *qpad, const float *seed, hipStream_t st,
                    const KnnCentring *cen = nullptr);
    da_status prepare_fused(int prec);

    bool trained_ = false;
    // Training rows (column-major) read back from a stored model, until restore() uploads them.
    std::vector<T> stored_X_;
    // Labels / targets: host copies, class list, per-row class index in GPU memory.
    bool has_labels_ = false, has_targets_ = false, classes_ready_ = false;
    std::vector<da_int> y_class_, classes_;
    std::vector<T> y_reg_;
    DeviceBuffer cls_idx_, classes_d_, tgt_, vote_out_, vote_pred_, vote_flag_;
    da_int n_samples_ = 0, n_features_ = 0;
    // Radius results: per-query counts, neighbours in query order, their distances.
    bool rnn_done_ = false, rnn_dist_ = false;
    std::vector<da_int> rnn_count_, rnn_ind_;
    std::vector<T> rnn_d_;
    DeviceBuffer rnn_keys_, rnn_tile_, rnn_cnt_, rnn_off_, rnn_segoff_, rnn_ind_buf_, rnn_key_buf_, rnn_ind2_, rnn_key2_,
        rnn_tmp_;
    // Options as read at set_data (nearest_neighbors.cpp:295-352).
    da_int n_neighbors_ = 5, algo_ = 3, metric_ = 0;
    double p_ = 2.0;
    da_int leaf_size_ = 30;
    DeviceBuffer train_;     // row-major n_samples x n_features, type T
    DeviceBuffer train_sq_;  // squared row norms (double)
    DeviceBuffer train_cos_; // cosine norms (double)
    DeviceBuffer q_, q_stage_, out_ind_, out_dist_;
    // Fused-path state: zero-padded training rows in the tile's operand precision, -|y|^2/2
    // per row, the largest row norm.
    DeviceBuffer ypad_, nyh_;
    int kp_ = 0;
    int ypad_prec_ = -1; // operand precision the padded copy was built for
    double ymax_ = 0;
    // Centred fp16 form: the stored rows are residuals about the centres of cen_G_ groups
    // (a k-means of the training rows), kept in group order, each group padded to whole
    // 32-row steps (cen_rows_ stored rows). Group centres (G x d, double), the group of each
    // step, the original index of each stored row, the largest residual and centre norms.
    DeviceBuffer cen_C_, cen_step_, cen_perm_;
    int cen_G_ = 0, cen_rows_ = 0;
    double rmax_ = 0, cmax_ = 0;
    da_status prepare_centred();
    // Threshold-seed sample (training sets of at least 131,072 rows): seed_rows_ rows taken
    // at an even stride through the training rows, in the tile layout of the seed form
    // seed_prec_ (the tile's form, plain fp16 for the centred form), and their -|y|^2/2.
    DeviceBuffer seed_y_, seed_nyh_;
    int seed_rows_ = 0, seed_prec_ = -1;
    da_status prepare_seed_sample(int prec, const std::vector<double> &sq);
    KnnKeyError key_error_model(int prec) const;
    // Rows the fused tile sweeps: the padded stored rows of the centred form, else n_samples.
    int tile_rows(int prec) const { return prec == KP_F16C ? cen_rows_ : (int)n_samples_; }
    bool fused_ready_ = false;
    int tile_prec() const;
};

} // namespace dagpu
