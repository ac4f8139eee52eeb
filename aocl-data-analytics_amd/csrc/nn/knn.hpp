// Brute-force k-nearest-neighbour handle (the role of da_neighbors::neighbors<T>,
// source/core/nearest_neighbors/nearest_neighbors.hpp:43-243 of the reference), with the
// training set resident in GPU memory.
#pragma once
#include "../core/handle.hpp"

namespace dagpu {

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
    da_status get_result(da_result query, da_int *dim, da_int *result) override;
    void refresh() override {}

    da_int last_fallback = 0, last_candidates = 0;

  private:
    da_status set_params();
    da_status check_options_update();
    da_status resolve_k(da_int &k);
    // Runs the search for device queries (row-major, ld = n_features) into device outputs
    // (row-major, leading dimension k).
    da_status search(const T *dQ, da_int nq, da_int k, bool return_distance, da_int *d_ind,
                     T *d_dist);
    da_status search_tiles(const T *dQ, da_int nq, da_int k, bool return_distance,
                           da_int *d_ind, T *d_dist);
    // Fused MFMA path for the squared-euclidean family (fp32 data).
    bool fused_eligible(da_int k) const;
    da_status search_fused(const T *dQ, da_int nq, da_int k, bool return_distance,
                           da_int *d_ind, T *d_dist);
    da_status prepare_fused(int prec);

    bool trained_ = false;
    da_int n_samples_ = 0, n_features_ = 0;
    // Options as read at set_data (nearest_neighbors.cpp:295-352).
    da_int n_neighbors_ = 5, algo_ = 3, metric_ = 0;
    double p_ = 2.0;
    da_int leaf_size_ = 30;
    DeviceBuffer train_;     // row-major n_samples x n_features, type T
    DeviceBuffer train_sq_;  // squared row norms (double)
    DeviceBuffer train_cos_; // cosine norms (double)
    DeviceBuffer q_, q_stage_, q_norm_, keys_, sel_key_, sel_idx_, out_ind_, out_dist_;
    // Fused-path state: zero-padded training rows (n x Kp, fp32), fp32 squared norms,
    // the largest row norm, and per-call candidate buffers.
    DeviceBuffer ypad_, nyh_, qpad_, part_key_, part_idx_, part_tau_, fb_, fb_q_, fb_ind_,
        fb_dist_;
    int kp_ = 0;
    int ypad_prec_ = -1; // operand precision the padded copy was built for
    double ymax_ = 0;
    bool fused_ready_ = false;
    int tile_prec() const;
};

} // namespace dagpu
