// IVF-Flat approximate nearest neighbours (the role of
// da_approx_nn::approximate_neighbors<T>, source/core/approximate_neighbors/
// approximate_neighbors.hpp:43-193 of the reference): k-means centroids trained on the GPU,
// inverted lists of the added vectors in GPU memory, probed-list search on the GPU.
#pragma once
#include "../core/handle.hpp"
#include <random>
#include <vector>

namespace dagpu {

template <typename T> class ApproxNNHandle : public AlgHandle {
  public:
    ApproxNNHandle(ErrorStack &err, GpuContext &gpu);
    da_status set_training_data(da_int n_samples, da_int n_features, const T *X, da_int ldx);
    da_status train();
    da_status add(da_int n_add, da_int n_features, const T *X, da_int ldx);
    da_status train_and_add();
    da_status kneighbors(da_int n_queries, da_int n_features, const T *X_test, da_int ldx_test,
                         da_int *n_ind, T *n_dist, da_int k, da_int return_distance);
    using AlgHandle::get_result;
    da_status get_result(da_result query, da_int *dim, T *result) override;
    da_status get_result(da_result query, da_int *dim, da_int *result) override;
    da_status serialize(ModelBuffer &b) override;
    da_status save_check() override;
    da_status restore() override;

  private:
    da_status check_options_update();
    da_status upload_rows(const T *X, da_int n, da_int ld, DeviceBuffer &dst, size_t offset_rows);

    // Training data (the caller's pointer, as the reference keeps it; a loaded model points
    // at its own copy, stored_train_).
    const T *X_train_ = nullptr;
    std::vector<T> stored_train_;
    // Added vectors of a loaded model in order of addition (row-major), until restore().
    std::vector<T> stored_rows_;
    da_int ldx_train_ = 0, n_samples_ = 0, n_features_ = 0;
    bool data_set_ = false, data_added_ = false;
    da_order order_ = column_major;
    // Options fixed at train time.
    da_int n_list_ = 1, algo_ = 1, metric_ = 1, kmeans_iter_ = 0;
    std::mt19937 engine_;
    // Index: centroids (n_list x d, row-major, T), list sizes and offsets, the global ids in
    // list order, the added vectors in order of addition.
    std::vector<T> centroids_;
    std::vector<da_int> list_sizes_;
    std::vector<std::vector<int>> lists_;
    da_int n_index_ = 0;
    // Buffers of the MFMA assignment of added vectors: padded rows, padded centroids and
    // their -|c|^2/2, counters, the re-check list and the transposed centroids.
    DeviceBuffer as_x_, as_c_, as_h_, as_small_, as_list_, as_ct_;
    DeviceBuffer C_d_, X_d_, X_spare_, perm_d_, off_d_, q_d_, q_stage_, ckeys_, probe_key_,
        probe_d_, seg_d_, cand_key_, cand_idx_, sel_key_, sel_idx_, out_ind_, out_dist_, lab_d_,
        cdouble_d_;
};

} // namespace dagpu
