// DBSCAN clustering handle (the role of da_dbscan::dbscan<T>,
// source/core/clustering/dbscan/dbscan.hpp:43-111 of the reference), with the samples
// resident in GPU memory and every stage, from the eps-neighbourhoods to the labels, on the
// GPU.
#pragma once
#include "../core/handle.hpp"
#include <vector>

namespace dagpu {

template <typename T> class DBSCANHandle : public AlgHandle {
  public:
    DBSCANHandle(ErrorStack &err, GpuContext &gpu);
    da_status set_data(da_int n_samples, da_int n_features, const T *A, da_int lda);
    da_status compute();
    using AlgHandle::get_result;
    da_status get_result(da_result query, da_int *dim, T *result) override;
    da_status get_result(da_result query, da_int *dim, da_int *result) override;
    void refresh() override;

    // Edges (ordered neighbour pairs) of the last neighbour graph.
    long long last_edges() const { return edges_; }

  private:
    da_status neighbour_graph(int km, T thr);

    bool initdone_ = false;
    da_int n_samples_ = 0, n_features_ = 0, lda_in_ = 0;
    // Options as read by compute (dbscan.cpp:221-233).
    T eps_ = 0, p_ = 2;
    da_int min_samples_ = 5, leaf_size_ = 30;
    // Results.
    std::vector<da_int> labels_;
    std::vector<unsigned char> core_;
    std::vector<int> count_;
    da_int n_core_ = 0, n_clusters_ = 0;
    long long edges_ = 0;
    // Device state: samples (row-major n x d), norms, key tile, neighbour graph (CSR),
    // union-find links and the cluster numbering.
    DeviceBuffer X_, norms_, keys_, tile_, cnt_, boff_, ind_, ind_spare_, offs_, core_d_, parent_,
        is_root_, rid_, labels_d_, tmp_;
};

} // namespace dagpu
