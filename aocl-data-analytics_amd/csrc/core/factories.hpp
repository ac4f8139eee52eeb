// Constructors of the algorithm objects that back da_handle_nn, da_handle_kmeans,
// da_handle_dbscan and da_handle_approx_nn.
#pragma once
#include "handle.hpp"

namespace dagpu {
std::unique_ptr<AlgHandle> make_nn_handle(da_precision p, ErrorStack &err, GpuContext &gpu);
std::unique_ptr<AlgHandle> make_kmeans_handle(da_precision p, ErrorStack &err, GpuContext &gpu);
std::unique_ptr<AlgHandle> make_dbscan_handle(da_precision p, ErrorStack &err, GpuContext &gpu);
std::unique_ptr<AlgHandle> make_approx_nn_handle(da_precision p, ErrorStack &err,
                                                 GpuContext &gpu);
} // namespace dagpu
