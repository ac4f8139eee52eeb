// Internal layout of the opaque da_handle and the base class of the algorithm objects
// behind it (the role of basic_handle<T> in source/core/utilities/basic_handle.hpp:49-222
// of the reference).
#pragma once
#include "aoclda_handle.h"
#include "aoclda_result.h"
#include "errors.hpp"
#include "gpu.hpp"
#include "options.hpp"
#include "persistence.hpp"
#include <memory>
#include <string>
#include <type_traits>

namespace dagpu {

class AlgHandle {
  public:
    AlgHandle(ErrorStack &err, GpuContext &gpu);
    virtual ~AlgHandle() = default;

    virtual da_status get_result(da_result query, da_int *dim, double *result);
    virtual da_status get_result(da_result query, da_int *dim, float *result);
    virtual da_status get_result(da_result query, da_int *dim, da_int *result);
    // Called when an option changes (basic_handle::refresh).
    virtual void refresh() {}

    // "storage order" option as a da_order value.
    da_order order() const { return (da_order)opts.skey("storage order"); }
    bool check_data() const { return opts.skey("check data") == 1; }

    // Argument checks of a 2-D array as in source/core/utilities/da_utils.cpp:571-630
    // (dimensions, null pointer, leading dimension, then the optional NaN scan).
    template <typename T>
    da_status check_2D_array(da_order ord, da_int n_rows, da_int n_cols, const T *data,
                             da_int ld, const std::string &rows_name,
                             const std::string &cols_name, const std::string &data_name,
                             const std::string &ld_name, da_int rows_min = 1,
                             da_int cols_min = 1);

    // Stored models (basic_handle::serialize, save_model and load_model of the reference,
    // basic_handle.hpp:82-101, basic_handle.cpp:276-325). serialize() visits the
    // algorithm's fields in their stored order, writing them (saving) or reading them back
    // (loading); save_check() refuses a model that cannot be saved yet; restore() rebuilds
    // the GPU-resident state of a loaded model from the fields read back.
    virtual da_status serialize(ModelBuffer &b) {
        (void)b;
        return da_status_not_implemented;
    }
    virtual da_status save_check() { return da_status_success; }
    virtual da_status restore() { return da_status_success; }

    OptionRegistry opts;
    ErrorStack *err;
    GpuContext *gpu;
    bool model_trained = false;
    // Serialization version of this build's models, and the library version a loaded model
    // was saved by ("" until a model is loaded).
    da_int serialization_version = MODEL_MIN_VERSION;
    std::string saved_version;
};

} // namespace dagpu

struct _da_handle {
    da_handle_type handle_type = da_handle_uninitialized;
    da_precision precision = da_unknown;
    dagpu::ErrorStack err;
    dagpu::GpuContext gpu;
    std::unique_ptr<dagpu::AlgHandle> alg;
    std::string last_message;

    template <typename T> da_status check_precision() {
        constexpr da_precision p = std::is_same_v<T, double> ? da_double : da_single;
        if (precision != p) {
            return err.error(da_status_wrong_type,
                             std::string("The handle was initialized with a different "
                                         "precision type than ") +
                                 (std::is_same_v<T, double> ? "double." : "float."));
        }
        return da_status_success;
    }
};
