#include "handle.hpp"
#include <cmath>

namespace dagpu {

AlgHandle::AlgHandle(ErrorStack &e, GpuContext &g) : err(&e), gpu(&g) {
    // Common options, source/core/utilities/basic_handle_options.hpp:40-54.
    opts.add_string("storage order",
                    "Whether data is supplied and returned in row- or column-major order.",
                    {{"row-major", row_major},
                     {"column-major", column_major},
                     {"fortran", column_major},
                     {"f", column_major},
                     {"c", row_major}},
                    "column-major");
    opts.add_string("check data", "Check input data for NaNs prior to performing computation.",
                    {{"yes", 1}, {"no", 0}}, "no");
}

da_status AlgHandle::get_result(da_result, da_int *, double *) {
    return err->warn(da_status_unknown_query, "The requested result could not be found.");
}
da_status AlgHandle::get_result(da_result, da_int *, float *) {
    return err->warn(da_status_unknown_query, "The requested result could not be found.");
}
da_status AlgHandle::get_result(da_result query, da_int *dim, da_int *result) {
    // da_trained is common to every handle (basic_handle.cpp get_result_common).
    if (query == da_trained) {
        if (*dim < 1) {
            *dim = 1;
            return err->warn(da_status_invalid_array_dimension,
                             "Size of the result array is too small, provide an array of at "
                             "least size: 1.");
        }
        result[0] = (da_int)model_trained;
        return da_status_success;
    }
    return err->warn(da_status_unknown_query, "The requested result could not be found.");
}

template <typename T>
da_status AlgHandle::check_2D_array(da_order ord, da_int n_rows, da_int n_cols, const T *data,
                                    da_int ld, const std::string &rows_name,
                                    const std::string &cols_name, const std::string &data_name,
                                    const std::string &ld_name, da_int rows_min,
                                    da_int cols_min) {
    if (n_rows < rows_min)
        return err->error(da_status_invalid_array_dimension,
                          "The function was called with " + rows_name + " = " +
                              std::to_string(n_rows) + ". Constraint: " + rows_name +
                              " >= " + std::to_string(rows_min) + ".");
    if (n_cols < cols_min)
        return err->error(da_status_invalid_array_dimension,
                          "The function was called with " + cols_name + " = " +
                              std::to_string(n_cols) + ". Constraint: " + cols_name +
                              " >= " + std::to_string(cols_min) + ".");
    if (data == nullptr)
        return err->error(da_status_invalid_pointer, "The array " + data_name + " is null.");
    // The leading dimension is validated before the NaN scan reads the array, so that a
    // bad leading dimension never leads to reads outside the caller's array.
    if (ord == column_major && ld < n_rows)
        return err->error(da_status_invalid_leading_dimension,
                          "The function was called with " + rows_name + " = " +
                              std::to_string(n_rows) + " and " + ld_name + " = " +
                              std::to_string(ld) + ". Constraint: " + ld_name +
                              " >= " + rows_name + ".");
    if (ord == row_major && ld < n_cols)
        return err->error(da_status_invalid_leading_dimension,
                          "The function was called with " + cols_name + " = " +
                              std::to_string(n_cols) + " and " + ld_name + " = " +
                              std::to_string(ld) + ". Constraint: " + ld_name +
                              " >= " + cols_name + ".");
    if (check_data()) {
        for (da_int i = 0; i < n_rows; i++)
            for (da_int j = 0; j < n_cols; j++) {
                T v = ord == column_major ? data[i + (size_t)j * ld] : data[(size_t)i * ld + j];
                if (std::isnan(v))
                    return err->error(da_status_invalid_input,
                                      "The array " + data_name + " contains at least one NaN.");
            }
    }
    return da_status_success;
}

template da_status AlgHandle::check_2D_array<float>(da_order, da_int, da_int, const float *,
                                                    da_int, const std::string &,
                                                    const std::string &, const std::string &,
                                                    const std::string &, da_int, da_int);
template da_status AlgHandle::check_2D_array<double>(da_order, da_int, da_int, const double *,
                                                     da_int, const std::string &,
                                                     const std::string &, const std::string &,
                                                     const std::string &, da_int, da_int);

} // namespace dagpu
