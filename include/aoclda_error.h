/*
 * Status codes returned by every AOCL-DA call. Numeric values are identical to
 * source/include/aoclda_error.h:38-110 of the reference.
 */
#ifndef AOCLDA_ERROR_H
#define AOCLDA_ERROR_H

enum da_status_ {
    da_status_success = 0,
    da_status_internal_error,
    da_status_memory_error,
    da_status_invalid_pointer,
    da_status_invalid_input,
    da_status_not_implemented,
    da_status_out_of_date,
    da_status_wrong_type,
    da_status_overflow,
    da_status_invalid_handle_type,
    da_status_handle_not_initialized,
    da_status_store_not_initialized,
    da_status_invalid_option,
    da_status_incompatible_options,
    da_status_operation_failed,
    da_status_invalid_leading_dimension,
    da_status_negative_data,
    da_status_invalid_array_dimension,
    da_status_unknown_query,
    da_status_incorrect_output,
    da_status_no_data,
    da_status_maxit,
    da_status_maxtime,
    da_status_numerical_difficulties,
    da_status_bad_derivatives,
    da_status_arch_not_supported,
    da_status_io_error,

    da_status_file_reading_error = 100,
    da_status_parsing_error,
    da_status_missing_data,

    da_status_option_not_found = 400,
    da_status_option_locked,
    da_status_option_wrong_type,
    da_status_option_invalid_bounds,
    da_status_option_invalid_value,

    da_status_optimization_usrstop = 500,
    da_status_optimization_infeasible,
    da_status_optimization_empty_space,

    da_status_missing_block = 600,
    da_status_full_extraction,

    da_status_outlier_warning = 700,

    da_status_serialization_error = 800,
    da_status_invalid_file_data,
    da_status_version_mismatch,

    da_status_empty_clusters = 900,
};
typedef enum da_status_ da_status;

typedef enum da_severity_ {
    DA_NOTSET = 0,
    DA_WARNING = 1,
    DA_ERROR = 2,
} da_severity;

#endif
