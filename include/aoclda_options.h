/*
 * Option setting on a handle. Declarations follow source/include/aoclda_options.h of the
 * reference (da_options_set_int/string/real_s/real_d, da_options_get_*, da_options_print).
 * Option names and string values are matched after trimming, squeezing blanks and
 * lower-casing, as source/core/utilities/options.hpp:192-201 does.
 */
#ifndef AOCLDA_OPTIONS_H
#define AOCLDA_OPTIONS_H

#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

da_status da_options_set_int(da_handle handle, const char *option, da_int value);
da_status da_options_set_string(da_handle handle, const char *option, const char *value);
da_status da_options_set_real_s(da_handle handle, const char *option, float value);
da_status da_options_set_real_d(da_handle handle, const char *option, double value);

da_status da_options_get_int(da_handle handle, const char *option, da_int *value);
da_status da_options_get_real_s(da_handle handle, const char *option, float *value);
da_status da_options_get_real_d(da_handle handle, const char *option, double *value);
da_status da_options_get_string(da_handle handle, const char *option, char *value,
                                da_int *lvalue);
da_status da_options_get_string_key(da_handle handle, const char *option, char *value,
                                    da_int *lvalue, da_int *key);
da_status da_options_print(da_handle handle);

#ifdef __cplusplus
}
#endif

#endif
