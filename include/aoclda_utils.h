/*
 * Utility functions of the MI355X build of AOCL-DA's distance-bound path. Of the reference's
 * aoclda_utils.h only the stored-model metadata printer belongs to this path.
 */
#ifndef AOCLDA_UTILS_H
#define AOCLDA_UTILS_H

#include "aoclda_error.h"

#ifdef __cplusplus
extern "C" {
#endif

/* aoclda_utils.h:213: prints the metadata block of a stored model file. */
da_status da_print_model_metadata(const char *filename);

#ifdef __cplusplus
}
#endif

#endif
