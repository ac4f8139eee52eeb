/*
 * Opaque handle of the AOCL-DA C API. Enum values follow
 * source/include/aoclda_handle.h:45-88 and the functions :118-150 of the reference.
 * Only da_handle_nn, da_handle_kmeans, da_handle_dbscan and da_handle_approx_nn are backed
 * by an implementation in this library; initialising any other type returns da_status_not_implemented.
 */
#ifndef AOCLDA_HANDLE
#define AOCLDA_HANDLE

#include "aoclda_error.h"
#include "aoclda_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum da_handle_type_ {
    da_handle_uninitialized,
    da_handle_linmod,
    da_handle_pca,
    da_handle_kmeans,
    da_handle_dbscan,
    da_handle_decision_tree,
    da_handle_decision_forest,
    da_handle_nlls,
    da_handle_nn,
    da_handle_svm,
    da_handle_interpolation,
    da_handle_approx_nn,
    da_handle_kernel_pca,
    da_handle_tsne,
};
typedef enum da_handle_type_ da_handle_type;

typedef struct _da_handle *da_handle;

/* aoclda_handle.h:118-119 */
da_status da_handle_init_d(da_handle *handle, da_handle_type handle_type);
da_status da_handle_init_s(da_handle *handle, da_handle_type handle_type);

/* aoclda_handle.h:133 */
da_status da_handle_print_error_message(da_handle handle);

/* aoclda_handle.h:146 */
void da_handle_destroy(da_handle *handle);

/* aoclda_handle.h:149-150 (used by language bindings) */
da_status da_handle_get_error_message(da_handle handle, char **message);
da_status da_handle_get_error_severity(da_handle handle, da_severity *severity);

/* aoclda_handle.h:153 */
void da_handle_refresh(da_handle handle);

/*
 * aoclda_handle.h:177, 201, 217: stored models. Saving writes the handle's options and
 * model to file_name in the reference's stored-model format ("AOCLDA_STORED_MODEL"
 * metadata, option registry, the algorithm's fields); loading creates *handle (which must be
 * NULL) from such a file and rebuilds its GPU-resident state. Handles with models:
 * da_handle_nn (training data, labels, targets, radius-search results), da_handle_kmeans
 * (computed centres and labels) and da_handle_approx_nn (trained and filled index).
 */
da_status da_handle_save_model(da_handle handle, const char *file_name);
da_status da_handle_load_model(da_handle *handle, const char *file_name);
da_status da_handle_print_model_versions(da_handle handle);

#ifdef __cplusplus
}
#endif

#endif
