// C entry points of the DBSCAN chapter (aoclda_dbscan.h), wrapped as in
// source/core/clustering/dbscan/dbscan_public.cpp:30-63 of the reference: null handle, log
// clearing, precision check, handle-type check.
#include "aoclda_dbscan.h"
#include "dbscan.hpp"

using namespace dagpu;

template <typename T> static DBSCANHandle<T> *db_of(da_handle h, da_status &s) {
    s = h->check_precision<T>();
    if (s != da_status_success)
        return nullptr;
    auto *db = dynamic_cast<DBSCANHandle<T> *>(h->alg.get());
    if (!db)
        s = h->err.error(da_status_invalid_handle_type,
                         "handle was not initialized with handle_type=da_handle_dbscan or handle "
                         "is invalid.");
    return db;
}

#define DA_DB_PROLOGUE(T)                                                                        \
    if (!handle)                                                                                 \
        return da_status_handle_not_initialized;                                                 \
    handle->err.clear();                                                                         \
    da_status s_;                                                                                \
    DBSCANHandle<T> *db = db_of<T>(handle, s_);                                                  \
    if (!db)                                                                                     \
        return s_;

extern "C" {

da_status da_dbscan_set_data_d(da_handle handle, da_int n_samples, da_int n_features,
                               const double *A, da_int lda) {
    DA_DB_PROLOGUE(double)
    return db->set_data(n_samples, n_features, A, lda);
}
da_status da_dbscan_set_data_s(da_handle handle, da_int n_samples, da_int n_features,
                               const float *A, da_int lda) {
    DA_DB_PROLOGUE(float)
    return db->set_data(n_samples, n_features, A, lda);
}
da_status da_dbscan_compute_d(da_handle handle) {
    DA_DB_PROLOGUE(double)
    return db->compute();
}
da_status da_dbscan_compute_s(da_handle handle) {
    DA_DB_PROLOGUE(float)
    return db->compute();
}

} // extern "C"
