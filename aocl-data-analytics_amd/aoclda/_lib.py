"""ctypes binding of libaoclda_amd.so, the C ABI of the MI355X build.

The shared library is loaded from ``../lib`` relative to this file; there is no other
implementation behind these calls, so importing fails loudly when it has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libaoclda_amd.so"))

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C aocl-data-analytics_amd` "
        "(or __graft_entry__.build()).")

lib = C.CDLL(LIB_PATH)

da_int = C.c_int32
da_handle = C.c_void_p
P = C.POINTER

# Status codes (aoclda_error.h).
STATUS = {
    0: "success", 1: "internal_error", 2: "memory_error", 3: "invalid_pointer",
    4: "invalid_input", 5: "not_implemented", 6: "out_of_date", 7: "wrong_type",
    8: "overflow", 9: "invalid_handle_type", 10: "handle_not_initialized",
    11: "store_not_initialized", 12: "invalid_option", 13: "incompatible_options",
    14: "operation_failed", 15: "invalid_leading_dimension", 16: "negative_data",
    17: "invalid_array_dimension", 18: "unknown_query", 19: "incorrect_output",
    20: "no_data", 21: "maxit", 22: "maxtime", 23: "numerical_difficulties",
    24: "bad_derivatives", 25: "arch_not_supported", 26: "io_error",
    700: "outlier_warning", 400: "option_not_found", 401: "option_locked", 402: "option_wrong_type",
    403: "option_invalid_bounds", 404: "option_invalid_value", 900: "empty_clusters",
    800: "serialization_error", 801: "invalid_file_data", 802: "version_mismatch",
}
CODE = {v: k for k, v in STATUS.items()}

# Enums.
ROW_MAJOR, COLUMN_MAJOR = 0, 1
HANDLE_KMEANS, HANDLE_DBSCAN, HANDLE_NN, HANDLE_APPROX_NN = 3, 4, 8, 11
METRICS = {"euclidean": 0, "l2": 0, "sqeuclidean": 1, "minkowski": 2, "manhattan": 3,
           "l1": 3, "cityblock": 3, "cosine": 4, "euclidean_gemm": 5,
           "sqeuclidean_gemm": 6, "inner_product": 7}
RINFO, TRAINED, KMEANS_CENTRES, KMEANS_LABELS = 1, 2, 501, 502
DBSCAN_LABELS, DBSCAN_N_CLUSTERS, DBSCAN_N_CORE, DBSCAN_CORE_INDICES = 503, 504, 505, 506
ANN_CENTROIDS, ANN_LIST_SIZES = 801, 802
(RNN_COUNT, RNN_OFFSETS, RNN_INDICES, RNN_DISTANCES, RNN_INDICES_INDEX,
 RNN_DISTANCES_INDEX) = range(601, 607)


def _proto(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


S = C.c_int  # da_status as int
for _t, _ct in (("d", C.c_double), ("s", C.c_float)):
    _proto(f"da_handle_init_{_t}", S, P(da_handle), C.c_int)
    _proto(f"da_pairwise_distances_{_t}", S, C.c_int, da_int, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p, da_int, C.c_void_p, da_int, _ct, C.c_int)
    _proto(f"da_nn_set_data_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int)
    _proto(f"da_nn_kneighbors_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p, C.c_void_p, da_int, da_int)
    _proto(f"da_nn_kneighbors_device_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p, C.c_void_p, da_int, da_int)
    _proto(f"da_nn_radius_neighbors_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           _ct, da_int, da_int)
    _proto(f"da_nn_set_labels_{_t}", S, da_handle, da_int, C.c_void_p)
    _proto(f"da_nn_set_targets_{_t}", S, da_handle, da_int, C.c_void_p)
    _proto(f"da_nn_classes_{_t}", S, da_handle, P(da_int), C.c_void_p)
    for _f in ("classifier_predict_proba", "classifier_predict", "regressor_predict"):
        _proto(f"da_nn_{_f}_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int, C.c_void_p,
               C.c_int)
    _proto(f"da_approx_nn_set_training_data_{_t}", S, da_handle, da_int, da_int, C.c_void_p,
           da_int)
    _proto(f"da_approx_nn_train_{_t}", S, da_handle)
    _proto(f"da_approx_nn_train_and_add_{_t}", S, da_handle)
    _proto(f"da_approx_nn_add_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int)
    _proto(f"da_approx_nn_kneighbors_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p, C.c_void_p, da_int, da_int)
    _proto(f"da_dbscan_set_data_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int)
    _proto(f"da_dbscan_compute_{_t}", S, da_handle)
    _proto(f"da_kmeans_set_data_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int)
    _proto(f"da_kmeans_set_init_centres_{_t}", S, da_handle, C.c_void_p, da_int)
    _proto(f"da_kmeans_compute_{_t}", S, da_handle)
    _proto(f"da_kmeans_transform_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p, da_int)
    _proto(f"da_kmeans_predict_{_t}", S, da_handle, da_int, da_int, C.c_void_p, da_int,
           C.c_void_p)
    _proto(f"da_handle_get_result_{_t}", S, da_handle, C.c_int, P(da_int), C.c_void_p)
    _proto(f"da_options_set_real_{_t}", S, da_handle, C.c_char_p, _ct)
    _proto(f"da_options_get_real_{_t}", S, da_handle, C.c_char_p, P(_ct))
_proto("da_handle_get_result_int", S, da_handle, C.c_int, P(da_int), C.c_void_p)
_proto("da_handle_destroy", None, P(da_handle))
_proto("da_handle_get_error_message", S, da_handle, P(C.c_char_p))
_proto("da_handle_print_error_message", S, da_handle)
_proto("da_handle_get_error_severity", S, da_handle, P(C.c_int))
_proto("da_options_set_int", S, da_handle, C.c_char_p, da_int)
_proto("da_options_set_string", S, da_handle, C.c_char_p, C.c_char_p)
_proto("da_options_get_int", S, da_handle, C.c_char_p, P(da_int))
_proto("da_options_get_string", S, da_handle, C.c_char_p, C.c_char_p, P(da_int))
_proto("da_amd_malloc", S, C.c_size_t, P(C.c_void_p))
_proto("da_amd_free", S, C.c_void_p)
_proto("da_amd_memcpy_htod", S, C.c_void_p, C.c_void_p, C.c_size_t)
_proto("da_amd_memcpy_dtoh", S, C.c_void_p, C.c_void_p, C.c_size_t)
_proto("da_amd_memcpy_dtod", S, C.c_void_p, C.c_void_p, C.c_size_t)
_proto("da_amd_device_synchronize", S)
_proto("da_amd_device_name", S, C.c_char_p, da_int)
_proto("da_amd_synchronize", S, da_handle)
_proto("da_amd_profile_enable", S, da_handle, da_int)
_proto("da_amd_profile_reset", S, da_handle)
_proto("da_amd_pairwise_profile_enable", S, da_int)
_proto("da_amd_pairwise_profile_reset", S)
_proto("da_amd_pairwise_profile_get", S, C.c_char_p, P(C.c_double), P(da_int))
_proto("da_amd_profile_get", S, da_handle, C.c_char_p, P(C.c_double), P(da_int))
_proto("da_nn_last_query_stats", S, da_handle, P(da_int), P(da_int))
_proto("da_kmeans_last_compute_stats", S, da_handle, P(da_int))
_proto("da_kmeans_batch_device_s", S, da_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
       C.c_void_p, C.c_void_p, da_int, C.c_float, da_int, C.c_void_p, C.c_void_p)
_proto("da_kmeans_batch_tile_shape", S, P(da_int))
_proto("da_kmeans_batch_release", S)
_proto("da_nn_server_create", S, da_handle, da_int, C.c_size_t, P(C.c_void_p))
_proto("da_nn_server_submit_s", S, C.c_void_p, da_int, da_int, C.c_void_p, da_int, da_int, P(da_int))
_proto("da_nn_server_submit_d", S, C.c_void_p, da_int, da_int, C.c_void_p, da_int, da_int, P(da_int))
_proto("da_nn_server_wait_s", S, C.c_void_p, da_int, C.c_void_p, C.c_void_p, P(C.c_double))
_proto("da_nn_server_wait_d", S, C.c_void_p, da_int, C.c_void_p, C.c_void_p, P(C.c_double))
_proto("da_nn_server_stats", S, C.c_void_p, C.c_void_p, P(da_int))
_proto("da_nn_server_destroy", None, P(C.c_void_p))


class DAError(RuntimeError):
    """A non-success status returned by the C API."""

    def __init__(self, status, message=""):
        self.status = status
        name = STATUS.get(status, str(status))
        super().__init__(f"da_status_{name} ({status}) {message}".strip())


def error_message(handle):
    msg = C.c_char_p()
    if handle and lib.da_handle_get_error_message(handle, C.byref(msg)) == 0 and msg.value:
        text = msg.value.decode()
        C.CDLL(None).free(msg)
        return text
    return ""


def check(status, handle=None, allow=()):
    """Raise DAError unless status is success or in `allow` (warnings)."""
    if status != 0 and status not in allow:
        raise DAError(status, error_message(handle))
    return status


def suffix(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def prepare(a, dtype=None, order=None):
    """Return (array, storage order) for a 2-D input: F-ordered arrays stay column-major,
    everything else is made C-contiguous (row-major)."""
    a = np.asarray(a)
    if dtype is None:
        dtype = np.float32 if a.dtype == np.float32 else np.float64
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    if order is None:
        order = COLUMN_MAJOR if (a.flags.f_contiguous and not a.flags.c_contiguous) else ROW_MAJOR
    if order == COLUMN_MAJOR:
        a = np.asfortranarray(a, dtype=dtype)
    else:
        a = np.ascontiguousarray(a, dtype=dtype)
    return a, order, np.dtype(dtype)


class Handle:
    """Owns a da_handle and destroys it with the object."""

    def __init__(self, handle_type, dtype):
        self.dtype = np.dtype(dtype)
        self.h = da_handle()
        check(getattr(lib, f"da_handle_init_{suffix(self.dtype)}")(C.byref(self.h), handle_type),
              self.h)

    def set_int(self, name, v):
        check(lib.da_options_set_int(self.h, name.encode(), int(v)), self.h)

    def set_string(self, name, v):
        check(lib.da_options_set_string(self.h, name.encode(), v.encode()), self.h)

    def set_real(self, name, v):
        f = getattr(lib, f"da_options_set_real_{suffix(self.dtype)}")
        check(f(self.h, name.encode(), float(v)), self.h)

    def call(self, fname, *args, allow=()):
        f = getattr(lib, f"{fname}_{suffix(self.dtype)}")
        return check(f(self.h, *args), self.h, allow)

    def __del__(self):
        try:
            if self.h:
                lib.da_handle_destroy(C.byref(self.h))
        except Exception:
            pass

# Stored models (aoclda_handle.h:177-217, aoclda_utils.h:213).
_proto("da_handle_save_model", S, da_handle, C.c_char_p)
_proto("da_handle_load_model", S, P(da_handle), C.c_char_p)
_proto("da_handle_print_model_versions", S, da_handle)
_proto("da_print_model_metadata", S, C.c_char_p)
