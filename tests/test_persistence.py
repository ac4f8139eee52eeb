"""Stored models without a GPU: the argument and metadata checks of da_handle_save_model /
da_handle_load_model / da_print_model_metadata (mirroring the reference's
tests/unit_tests/model_persistence/handle_errors_public.cpp), and the byte format itself: a
k-means model written here byte by byte in the reference's layout
(model_persistence.cpp:108-182 metadata, options.hpp:956-975 registry, kmeans.cpp:1882-1912
fields) is read back field by field through da_handle_get_result."""
import ctypes as C
import struct

import numpy as np
import pytest

from aoclda import _lib as L

st = L.CODE
lib = L.lib


def i64(v):
    return struct.pack("<q", v)


def string(s):
    return i64(len(s)) + s.encode()


def metadata(magic="AOCLDA_STORED_MODEL", int_size=4, version=50301, handle_type=L.HANDLE_KMEANS,
             precision=8, lib_version="5.3.1", pad=50):
    return (string(magic) + i64(int_size) + i64(version) + i64(handle_type) + i64(precision) +
            string(lib_version) + b"\0" * pad)


def load(path):
    h = L.da_handle()
    s = lib.da_handle_load_model(C.byref(h), str(path).encode())
    return s, h


def destroy(h):
    lib.da_handle_destroy(C.byref(h))


def test_save_argument_errors(tmp_path):
    assert lib.da_handle_save_model(None, str(tmp_path / "m.bin").encode()) == st["invalid_pointer"]
    h = L.da_handle()
    assert lib.da_handle_init_d(C.byref(h), L.HANDLE_KMEANS) == 0
    assert lib.da_handle_save_model(h, None) == st["invalid_pointer"]
    # nothing computed yet: kmeans.cpp:1914-1921
    assert lib.da_handle_save_model(h, str(tmp_path / "m.bin").encode()) == st["no_data"]
    destroy(h)
    h = L.da_handle()
    assert lib.da_handle_init_s(C.byref(h), L.HANDLE_NN) == 0
    assert lib.da_handle_save_model(h, str(tmp_path / "m.bin").encode()) == st["no_data"]
    destroy(h)
    h = L.da_handle()
    assert lib.da_handle_init_d(C.byref(h), L.HANDLE_APPROX_NN) == 0
    assert lib.da_handle_save_model(h, str(tmp_path / "m.bin").encode()) == st["no_data"]
    destroy(h)


def test_load_argument_errors(tmp_path):
    assert lib.da_handle_load_model(None, str(tmp_path / "x.bin").encode()) == st["invalid_pointer"]
    h = L.da_handle()
    assert lib.da_handle_load_model(C.byref(h), None) == st["invalid_pointer"]
    assert not h
    # an initialised handle is refused
    assert lib.da_handle_init_d(C.byref(h), L.HANDLE_KMEANS) == 0
    assert lib.da_handle_load_model(C.byref(h), str(tmp_path / "x.bin").encode()) == \
        st["invalid_pointer"]
    destroy(h)
    s, h = load(tmp_path / "this_file_does_not_exist_12345.bin")
    assert s == st["io_error"] and not h
    (tmp_path / "empty.bin").write_bytes(b"")
    s, h = load(tmp_path / "empty.bin")
    assert s == st["io_error"] and not h
    (tmp_path / "small.bin").write_bytes(b"\x01\x02\x03")
    s, h = load(tmp_path / "small.bin")
    assert s == st["invalid_file_data"] and not h


@pytest.mark.parametrize("kw,expect,created", [
    (dict(int_size=16), "invalid_file_data", False),      # neither 4 nor 8
    (dict(int_size=8), "invalid_file_data", False),       # int64 model into a 32-bit build
    (dict(version=90000), "version_mismatch", True),      # accepted metadata, refused by the handle
    (dict(version=100), "version_mismatch", False),       # older than the oldest accepted
    (dict(precision=2), "invalid_file_data", False),
    (dict(magic="WRONG_MAGIC_KEYWORD"), "invalid_file_data", False),
])
def test_corrupt_metadata(tmp_path, kw, expect, created):
    f = tmp_path / "corrupt.bin"
    f.write_bytes(metadata(**kw))
    s, h = load(f)
    assert s == st[expect]
    assert bool(h) == created
    if h:
        destroy(h)


def test_wrong_magic_length(tmp_path):
    # handle_errors_public.cpp Test 6: the keyword's stored length is 5, not 19
    buf = i64(5) + b"AOCLDA_STORED_MODEL" + i64(4) + i64(50301) + i64(3) + i64(8) + b"\0" * 50
    (tmp_path / "c.bin").write_bytes(buf)
    s, h = load(tmp_path / "c.bin")
    assert s == st["invalid_file_data"] and not h


def test_print_errors(tmp_path):
    assert lib.da_handle_print_model_versions(None) == st["invalid_pointer"]
    h = L.da_handle()
    assert lib.da_handle_init_d(C.byref(h), L.HANDLE_KMEANS) == 0
    assert lib.da_handle_print_model_versions(h) == st["invalid_input"]
    destroy(h)
    assert lib.da_print_model_metadata(None) == st["invalid_pointer"]
    assert lib.da_print_model_metadata(str(tmp_path / "none.bin").encode()) == st["io_error"]
    (tmp_path / "empty.bin").write_bytes(b"")
    assert lib.da_print_model_metadata(str(tmp_path / "empty.bin").encode()) == \
        st["invalid_file_data"]
    (tmp_path / "small.bin").write_bytes(b"\x01\x02\x03")
    assert lib.da_print_model_metadata(str(tmp_path / "small.bin").encode()) == \
        st["invalid_file_data"]
    (tmp_path / "ok.bin").write_bytes(metadata())
    assert lib.da_print_model_metadata(str(tmp_path / "ok.bin").encode()) == 0


def reference_kmeans_file(centres, labels, n_iter, inertia, precision=8, options=()):
    """A stored k-means in the reference's byte layout: metadata, the option registry (name,
    value pairs), then the fields of kmeans<T>::serialize with the centres column-major."""
    k, d = centres.shape
    flt = "<d" if precision == 8 else "<f"
    reg = i64(len(options))
    for name, kind, v in options:
        reg += string(name)
        reg += i64(v) if kind == "int" else (struct.pack(flt, v) if kind == "real" else string(v))
    f = b""
    f += b"\x01"                       # model_trained
    f += i64(0)                        # order (row-major)
    f += i64(1)                        # algorithm
    f += i64(len(labels))              # n_samples
    f += i64(d)                        # n_features
    f += i64(n_iter)                   # best_n_iter
    f += struct.pack(flt, inertia)     # best_inertia
    f += i64(7)                        # seed
    f += i64(k)                        # n_clusters
    f += struct.pack(flt, 1e-4)        # tol
    f += i64(300)                      # max_iter
    f += i64(1)                        # n_init
    f += i64(len(labels)) + b"".join(i64(int(v)) for v in labels)
    cm = centres.T.reshape(-1)         # column-major k x d
    f += i64(cm.size) + b"".join(struct.pack(flt, float(v)) for v in cm)
    f += i64(k + 8) + b"".join(struct.pack(flt, float(v)) for v in
                               list((centres ** 2).sum(1)) + [0.0] * 8)
    f += b"\x00" + b"\x00"             # do_spherical, normalize_data
    return metadata(handle_type=L.HANDLE_KMEANS, precision=precision, pad=0) + reg + f


@pytest.mark.parametrize("precision", [8, 4])
def test_reference_layout_kmeans(tmp_path, precision):
    """A model laid out by the reference's rules loads, field by field: the results read
    back are the stored ones, the centres transposed from the stored column-major order."""
    rng = np.random.default_rng(1)
    centres = rng.standard_normal((3, 4))
    if precision == 4:
        centres = centres.astype(np.float32).astype(np.float64)
    labels = [0, 2, 1, 1, 0, 2, 2]
    opts = [("n_clusters", "int", 3), ("max_iter", "int", 300),
            ("convergence tolerance", "real", 1e-4), ("algorithm", "string", "lloyd"),
            ("storage order", "string", "row-major")]
    path = tmp_path / "ref.bin"
    path.write_bytes(reference_kmeans_file(centres, labels, 11, 2.5, precision, opts))
    s, h = load(path)
    # Without a GPU the centres cannot be made resident: the fields are read all the same.
    assert s in (0, st["arch_not_supported"])
    assert h
    dt = np.float64 if precision == 8 else np.float32
    sfx = "d" if precision == 8 else "s"
    get = getattr(lib, f"da_handle_get_result_{sfx}")
    get.argtypes = [L.da_handle, C.c_int, C.POINTER(L.da_int), C.c_void_p]
    out = np.zeros(12, dt)
    dim = L.da_int(12)
    assert get(h, L.KMEANS_CENTRES, C.byref(dim), out.ctypes.data_as(C.c_void_p)) == 0
    np.testing.assert_array_equal(out.reshape(3, 4), centres.astype(dt))
    info = np.zeros(6, dt)
    dim = L.da_int(6)
    assert get(h, L.RINFO, C.byref(dim), info.ctypes.data_as(C.c_void_p)) == 0
    assert list(info[:5]) == [7, 4, 3, 11, 2.5]
    lab = np.zeros(7, np.int32)
    dim = L.da_int(7)
    geti = lib.da_handle_get_result_int
    geti.argtypes = [L.da_handle, C.c_int, C.POINTER(L.da_int), C.c_void_p]
    assert geti(h, L.KMEANS_LABELS, C.byref(dim), lab.ctypes.data_as(C.c_void_p)) == 0
    assert list(lab) == labels
    v = L.da_int(0)
    assert lib.da_options_get_int(h, b"n_clusters", C.byref(v)) == 0 and v.value == 3
    assert lib.da_handle_print_model_versions(h) == 0
    destroy(h)


def test_unknown_option_refused(tmp_path):
    rng = np.random.default_rng(2)
    centres = rng.standard_normal((2, 2))
    path = tmp_path / "bad.bin"
    path.write_bytes(reference_kmeans_file(centres, [0, 1], 3, 1.0, 8,
                                           [("no such option", "int", 1)]))
    s, h = load(path)
    assert s == st["invalid_file_data"] and h
    destroy(h)


def test_truncated_fields(tmp_path):
    rng = np.random.default_rng(3)
    data = reference_kmeans_file(rng.standard_normal((2, 3)), [0, 1, 1], 3, 1.0, 8)
    (tmp_path / "t.bin").write_bytes(data[:-20])
    s, h = load(tmp_path / "t.bin")
    assert s == st["invalid_file_data"] and h
    destroy(h)
