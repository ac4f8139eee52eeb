"""ctypes loader of the CPU baseline of the brute-force kNN (oracle/cpu_knn.cpp).

Measurement and test infrastructure only: bench.py's cpu_baseline leg times it on the host
cores beside the GPU path, and tests/ check it against the oracle. The product library never
uses it. `load(native=True)` rebuilds it with -march=native for the host it runs on (into a
temporary directory) and falls back to the prebuilt oracle/_build/libcpu_knn.so.
"""
import ctypes as C
import hashlib
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PREBUILT = os.path.join(HERE, "_build", "libcpu_knn.so")
_cache = {}


def _bind(path):
    lib = C.CDLL(path)
    for name, t in (("cpu_knn_f32", C.c_float), ("cpu_knn_f64", C.c_double)):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = [C.c_int, C.c_int, C.POINTER(t), C.c_int, C.POINTER(t), C.c_int, C.c_int,
                      C.POINTER(C.c_int), C.POINTER(t)]
    lib.cpu_knn_vector_bits.restype = C.c_int
    return lib


def load(native=True):
    """The baseline library: built for this host's CPU (-march=native) when `native` and a
    compiler is there, else the prebuilt one."""
    if native in _cache:
        return _cache[native]
    path = PREBUILT
    if native:
        src = os.path.join(HERE, "cpu_knn.cpp")
        with open(src, "rb") as f:
            tag = hashlib.sha256(f.read()).hexdigest()[:12]
        out = os.path.join(tempfile.gettempdir(), f"aoclda_cpu_knn_{os.getuid()}_{tag}.so")
        if not os.path.exists(out):
            try:
                subprocess.run(["make", "-s", "-C", HERE, f"CPU_OUT={out}", out], check=True,
                               capture_output=True, timeout=180)
            except (OSError, subprocess.SubprocessError):
                out = PREBUILT
        path = out
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", HERE])
        path = PREBUILT
    _cache[native] = _bind(path)
    return _cache[native]


def knn(X_train, X_query, k, threads, native=True):
    """(ind, dist) of the k nearest training rows of every query (euclidean, ascending),
    computed by the restated blocked host algorithm on `threads` OpenMP threads. Inputs are
    copied to column-major, as the reference keeps them."""
    lib = load(native)
    dt = np.float32 if X_train.dtype == np.float32 else np.float64
    Xt = np.asfortranarray(X_train, dtype=dt)
    Xq = np.asfortranarray(X_query, dtype=dt)
    n, d = Xt.shape
    nq = Xq.shape[0]
    ind = np.empty((nq, k), dtype=np.int32)
    dist = np.empty((nq, k), dtype=dt)
    t = C.c_float if dt == np.float32 else C.c_double
    f = lib.cpu_knn_f32 if dt == np.float32 else lib.cpu_knn_f64
    p = lambda a, ty: a.ctypes.data_as(C.POINTER(ty))
    st = f(n, d, p(Xt, t), nq, p(Xq, t), k, int(threads), p(ind, C.c_int), p(dist, t))
    assert st == 0
    return ind, dist


def vector_bits(native=True):
    return load(native).cpu_knn_vector_bits()


def physical_cores():
    """Physical cores of the host (sockets x cores per socket, from lscpu), or None."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    info = {}
    for line in out.splitlines():
        if ":" in line:
            a, b = line.split(":", 1)
            info[a.strip()] = b.strip()
    try:
        return int(info["Socket(s)"]) * int(info["Core(s) per socket"])
    except (KeyError, ValueError):
        return None
