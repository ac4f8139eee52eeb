"""The allocation-free variant of da_pairwise_distances_s
(da_pairwise_distances_device_noalloc_s: device pointers in and out, a fixed per-GPU scratch
array instead of allocations) returns what da_pairwise_distances_s returns, call after call,
including a second, smaller call that reuses the scratch the first one filled."""
import ctypes as C

import numpy as np
import pytest

from aoclda import _lib as L

pytestmark = pytest.mark.gpu

F = L.lib.da_pairwise_distances_device_noalloc_s
F.restype = C.c_int
F.argtypes = [C.c_int, L.da_int, L.da_int, L.da_int, C.c_void_p, L.da_int, C.c_void_p, L.da_int,
              C.c_void_p, L.da_int, C.c_float, C.c_int]
SCRATCH = 32 << 20


class Dev:
    """A device copy of a host array (and back)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a) if a.flags.c_contiguous else np.asfortranarray(a)
        self.p = C.c_void_p()
        L.check(L.lib.da_amd_malloc(max(self.a.nbytes, 4), C.byref(self.p)))
        L.check(L.lib.da_amd_memcpy_htod(self.p, L.ptr(self.a), self.a.nbytes))

    def get(self):
        out = np.empty_like(self.a)
        L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(out), self.p, out.nbytes))
        return out

    def free(self):
        L.lib.da_amd_free(self.p)


def _host(order, m, n, k, X, ldx, Y, ldy, D, ldd, metric, p=2.0):
    st = L.lib.da_pairwise_distances_s(order, m, n, k, L.ptr(X), ldx,
                                       None if Y is None else L.ptr(Y), ldy, L.ptr(D), ldd,
                                       p, L.METRICS[metric])
    assert st == 0
    return D


def _case(rng, order, m, n, k, metric, pad, x_is_y):
    """Runs both entry points on the same operands (leading dimensions padded by `pad`)
    and returns (noalloc result, host result)."""
    if order == L.ROW_MAJOR:
        ldx, ldy, ldd = k + pad, k + pad, (m if x_is_y else n) + pad
        X = rng.random((m, ldx)).astype(np.float32)
        Y = None if x_is_y else rng.random((n, ldy)).astype(np.float32)
        D = np.full((m, ldd), -7.0, dtype=np.float32)
    else:
        ldx, ldy, ldd = m + pad, n + pad, m + pad
        X = np.asfortranarray(rng.random((ldx, k)).astype(np.float32))
        Y = None if x_is_y else np.asfortranarray(rng.random((ldy, k)).astype(np.float32))
        D = np.asfortranarray(np.full((ldd, m if x_is_y else n), -7.0, dtype=np.float32))
    nn = m if x_is_y else n
    dX, dD = Dev(X), Dev(D)
    dY = None if x_is_y else Dev(Y)
    st = F(order, m, nn, k, dX.p, ldx, None if dY is None else dY.p, ldy, dD.p, ldd, 2.0,
           L.METRICS[metric])
    assert st == 0, L.STATUS.get(st, st)
    got = dD.get()
    for b in (dX, dD, dY):
        if b is not None:
            b.free()
    ref = _host(order, m, nn, k, X, ldx, Y, ldy, D.copy(order="A"), ldd, metric)
    return got, ref


@pytest.mark.parametrize("order,metric,pad,x_is_y", [
    (L.COLUMN_MAJOR, "cosine", 3, False),         # transposed copies and norms in scratch
    (L.ROW_MAJOR, "sqeuclidean_gemm", 5, False),  # strided rows copied, squared norms
    (L.ROW_MAJOR, "euclidean", 0, False),         # compact rows read in place
    (L.COLUMN_MAJOR, "euclidean_gemm", 0, True),  # X against itself: zero diagonal
    (L.ROW_MAJOR, "manhattan", 2, True),
])
def test_matches_host_entry_point_twice(order, metric, pad, x_is_y):
    rng = np.random.default_rng(123)
    # The second call is smaller and follows the first on the same scratch.
    for m, n, k in ((700, 900, 40), (130, 70, 17)):
        got, ref = _case(rng, order, m, n, k, metric, pad, x_is_y)
        np.testing.assert_array_equal(got, ref)
        # padding of D beyond the m x n block is untouched
        if pad:
            assert (got[m:, :] == -7.0).all() if order == L.COLUMN_MAJOR else \
                (got[:, (m if x_is_y else n):] == -7.0).all()


def test_rejects_sizes_beyond_the_scratch():
    """Column-major operands whose copies exceed the 32 MiB scratch are refused (status
    invalid_input) before any work is enqueued."""
    m = n = 4096
    k = SCRATCH // (4 * (m + n)) + 1
    dX, dY, dD = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p, b in ((dX, m * k * 4), (dY, n * k * 4), (dD, m * n * 4)):
        L.check(L.lib.da_amd_malloc(b, C.byref(p)))
    st = F(L.COLUMN_MAJOR, m, n, k, dX, m, dY, n, dD, m, 2.0, L.METRICS["euclidean"])
    assert L.STATUS[st] == "invalid_input"
    for p in (dX, dY, dD):
        L.lib.da_amd_free(p)


def test_rejects_host_pointers():
    X = np.ones((4, 3), dtype=np.float32)
    D = np.zeros((4, 4), dtype=np.float32)
    st = F(L.ROW_MAJOR, 4, 4, 3, L.ptr(X), 3, None, 3, L.ptr(D), 4, 2.0, L.METRICS["euclidean"])
    assert L.STATUS[st] == "invalid_pointer"
