"""Pairwise distances (python_interface/python_package/aoclda/metrics.py:35 of the reference)."""
import numpy as np

from . import _lib as L


def pairwise_distances(X, Y=None, metric="euclidean", p=2.0):
    """Distance matrix between the rows of X and the rows of Y (or of X when Y is None).

    metric is one of 'euclidean', 'l2', 'sqeuclidean', 'manhattan', 'l1', 'cityblock',
    'cosine', 'minkowski', 'euclidean_gemm', 'sqeuclidean_gemm', 'inner_product'.
    """
    X, order, dtype = L.prepare(X)
    if Y is not None:
        Y, _, _ = L.prepare(Y, dtype=dtype, order=order)
    if metric not in L.METRICS:
        raise ValueError(f"unknown metric {metric!r}")
    m, k = X.shape
    n = m if Y is None else Y.shape[0]
    if Y is not None and Y.shape[1] != k:
        raise ValueError("X and Y must have the same number of columns")
    D = np.empty((m, n), dtype=dtype, order="F" if order == L.COLUMN_MAJOR else "C")
    ldx = m if order == L.COLUMN_MAJOR else k
    ldy = (n if order == L.COLUMN_MAJOR else k)
    ldd = m if order == L.COLUMN_MAJOR else n
    f = getattr(L.lib, f"da_pairwise_distances_{L.suffix(dtype)}")
    L.check(f(order, m, n, k, L.ptr(X), ldx, L.ptr(Y), ldy, L.ptr(D), ldd, p, L.METRICS[metric]))
    return D
