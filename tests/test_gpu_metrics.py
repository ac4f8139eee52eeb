"""GPU parity of da_pairwise_distances against the oracle and the reference's known answers.
Tolerances: fp64 1e-12 relative, fp32 1e-5 relative (with an absolute floor of the same
size times the data scale for entries near zero)."""
import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu

METRICS = ["euclidean", "sqeuclidean", "manhattan", "cosine", "minkowski", "inner_product",
           "euclidean_gemm", "sqeuclidean_gemm"]
RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def close(a, b, dtype, scale=1.0):
    rtol = RTOL[dtype]
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * scale)


@pytest.mark.parametrize("case", ["pairwise_fixed_xy", "pairwise_fixed_xx"])
@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
def test_known_answers(known, case, metric, dtype, order):
    c = known[case]
    X = np.array(c["X_row"], dtype=dtype, order=order)
    Y = np.array(c["Y_row"], dtype=dtype, order=order) if "Y_row" in c else None
    key = {"euclidean": "sqeuclidean", "euclidean_gemm": "sqeuclidean",
           "sqeuclidean_gemm": "sqeuclidean"}.get(metric, metric)
    exp = np.array(c["expected"][key])
    if metric in ("euclidean", "euclidean_gemm"):
        exp = np.sqrt(exp)
    D = aoclda.metrics.pairwise_distances(X, Y, metric=metric, p=c["p"])
    eps = np.finfo(dtype).eps
    np.testing.assert_allclose(D, exp, rtol=0, atol=100 * eps * max(1.0, np.abs(exp).max()))


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("same", [False, True])
def test_vs_oracle(oracle, metric, dtype, order, same):
    rng = np.random.default_rng(42)
    X = rng.uniform(-10, 10, (203, 37)).astype(dtype)
    Y = None if same else rng.uniform(-10, 10, (131, 37)).astype(dtype)
    Xo = np.asarray(X, order=order)
    Yo = None if same else np.asarray(Y, order=order)
    D = aoclda.metrics.pairwise_distances(Xo, Yo, metric=metric, p=3.0)
    E = oracle.pairwise(X.astype(np.float64), None if same else Y.astype(np.float64), metric,
                        p=3.0, eps=np.finfo(dtype).eps)
    if metric in ("euclidean_gemm", "sqeuclidean_gemm"):
        E = np.nan_to_num(E, nan=0.0)  # the GPU clamps the GEMM form at zero
    scale = np.abs(E).max()
    if metric in ("cosine",):
        scale = 1.0
    close(D, E, dtype, scale if metric in ("inner_product", "sqeuclidean_gemm",
                                           "euclidean_gemm") else 1.0)


def test_leading_dimensions(oracle):
    """Sub-arrays with leading dimensions larger than the row/column count."""
    rng = np.random.default_rng(1)
    m, n, k = 17, 13, 9
    big = rng.standard_normal((40, 30))
    lib = L.lib
    X = np.ascontiguousarray(big[:m, :k])
    Y = np.ascontiguousarray(big[20:20 + n, :k])
    E = oracle.pairwise(X, Y, "euclidean")
    # row-major with ldx = ldy = 30 over the big buffer, ldd = n + 5
    Xb = np.ascontiguousarray(big)
    D = np.zeros((m, n + 5))
    assert lib.da_pairwise_distances_d(L.ROW_MAJOR, m, n, k, L.ptr(Xb), 30,
                                       L.ptr(Xb[20:]), 30, L.ptr(D), n + 5, 2.0, 0) == 0
    np.testing.assert_allclose(D[:, :n], E, rtol=1e-12)
    assert np.all(D[:, n:] == 0)
    # column-major with ldx = 40 over the big buffer (rows 0..m of the first k columns)
    Fb = np.asfortranarray(big)
    Dc = np.zeros((m + 3) * n)
    assert lib.da_pairwise_distances_d(L.COLUMN_MAJOR, m, n, k, L.ptr(Fb), 40,
                                       L.ptr(Fb[20:]), 40, L.ptr(Dc), m + 3, 2.0, 0) == 0
    Dm = Dc.reshape((n, m + 3)).T[:m]
    np.testing.assert_allclose(Dm, E, rtol=1e-12)


def test_config_plain_small(oracle):
    """X 2000x64 and Y 1500x64 fp64, iid U[0,1) from numpy default_rng(0) (X drawn first):
    the full 2000x1500 euclidean distance matrix."""
    rng = np.random.default_rng(0)
    X = rng.random((2000, 64))
    Y = rng.random((1500, 64))
    D = aoclda.metrics.pairwise_distances(X, Y)
    E = oracle.pairwise(X, Y, "euclidean")
    np.testing.assert_allclose(D, E, rtol=1e-12, atol=0)
    Dx = aoclda.metrics.pairwise_distances(X)
    assert np.all(np.diag(Dx) == 0)
    np.testing.assert_array_equal(Dx, Dx.T)


def test_zero_data():
    """metric_public.cpp:210-262: all-zero data gives zero distances (cosine: norms -> 1)."""
    X = np.zeros((5, 4))
    for metric in METRICS:
        D = aoclda.metrics.pairwise_distances(X, np.zeros((3, 4)), metric=metric, p=3.0)
        exp = 1.0 if metric == "cosine" else 0.0
        np.testing.assert_array_equal(D, np.full((5, 3), exp))


def test_minkowski_equivalents():
    """metric_public.cpp:264-295: minkowski p=2 is euclidean and p=1 manhattan."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((30, 7))
    Y = rng.standard_normal((20, 7))
    pd = aoclda.metrics.pairwise_distances
    np.testing.assert_array_equal(pd(X, Y, "minkowski", 2.0), pd(X, Y, "euclidean"))
    np.testing.assert_array_equal(pd(X, Y, "minkowski", 1.0), pd(X, Y, "manhattan"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_gemm_form_recheck_near_duplicates(oracle, dtype, metric):
    """Rows of Y are X's rows moved by 1e-7 .. 1e-2 of their norm, far from the origin: the
    GEMM form |x|^2 + |y|^2 - 2 x.y loses those distances to cancellation, so the tile must
    recompute them with direct differences and keep the direct form's relative accuracy."""
    rng = np.random.default_rng(9)
    X = rng.uniform(50, 60, (300, 48))
    scale = 10.0 ** rng.uniform(-7, -2, (300, 1))
    Y = X + scale * rng.standard_normal((300, 48))
    X, Y = X.astype(dtype), Y.astype(dtype)
    D = aoclda.metrics.pairwise_distances(X, Y, metric=metric)
    E = oracle.pairwise(X.astype(np.float64), Y.astype(np.float64), metric)
    np.testing.assert_allclose(D, E, rtol=RTOL[dtype], atol=0)
    # X against itself: an exactly zero diagonal, a symmetric matrix
    Dx = aoclda.metrics.pairwise_distances(X, metric=metric)
    assert np.all(np.diag(Dx) == 0)
    np.testing.assert_array_equal(Dx, Dx.T)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "cosine", "inner_product",
                                    "euclidean_gemm", "sqeuclidean_gemm"])
def test_gemm_form_tile_edges(oracle, dtype, metric):
    """Shapes that are not multiples of the 64 x 64 tile or of the 16-deep k slice, k from 1
    to 133, against the oracle."""
    rng = np.random.default_rng(10)
    for m, n, k in ((1, 1, 1), (65, 63, 3), (130, 71, 17), (64, 128, 133), (200, 9, 64)):
        X = rng.standard_normal((m, k)).astype(dtype)
        Y = rng.standard_normal((n, k)).astype(dtype)
        D = aoclda.metrics.pairwise_distances(X, Y, metric=metric)
        E = oracle.pairwise(X.astype(np.float64), Y.astype(np.float64), metric,
                            eps=np.finfo(dtype).eps)
        if metric in ("euclidean_gemm", "sqeuclidean_gemm"):
            E = np.nan_to_num(E, nan=0.0)
        scale = 1.0 if metric in ("euclidean", "sqeuclidean", "cosine") else max(1.0, np.abs(E).max())
        close(D, E, dtype, scale)
