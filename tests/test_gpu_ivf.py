"""GPU parity of IVF-Flat approximate neighbours (da_approx_nn_*) against the reference's
fixed cases and the oracle's add / probe / search restatement applied to the centroids the
GPU trained. Indices identical; distances within 1e-12 (fp64) and 1e-6 (fp32) relative."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu


def _case_arrays(c, dtype):
    order = "F" if c["order"] == "column-major" else "C"
    if order == "F":
        Xt = np.array(c["X_train"], dtype=dtype).reshape(c["n_features"], c["n_samples"]).T
        Xq = np.array(c["X_test"], dtype=dtype).reshape(c["n_features"], c["n_queries"]).T
    else:
        Xt = np.array(c["X_train"], dtype=dtype).reshape(c["n_samples"], c["n_features"])
        Xq = np.array(c["X_test"], dtype=dtype).reshape(c["n_queries"], c["n_features"])
    return np.asarray(Xt, order=order), np.asarray(Xq, order=order)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_cases(known, dtype):
    for c in known["ivf"]:
        Xt, Xq = _case_arrays(c, dtype)
        ann = aoclda.neighbors.approximate_neighbors(n_list=c["n_list"], n_probe=c["n_probe"],
                                                     n_neighbors=c["k"]).train_and_add(Xt)
        dist, ind = ann.kneighbors(Xq)
        got = dist.T.ravel() if c["order"] == "column-major" else dist.ravel()
        np.testing.assert_allclose(got, c["kdist"], rtol=1e-6)
        if c["kind"] is not None:
            gi = ind.T.ravel() if c["order"] == "column-major" else ind.ravel()
            assert gi.tolist() == c["kind"]
        r = ann.rinfo
        assert r[0] == c["n_list"] and r[1] == c["n_samples"] and r[2] == c["n_features"]
        assert ann.list_sizes.sum() == c["n_samples"]
        # the same data added twice: every neighbour comes with its copy
        ann.add(Xt)
        dist2, ind2 = ann.kneighbors(Xq, n_neighbors=2 * c["k"])
        np.testing.assert_allclose(np.sort(dist2, axis=1)[:, ::2], np.sort(dist2, axis=1)[:, 1::2])
        assert ann.rinfo[1] == 2 * c["n_samples"]


def test_centroids_match_kmeans():
    """Training is the k-means handle with random initial centres, one initialisation,
    tolerance sqrt(eps) (GetExpectedResults of ann_tests.hpp)."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((3000, 8))
    ann = aoclda.neighbors.approximate_neighbors(n_list=20, kmeans_iter=15, seed=7).train(X)
    km = aoclda.clustering.kmeans(n_clusters=20, initialization_method="random", n_init=1,
                                  max_iter=15, seed=7, tol=np.sqrt(np.finfo(np.float64).eps)).fit(X)
    np.testing.assert_allclose(ann.cluster_centroids, km.cluster_centres, rtol=1e-12, atol=1e-12)
    assert ann.rinfo[3] == km.n_iter


@pytest.mark.parametrize("metric", ["sqeuclidean", "euclidean"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle(oracle, metric, dtype, order):
    rng = np.random.default_rng(6)
    means = rng.uniform(-6, 6, (24, 10))
    X = (means[rng.integers(0, 24, 6000)] + rng.standard_normal((6000, 10))).astype(dtype)
    Q = (means[rng.integers(0, 24, 300)] + rng.standard_normal((300, 10))).astype(dtype)
    single = dtype == np.float32
    rtol = 1e-6 if single else 1e-12
    ann = aoclda.neighbors.approximate_neighbors(n_list=32, metric=metric, kmeans_iter=8,
                                                 seed=3).train_and_add(np.asarray(X, order=order))
    C = np.asarray(ann.cluster_centroids, dtype=np.float64)
    for n_probe in (1, 4, 32):
        ann.n_probe = n_probe
        for k in (1, 10, 100):
            dist, ind = ann.kneighbors(np.asarray(Q, order=order), n_neighbors=k)
            ei, ed, sizes = oracle.ivf_search(C, X, Q, k, n_probe, metric, single=single)
            np.testing.assert_array_equal(ind, ei, err_msg=f"n_probe {n_probe} k {k}")
            np.testing.assert_allclose(dist, ed, rtol=rtol)
    np.testing.assert_array_equal(ann.list_sizes, sizes)


@pytest.mark.parametrize("d", [12, 40, 70, 96, 110, 128])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle_chunked_rows(oracle, d, dtype, order):
    """Rows of whole 16-byte chunks (d a multiple of 4 for fp32, of 2 for fp64, up to 128)
    are scanned in chunks: one to four chunks per lane for fp32, one to eight for fp64,
    including partly used ones; fp32 rows of 70 and 110 features take the general scan."""
    rng = np.random.default_rng(60 + d)
    means = rng.uniform(-6, 6, (24, d))
    X = (means[rng.integers(0, 24, 4000)] + rng.standard_normal((4000, d))).astype(dtype)
    Q = (means[rng.integers(0, 24, 200)] + rng.standard_normal((200, d))).astype(dtype)
    single = dtype == np.float32
    ann = aoclda.neighbors.approximate_neighbors(n_list=32, kmeans_iter=8,
                                                 seed=3).train_and_add(np.asarray(X, order=order))
    C = np.asarray(ann.cluster_centroids, dtype=np.float64)
    for n_probe in (1, 5, 32):
        ann.n_probe = n_probe
        for k in (1, 10, 50):
            dist, ind = ann.kneighbors(np.asarray(Q, order=order), n_neighbors=k)
            ei, ed, _ = oracle.ivf_search(C, X, Q, k, n_probe, "sqeuclidean", single=single)
            np.testing.assert_array_equal(ind, ei, err_msg=f"n_probe {n_probe} k {k}")
            np.testing.assert_allclose(dist, ed, rtol=1e-6 if single else 1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_vs_oracle_many_features(oracle, dtype):
    """Rows of 5,000 features: the list assignment of add() reads centroids wider than its
    LDS stage from global memory."""
    rng = np.random.default_rng(77)
    d = 5000
    means = rng.uniform(-2, 2, (6, d))
    X = (means[rng.integers(0, 6, 400)] + rng.standard_normal((400, d))).astype(dtype)
    Q = (means[rng.integers(0, 6, 20)] + rng.standard_normal((20, d))).astype(dtype)
    single = dtype == np.float32
    ann = aoclda.neighbors.approximate_neighbors(n_list=6, kmeans_iter=5,
                                                 seed=3).train_and_add(X)
    C = np.asarray(ann.cluster_centroids, dtype=np.float64)
    for n_probe in (1, 3):
        ann.n_probe = n_probe
        dist, ind = ann.kneighbors(Q, n_neighbors=5)
        ei, ed, sizes = oracle.ivf_search(C, X, Q, 5, n_probe, "sqeuclidean", single=single)
        np.testing.assert_array_equal(ind, ei, err_msg=f"n_probe {n_probe}")
        np.testing.assert_allclose(dist, ed, rtol=1e-6 if single else 1e-12)
    np.testing.assert_array_equal(ann.list_sizes, sizes)
