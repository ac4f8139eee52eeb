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


@pytest.mark.parametrize("d", [12, 40, 96, 128])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle_chunked_rows(oracle, d, order):
    """fp32 rows of whole 16-byte chunks (d a multiple of 4, up to 128) are scanned in
    chunks: one, two, three and four chunks per lane here, including partly used ones."""
    rng = np.random.default_rng(60 + d)
    means = rng.uniform(-6, 6, (24, d))
    X = (means[rng.integers(0, 24, 4000)] + rng.standard_normal((4000, d))).astype(np.float32)
    Q = (means[rng.integers(0, 24, 200)] + rng.standard_normal((200, d))).astype(np.float32)
    ann = aoclda.neighbors.approximate_neighbors(n_list=32, kmeans_iter=8,
                                                 seed=3).train_and_add(np.asarray(X, order=order))
    C = np.asarray(ann.cluster_centroids, dtype=np.float64)
    for n_probe in (1, 5, 32):
        ann.n_probe = n_probe
        for k in (1, 10, 50):
            dist, ind = ann.kneighbors(np.asarray(Q, order=order), n_neighbors=k)
            ei, ed, _ = oracle.ivf_search(C, X, Q, k, n_probe, "sqeuclidean", single=True)
            np.testing.assert_array_equal(ind, ei, err_msg=f"n_probe {n_probe} k {k}")
            np.testing.assert_allclose(dist, ed, rtol=1e-6)
    # every list probed: the exact k nearest neighbours
    ann.n_probe = 32
    dist, ind = ann.kneighbors(Q, n_neighbors=10)
    ei, ed = oracle.knn(X.astype(np.float64), Q.astype(np.float64), 10)
    np.testing.assert_array_equal(ind, ei)


def test_fraction_two_adds_and_padding(oracle):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((2000, 5))
    X2 = rng.standard_normal((700, 5))
    Q = rng.standard_normal((50, 5))
    ann = aoclda.neighbors.approximate_neighbors(n_list=40, n_probe=2, train_fraction=0.5,
                                                 seed=11).train(X)
    ann.add(X).add(X2)
    C = ann.cluster_centroids
    dist, ind = ann.kneighbors(Q, n_neighbors=8)
    ei, ed, sizes = oracle.ivf_search(C, np.vstack([X, X2]), Q, 8, 2)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12)
    np.testing.assert_array_equal(ann.list_sizes, sizes)
    # more neighbours than the probed lists hold: index -1, distance +inf
    small = aoclda.neighbors.approximate_neighbors(n_list=40, n_probe=1, seed=1).train_and_add(X)
    k = 2000
    dist, ind = small.kneighbors(Q[:3], n_neighbors=k)
    ei, ed, _ = oracle.ivf_search(small.cluster_centroids, X, Q[:3], k, 1)
    np.testing.assert_array_equal(ind, ei)
    assert (ind == -1).any() and np.isinf(dist[ind == -1]).all()


def test_errors():
    X = np.ascontiguousarray(np.random.default_rng(1).standard_normal((100, 4)))
    Q = X[:5].copy()
    ind = np.zeros(5 * 3, np.int32)
    dist = np.zeros(5 * 3)
    h = L.Handle(L.HANDLE_APPROX_NN, np.float64)
    h.set_string("storage order", "row-major")
    assert L.lib.da_approx_nn_train_d(h.h) == L.CODE["no_data"]
    assert L.lib.da_approx_nn_set_training_data_d(h.h, 100, 4, L.ptr(X), 4) == 0
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == L.CODE["no_data"]
    h.set_int("n_list", 101)
    assert L.lib.da_approx_nn_train_d(h.h) == L.CODE["invalid_array_dimension"]
    h.set_int("n_list", 4)
    h.set_string("metric", "inner product")
    assert L.lib.da_approx_nn_train_d(h.h) == L.CODE["not_implemented"]
    h.set_string("metric", "sqeuclidean")
    assert L.lib.da_approx_nn_train_d(h.h) == 0
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == L.CODE["no_data"]
    assert L.lib.da_approx_nn_add_d(h.h, 100, 3, L.ptr(X), 4) == L.CODE["invalid_input"]
    assert L.lib.da_approx_nn_add_d(h.h, 100, 4, L.ptr(X), 4) == 0
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 101, 1) == L.CODE["invalid_input"]
    h.set_int("n_probe", 5)
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == L.CODE["invalid_input"]
    h.set_int("n_probe", 2)
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, None, L.ptr(dist), 3, 1) == L.CODE["invalid_pointer"]
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), None, 3, 1) == L.CODE["invalid_pointer"]
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 3, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == L.CODE["invalid_input"]
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == 0
    h.set_int("n_list", 8)
    assert L.lib.da_approx_nn_kneighbors_d(h.h, 5, 4, L.ptr(Q), 4, L.ptr(ind), L.ptr(dist), 3, 1) == L.CODE["option_locked"]
    n = L.da_int(1)
    sizes = np.zeros(1, np.int32)
    h.set_int("n_list", 4)
    assert L.lib.da_handle_get_result_int(h.h, L.ANN_LIST_SIZES, C.byref(n), L.ptr(sizes)) == L.CODE["invalid_array_dimension"] and n.value == 4
    hs = L.Handle(L.HANDLE_APPROX_NN, np.float32)
    assert L.lib.da_approx_nn_train_d(hs.h) == L.CODE["wrong_type"]
