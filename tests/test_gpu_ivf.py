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


def _run_case_capi(c, dtype):
    """The reference's test_functionality flow (ann_public.cpp:60-240) through the C API with
    the case's own leading dimensions: train, add, search; then the same data added again and
    twice the neighbours searched. Returns (ind, dist, rinfo, list_sizes, ind2, dist2) with
    the outputs in the case's storage order."""
    h = L.Handle(L.HANDLE_APPROX_NN, dtype)
    h.set_string("storage order", c["order"])
    h.set_string("metric", c["metric"])
    h.set_string("algorithm", c["algorithm"])
    h.set_int("n_list", c["n_list"])
    h.set_int("n_probe", c["n_probe"])
    h.set_int("number of neighbors", c["k"])
    h.set_int("seed", c["seed"])
    h.set_int("k-means_iter", c["kmeans_iter"])
    h.set_real("train fraction", c["train_fraction"])
    Xt = np.asarray(c["X_train"], dtype=dtype)
    Xq = np.asarray(c["X_test"], dtype=dtype)
    n, d, nq, k = c["n_samples"], c["n_features"], c["n_queries"], c["k"]
    h.call("da_approx_nn_set_training_data", n, d, L.ptr(Xt), c["ldx_train"])
    h.call("da_approx_nn_train")
    h.call("da_approx_nn_add", n, d, L.ptr(Xt), c["ldx_train"])
    ind = np.zeros(nq * k, dtype=np.int32)
    dist = np.zeros(nq * k, dtype=dtype)
    h.call("da_approx_nn_kneighbors", nq, d, L.ptr(Xq), c["ldx_test"], L.ptr(ind), L.ptr(dist),
           k, 1)
    rinfo = np.zeros(4, dtype=dtype)
    dim = L.da_int(4)
    h.call("da_handle_get_result", L.RINFO, C.byref(dim), L.ptr(rinfo))
    sizes = np.zeros(c["n_list"], dtype=np.int32)
    dim = L.da_int(c["n_list"])
    L.check(L.lib.da_handle_get_result_int(h.h, L.ANN_LIST_SIZES, C.byref(dim), L.ptr(sizes)),
            h.h)
    h.call("da_approx_nn_add", n, d, L.ptr(Xt), c["ldx_train"])
    ind2 = np.zeros(nq * 2 * k, dtype=np.int32)
    dist2 = np.zeros(nq * 2 * k, dtype=dtype)
    h.call("da_approx_nn_kneighbors", nq, d, L.ptr(Xq), c["ldx_test"], L.ptr(ind2),
           L.ptr(dist2), 2 * k, 1)
    return ind, dist, rinfo, sizes, ind2, dist2


def _qmajor(a, c, k):
    """nq x k view of an output array in the case's storage order."""
    nq = c["n_queries"]
    return a.reshape(k, nq).T if c["order"] == "column-major" else a.reshape(nq, k)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", range(52))
def test_reference_ivf_cases(known, dtype, idx):
    """Every fixed IVF-Flat case of ann_tests.hpp (squared euclidean, euclidean, inner
    product, cosine; zero data, 1 x 1, one list, octants, wide rows, padded leading
    dimensions, subsampled training, outliers): the expected indices exactly and the
    distances within 3000 eps (the reference's own tolerance), the list sizes summing to the
    rows added, and, after the same rows are added again, every neighbour followed by its
    copy at the same distance (ann_public.cpp:200-240)."""
    cases = known["ivf_cases"]
    if idx >= len(cases):
        pytest.skip("fewer cases")
    c = cases[idx]
    k = c["k"]
    ind, dist, rinfo, sizes, ind2, dist2 = _run_case_capi(c, dtype)
    eps = 3e3 * np.finfo(dtype).eps
    np.testing.assert_allclose(dist.astype(np.float64), c["expected_kdist"], rtol=0, atol=eps,
                               err_msg=c["name"])
    if c["expected_kind"] is not None:
        assert ind.tolist() == c["expected_kind"], c["name"]
    assert rinfo[0] == c["n_list"] and rinfo[1] == c["n_samples"] and rinfo[2] == c["n_features"]
    assert sizes.sum() == c["n_samples"]
    I2, D2 = _qmajor(ind2, c, 2 * k), _qmajor(dist2, c, 2 * k)
    D1 = _qmajor(dist, c, k)
    np.testing.assert_allclose(D2[:, ::2], D1, rtol=0, atol=eps)
    np.testing.assert_allclose(D2[:, 1::2], D1, rtol=0, atol=eps)
    if c["expected_kind"] is not None:
        I1 = _qmajor(ind, c, k)
        np.testing.assert_array_equal(np.sort(I2.reshape(-1, k, 2), axis=2)[:, :, 0], I1)
        np.testing.assert_array_equal(np.sort(I2.reshape(-1, k, 2), axis=2)[:, :, 1],
                                      I1 + c["n_samples"])


@pytest.mark.parametrize("metric", ["inner_product", "cosine"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle_ip_cosine(oracle, metric, dtype, order):
    """Inner-product and cosine IVF-Flat against the oracle's add / probe / search on the
    centroids the GPU trained (spherical k-means), rows of varied norms (zero rows included),
    n_probe from one list to all of them, k up to 100."""
    rng = np.random.default_rng(16)
    d = 24
    dirs = rng.standard_normal((20, d))
    X = dirs[rng.integers(0, 20, 5000)] + 0.4 * rng.standard_normal((5000, d))
    X *= rng.uniform(0.3, 3.0, (5000, 1))
    X[11] = 0.0
    Q = dirs[rng.integers(0, 20, 250)] + 0.4 * rng.standard_normal((250, d))
    Q[3] = 0.0
    X, Q = X.astype(dtype), Q.astype(dtype)
    single = dtype == np.float32
    ann = aoclda.neighbors.approximate_neighbors(n_list=24, metric=metric, kmeans_iter=8,
                                                 seed=5).train_and_add(np.asarray(X, order=order))
    C = np.asarray(ann.cluster_centroids, dtype=np.float64)
    # spherical k-means centroids are unit rows (or zero for an empty list)
    nrm = np.linalg.norm(C, axis=1)
    assert np.all((np.abs(nrm - 1) < 1e-5) | (nrm == 0))
    for n_probe in (1, 4, 24):
        ann.n_probe = n_probe
        for k in (1, 10, 100):
            dist, ind = ann.kneighbors(np.asarray(Q, order=order), n_neighbors=k)
            ei, ed, sizes = oracle.ivf_search(C, X, Q, k, n_probe, metric, single=single)
            np.testing.assert_array_equal(ind, ei, err_msg=f"n_probe {n_probe} k {k}")
            np.testing.assert_allclose(dist, ed, rtol=1e-5 if single else 1e-12,
                                       atol=1e-6 if single else 1e-13)
    np.testing.assert_array_equal(ann.list_sizes, sizes)


def test_ip_padding_is_minus_infinity():
    """Fewer candidates than k in the probed lists: index -1, and the distance the
    reference reports, +inf negated for the inner product (-inf) and 1 + inf for cosine."""
    rng = np.random.default_rng(3)
    X = np.concatenate([rng.standard_normal((10, 4)) + 5, rng.standard_normal((10, 4)) - 5])
    for metric, pad in (("inner_product", -np.inf), ("cosine", np.inf)):
        ann = aoclda.neighbors.approximate_neighbors(n_list=2, metric=metric,
                                                     seed=1).train_and_add(X)
        dist, ind = ann.kneighbors(X[:2], n_neighbors=15)
        assert (ind[:, 10:] == -1).all()
        assert (dist[:, 10:] == pad).all()
