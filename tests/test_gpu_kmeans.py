"""GPU parity of k-means (da_kmeans_*) against the reference's known answers and the oracle.
Labels must be identical (no ties in these data); centres and inertia within 1e-12 relative
(fp64) or 1e-5 relative (fp32)."""
import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps


def _colmajor(v, rows, cols, dtype=np.float64):
    return np.array(v, dtype=dtype).reshape((cols, rows)).T.copy(order="F")


@pytest.mark.parametrize("idx", range(5))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("algorithm", ["lloyd", "elkan"])
def test_known_answers(known, idx, dtype, algorithm):
    c = known["kmeans"][idx]
    n, d, k = c["n_samples"], c["n_features"], c["n_clusters"]
    A = _colmajor(c["A_colmajor"], n, d, dtype)
    init = c.get("init", "supplied")
    C0 = _colmajor(c["C_colmajor"], k, d, dtype) if init == "supplied" else None
    km = aoclda.clustering.kmeans(n_clusters=k, initialization_method=init, C=C0, n_init=1,
                                  max_iter=c["max_iter"], tol=c["tol"], seed=78,
                                  algorithm=algorithm, empty_clusters=c["empty_clusters"])
    km.fit(A)
    eps = np.finfo(dtype).eps
    tol = {"100*sqrt(eps)": 100 * np.sqrt(eps), "10000*sqrt(eps)": 1e4 * np.sqrt(eps),
           "10*eps": 10 * eps}[c["tol_check"]]
    tol = max(tol, 1e-6 if dtype == np.float64 else 2e-5)
    assert km.n_iter == c["rinfo"][3]
    assert abs(km.inertia - c["rinfo"][4]) <= tol
    np.testing.assert_allclose(km.cluster_centres, _colmajor(c["centres_colmajor"], k, d), rtol=0,
                               atol=tol)
    np.testing.assert_array_equal(km.labels, c["labels"])
    X = _colmajor(c["X_colmajor"], c["m_samples"], d, dtype)
    np.testing.assert_allclose(km.transform(X), _colmajor(c["X_transform_colmajor"], c["m_samples"], k),
                               rtol=0, atol=tol)
    Y = _colmajor(c["Y_colmajor"], c["k_samples"], d, dtype)
    np.testing.assert_array_equal(km.predict(Y), c["Y_labels"])


def test_empty_cluster_error(known):
    c = known["kmeans"][2]
    A = _colmajor(c["A_colmajor"], 5, 2)
    C0 = _colmajor(c["C_colmajor"], 2, 2)
    km = aoclda.clustering.kmeans(n_clusters=2, C=C0, max_iter=30, empty_clusters="error")
    with pytest.raises(aoclda.DAError) as e:
        km.fit(A)
    assert e.value.status == L.CODE["empty_clusters"]


def blobs(rng, n_per, centres, sigma=1.0):
    return np.concatenate([c + sigma * rng.standard_normal((n_per, len(c))) for c in centres])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("init", ["supplied", "random", "k-means++", "random partitions"])
def test_vs_oracle(oracle, dtype, init):
    rng = np.random.default_rng(9)
    centres = rng.uniform(-10, 10, (12, 6))
    X = blobs(rng, 150, centres).astype(dtype)
    rng.shuffle(X)
    C0 = X[rng.choice(len(X), 12, replace=False)] if init == "supplied" else None
    km = aoclda.clustering.kmeans(n_clusters=12, initialization_method=init, C=C0, n_init=3,
                                  max_iter=100, tol=1e-6, seed=123)
    km.fit(X)
    r = oracle.kmeans(X.astype(np.float64), 12, init, None if C0 is None else C0.astype(np.float64),
                      n_init=1 if init == "supplied" else 3, max_iter=100, tol=1e-6, seed=123)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)
    np.testing.assert_allclose(km.inertia, r["inertia"], rtol=rtol)


def test_row_major_and_column_major_agree():
    rng = np.random.default_rng(10)
    X = blobs(rng, 100, rng.uniform(-5, 5, (5, 3)))
    a = aoclda.clustering.kmeans(n_clusters=5, seed=7, n_init=2).fit(np.ascontiguousarray(X))
    b = aoclda.clustering.kmeans(n_clusters=5, seed=7, n_init=2).fit(np.asfortranarray(X))
    np.testing.assert_array_equal(a.labels, b.labels)
    np.testing.assert_array_equal(a.cluster_centres, b.cluster_centres)


def test_maxit_warning(oracle):
    rng = np.random.default_rng(11)
    X = rng.standard_normal((500, 4))
    km = aoclda.clustering.kmeans(n_clusters=8, max_iter=1, seed=3, n_init=1)
    km.fit(X)
    assert km.status == L.CODE["maxit"]
    r = oracle.kmeans(X, 8, "k-means++", max_iter=1, seed=3)
    assert r["status"] == 21
    np.testing.assert_array_equal(km.labels, r["labels"])


def test_many_clusters(oracle):
    """kmeans_test_data.hpp:809-858: 2000 points in pairs, 1000 clusters from supplied
    centres (the reference bounds only the inertia, <= 0.01, for this case)."""
    n, d, k = 2000, 2, 1000
    A = np.zeros((n, d))
    for i in range(0, n, 2):
        for j in range(d):
            A[i, j] = i + j * 0.00001
            A[i + 1, j] = i + 0.00001 + j * 0.00001
    C0 = np.array([[2 * i + j * 0.00001 for j in range(d)] for i in range(k)])
    for dtype in (np.float64, np.float32):
        km = aoclda.clustering.kmeans(n_clusters=k, C=C0.astype(dtype), max_iter=50).fit(
            np.asfortranarray(A.astype(dtype)))
        assert km.inertia <= 0.01
        r = oracle.kmeans(A.astype(dtype).astype(np.float64), k, "supplied",
                          C0.astype(dtype).astype(np.float64), max_iter=50)
        assert km.n_iter == r["n_iter"]
        np.testing.assert_array_equal(km.labels, r["labels"])


def test_mixed_precision(known):
    """kmeans_test_data.hpp:944-969: iterative refinement reports the low precision
    iteration count in rinfo[5]."""
    c = known["kmeans"][0]
    A = _colmajor(c["A_colmajor"], 10, 2)
    C0 = _colmajor(c["C_colmajor"], 3, 2)
    km = aoclda.clustering.kmeans(n_clusters=3, C=C0, max_iter=50, mixed_precision=True,
                                  low_precision_max_iter=1)
    km.fit(A)
    assert km.lp_n_iter == 1
    assert km.n_iter == 0
    np.testing.assert_array_equal(km.labels, c["labels"])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_assignment_ties_go_to_lowest_centre(dtype):
    """Points exactly halfway between two centres take the lower centre index (the scalar
    kernel's rule, kmeans_kernels.cpp:209-260); the fp32 MFMA path must re-check them."""
    X = np.array([[0.0, 0.0], [0.0, 1.0], [0.0, -1.0], [-1.0, 0.0], [1.0, 0.0]], dtype=dtype)
    C0 = np.array([[-1.0, 0.0], [1.0, 0.0]], dtype=dtype)
    km = aoclda.clustering.kmeans(n_clusters=2, C=C0, max_iter=1).fit(X)
    # the first three samples are tied in the first assignment and go to centre 0, which
    # becomes (-0.25, 0); centre 1 stays (1, 0)
    np.testing.assert_array_equal(km.labels, [0, 0, 0, 0, 1])
    np.testing.assert_allclose(km.cluster_centres, [[-0.25, 0.0], [1.0, 0.0]])
    # (0.375, 0) is exactly halfway between the final centres
    Y = np.array([[0.0, 5.0], [0.0, -3.0], [0.375, 0.0], [0.5, 0.0]], dtype=dtype)
    np.testing.assert_array_equal(km.predict(Y), [0, 0, 0, 1])


def test_larger_mfma_vs_oracle(oracle):
    """fp32 data large enough for several MFMA workgroups, many clusters, blobs close
    enough that some samples sit near decision boundaries."""
    rng = np.random.default_rng(21)
    centres = rng.uniform(-3, 3, (40, 32))
    X = blobs(rng, 400, centres, sigma=1.0).astype(np.float32)
    rng.shuffle(X)
    C0 = X[:40].copy()
    km = aoclda.clustering.kmeans(n_clusters=40, C=C0, max_iter=30, tol=0.0).fit(X)
    r = oracle.kmeans(X.astype(np.float64), 40, "supplied", C0.astype(np.float64), max_iter=30,
                      tol=0.0)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d,k", [(32, 40), (32, 300), (96, 260), (128, 90), (5, 700)])
def test_mfma_assignment_any_k_vs_oracle(oracle, dtype, d, k):
    """The MFMA assignment in the data's precision (fp64 matrix cores for double data) with
    the centres staged in LDS chunks when they exceed a workgroup's share (k x Kp elements
    above 80 KB: fp64 at (32, 300), (96, 260), (128, 90); fp32 at (96, 260), (5, 700)
    needs no chunk), against the oracle: labels, iteration count, centres and inertia."""
    rng = np.random.default_rng(d * 1000 + k)
    centres = rng.uniform(-4, 4, (k, d))
    X = blobs(rng, max(3, 3000 // k), centres, sigma=1.0).astype(dtype)
    rng.shuffle(X)
    C0 = X[:k].copy()
    km = aoclda.clustering.kmeans(n_clusters=k, C=C0, max_iter=12, tol=0.0).fit(X)
    r = oracle.kmeans(X.astype(np.float64), k, "supplied", C0.astype(np.float64), max_iter=12,
                      tol=0.0)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    tol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=tol, atol=tol)
    assert abs(km.inertia - r["inertia"]) <= tol * max(1.0, abs(r["inertia"]))
    np.testing.assert_array_equal(km.predict(X[: 3 * k]), r["labels"][: 3 * k])


@pytest.mark.parametrize("d", [8, 32, 40, 96])
def test_bf16_cross_term_matches_exact(oracle, d):
    """The bf16 cross term with exact re-check of near ties gives the labels, iteration count
    and centres of the exact computation."""
    rng = np.random.default_rng(30 + d)
    centres = rng.uniform(-20, 20, (24, d))
    X = blobs(rng, 300, centres, sigma=1.5).astype(np.float32)
    rng.shuffle(X)
    C0 = X[:24].copy()
    km = aoclda.clustering.kmeans(n_clusters=24, C=C0, max_iter=20, tol=0.0,
                                  precision="bf16").fit(X)
    r = oracle.kmeans(X.astype(np.float64), 24, "supplied", C0.astype(np.float64), max_iter=20,
                      tol=0.0)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-5, atol=1e-5)
    ref = aoclda.clustering.kmeans(n_clusters=24, C=C0, max_iter=20, tol=0.0).fit(X)
    np.testing.assert_array_equal(km.labels, ref.labels)
    np.testing.assert_array_equal(km.cluster_centres, ref.cluster_centres)


def test_bf16_cross_term_near_ties(oracle):
    """Blobs that overlap, so that many samples are within the bf16 error of a decision
    boundary and go through the exact re-check."""
    rng = np.random.default_rng(41)
    centres = rng.uniform(-1, 1, (30, 16))
    X = blobs(rng, 200, centres, sigma=0.8).astype(np.float32)
    rng.shuffle(X)
    C0 = X[:30].copy()
    km = aoclda.clustering.kmeans(n_clusters=30, C=C0, max_iter=10, tol=0.0,
                                  precision="bf16").fit(X)
    r = oracle.kmeans(X.astype(np.float64), 30, "supplied", C0.astype(np.float64), max_iter=10,
                      tol=0.0)
    np.testing.assert_array_equal(km.labels, r["labels"])
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("precision", ["working", "bf16", "bf16x3"])
def test_batch_matches_single_instances(oracle, precision):
    """Batched k-means over uneven instances equals each instance run alone (oracle with
    supplied centres): labels, iteration counts, centres and inertia, for both operand forms
    of the assignment cross term."""
    rng = np.random.default_rng(50)
    datasets, inits, cfg = [], [], []
    for i in range(14):
        d = int(rng.choice([3, 8, 16, 32, 40, 64]))
        k = int(rng.integers(2, 65))
        n = int(rng.integers(max(k, 50), 3000))
        centres = rng.uniform(-10, 10, (k, d))
        X = (centres[rng.integers(0, k, n)] + rng.standard_normal((n, d))).astype(np.float32)
        datasets.append(X)
        inits.append(X[:k].copy())
        cfg.append((n, d, k))
    for max_iter, tol in ((12, 0.0), (40, 1e-4)):
        cs, labs, n_iter, inertia = aoclda.clustering.kmeans_batch(datasets, inits, max_iter, tol,
                                                                   precision)
        for i, X in enumerate(datasets):
            r = oracle.kmeans(X.astype(np.float64), cfg[i][2], "supplied",
                              inits[i].astype(np.float64), max_iter=max_iter, tol=tol)
            np.testing.assert_array_equal(labs[i], r["labels"], err_msg=f"instance {i} {cfg[i]}")
            assert n_iter[i] == r["n_iter"], (i, cfg[i], n_iter[i], r["n_iter"])
            np.testing.assert_allclose(cs[i], r["centres"], rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(inertia[i], r["inertia"], rtol=1e-5)


def test_batch_argument_errors():
    X = np.zeros((10, 4), dtype=np.float32)
    with pytest.raises(aoclda.DAError):
        aoclda.clustering.kmeans_batch([X], [np.zeros((65, 4), dtype=np.float32)])
    with pytest.raises(aoclda.DAError):
        aoclda.clustering.kmeans_batch([np.zeros((10, 65), dtype=np.float32)],
                                       [np.zeros((2, 65), dtype=np.float32)])
    with pytest.raises(KeyError):
        aoclda.clustering.kmeans_batch([X], [X[:2].copy()], precision="fp8")
