"""GPU parity of the k-means initialisations, Elkan's bounds and mixed-precision refinement
against the reference's own fixtures (tests/unit_tests/kmeans/kmeans_test_data.hpp) and the
oracle, which the same fixtures pin (tests/test_oracle.py)."""
import numpy as np
import pytest

import aoclda
from _kmeans_cases import base_data, check_tol, colmajor

pytestmark = pytest.mark.gpu


def _fit(A, k, dtype, order, **kw):
    A = A.astype(dtype)
    A = np.asfortranarray(A) if order == "column-major" else np.ascontiguousarray(A)
    C = kw.pop("C", None)
    if C is not None:
        C = C.astype(dtype)
        C = np.asfortranarray(C) if order == "column-major" else np.ascontiguousarray(C)
    return aoclda.clustering.kmeans(n_clusters=k, C=C, **kw).fit(A)


def _n_random(known):
    return len(known["kmeans_random"])


@pytest.mark.parametrize("i", range(18))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_random_inits(known, i, dtype):
    """kmeans_public.cpp:272-276: the best inertia of n_init runs of a random initialisation
    (k-means++, random, random partitions, AFK-MC2; Lloyd or Elkan) is within the bound."""
    if i >= _n_random(known):
        pytest.skip("no such case")
    c = known["kmeans_random"][i]
    A, _, order, _ = base_data(known, c["base"])
    km = _fit(A, c["n_clusters"], dtype, order, initialization_method=c["init"],
              n_init=c["n_init"], max_iter=c["max_iter"], seed=c["seed"],
              algorithm=c["algorithm"])
    tol = check_tol(c.get("check_tol", "100*sqrt(eps)"), dtype)
    assert km.inertia <= c["max_inertia"] + tol


@pytest.mark.parametrize("i", range(7))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_fixed_random_inits(known, i, dtype):
    """All-zero data and the 1 x 1 matrix: every initialisation gives the reference's
    rinfo, centres, labels, transform and predict."""
    c = known["kmeans_fixed_random_init"][i]
    n, d, k = c["n_samples"], c["n_features"], c["n_clusters"]
    A = colmajor(c["A_colmajor"], n, d)
    km = _fit(A, k, dtype, "column-major", initialization_method=c["init"],
              n_init=c["n_init"], max_iter=c["max_iter"], tol=c["tol"], seed=c["seed"],
              algorithm=c["algorithm"])
    assert [n, d, k, km.n_iter, km.inertia, km.lp_n_iter] == c["rinfo"]
    np.testing.assert_array_equal(km.cluster_centres,
                                  colmajor(c["centres_colmajor"], k, d).astype(dtype))
    np.testing.assert_array_equal(km.labels, c["labels"])
    X = colmajor(c["X_colmajor"], c["m_samples"], d).astype(dtype)
    # the reference compares the transform within its tolerance: 3.3 - 2.1 is not 1.2 in
    # binary floating point
    np.testing.assert_allclose(km.transform(X),
                               colmajor(c["X_transform_colmajor"], c["m_samples"], k),
                               rtol=1e-12 if dtype == np.float64 else 1e-6, atol=0)
    Y = colmajor(c["Y_colmajor"], c["k_samples"], d).astype(dtype)
    np.testing.assert_array_equal(km.predict(Y), c["Y_labels"])


@pytest.mark.parametrize("i", range(3))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_elkan(known, i, dtype):
    """Elkan's algorithm on the reference's supplied-centre cases gives their answers."""
    c = known["kmeans_elkan"][i]
    A, C0, order, b = base_data(known, c["base"])
    k, d = C0.shape
    km = _fit(A, k, dtype, order, C=C0, max_iter=b["max_iter"], tol=b["tol"],
              empty_clusters=b["empty_clusters"], algorithm="elkan")
    tol = max(check_tol(b["tol_check"], dtype), 1e-6 if dtype == np.float64 else 2e-5)
    assert km.n_iter == b["rinfo"][3]
    assert abs(km.inertia - b["rinfo"][4]) <= tol
    np.testing.assert_allclose(km.cluster_centres, colmajor(b["centres_colmajor"], k, d),
                               rtol=0, atol=tol)
    np.testing.assert_array_equal(km.labels, b["labels"])


@pytest.mark.parametrize("i", range(2))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_mixed_precision(known, oracle, i, dtype):
    """kmeans_test_data.hpp:944-969: the iteration counts of mixed-precision refinement (the
    low-precision phase in float for double data, in half for float data), and the oracle's
    labels, centres and iteration counts."""
    c = known["kmeans_mixed"][i]
    A, C0, order, b = base_data(known, c["base"])
    k = C0.shape[0]
    km = _fit(A, k, dtype, order, C=C0, max_iter=b["max_iter"], tol=b["tol"],
              mixed_precision=True, low_precision_tol=c["lp_tol"],
              low_precision_max_iter=c["lp_max_iter"])
    key = "double" if dtype == np.float64 else "float"
    assert km.lp_n_iter == c["lp_n_iter"]
    assert km.n_iter == c["n_iter"][key]
    np.testing.assert_array_equal(km.labels, b["labels"])
    Ao = A.astype(dtype).astype(np.float64)
    r = oracle.kmeans(Ao, k, "supplied", C0.astype(dtype).astype(np.float64),
                      max_iter=b["max_iter"], tol=b["tol"],
                      mixed_precision={"double": "float", "float": "half"}[key],
                      lp_tol=c["lp_tol"], lp_max_iter=c["lp_max_iter"])
    assert (km.n_iter, km.lp_n_iter) == (r["n_iter"], r["lp_n_iter"])
    np.testing.assert_array_equal(km.labels, r["labels"])


def _blobs(seed, n, d, k, spread=3.0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)) + rng.integers(0, k, size=(n, 1)) * spread
    return X, rng


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,d,k", [(400, 3, 7), (300, 8, 20), (250, 2, 40)])
def test_elkan_vs_oracle(oracle, dtype, n, d, k):
    """Elkan on the GPU follows the oracle's Elkan: labels, iteration count, centres."""
    X, rng = _blobs(n + d + k, n, d, k)
    X = X.astype(dtype).astype(np.float64)
    C0 = X[rng.choice(n, k, replace=False)]
    km = _fit(X, k, dtype, "row-major", C=C0, max_iter=60, tol=1e-6, algorithm="elkan")
    r = oracle.kmeans(X, k, "supplied", C0, max_iter=60, tol=1e-6, algorithm="elkan")
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)
    np.testing.assert_allclose(km.inertia, r["inertia"], rtol=rtol)


@pytest.mark.parametrize("empty", ["split", "error"])
def test_elkan_empty_clusters_vs_oracle(oracle, empty):
    """A centre far from every sample empties at the first iteration: "split" moves the
    furthest sample into it and resets the bounds, "error" fails the run."""
    X, rng = _blobs(5, 200, 2, 4)
    C0 = np.concatenate([X[rng.choice(200, 4, replace=False)], [[1e3, 1e3]]])
    r = oracle.kmeans(X, 5, "supplied", C0, max_iter=40, tol=1e-6, empty_clusters=empty,
                      algorithm="elkan")
    if empty == "error":
        assert r["status"] == 900
        with pytest.raises(aoclda.DAError):
            _fit(X, 5, np.float64, "row-major", C=C0, max_iter=40, tol=1e-6,
                 empty_clusters=empty, algorithm="elkan")
        return
    km = _fit(X, 5, np.float64, "row-major", C=C0, max_iter=40, tol=1e-6, empty_clusters=empty,
              algorithm="elkan")
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("seed,k,m", [(0, 5, 1), (78, 9, 50), (3, 16, 20)])
def test_afkmc2_vs_oracle(oracle, dtype, seed, k, m):
    """AFK-MC2 seeding on the GPU draws the oracle's chain: the runs that follow agree."""
    X, _ = _blobs(seed + 1, 500, 4, k, spread=4.0)
    X = X.astype(dtype).astype(np.float64)
    km = _fit(X, k, dtype, "row-major", initialization_method="afk-mc2", n_init=2,
              afk_mcmc_samples=m, seed=seed, max_iter=100, tol=1e-6)
    r = oracle.kmeans(X, k, "afk-mc2", n_init=2, afk_samples=m, seed=seed, max_iter=100,
                      tol=1e-6)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)


def test_afkmc2_duplicate_rows(oracle):
    """Chains that start at an already chosen sample (d(x) = 0) accept without a uniform draw;
    the engine must advance exactly as the oracle's does for the later centres to agree."""
    rng = np.random.default_rng(4)
    P = rng.integers(-10, 10, size=(12, 3)).astype(np.float64)
    X = P[rng.integers(0, 12, size=300)]
    km = _fit(X, 8, np.float64, "row-major", initialization_method="afk-mc2", n_init=3,
              afk_mcmc_samples=7, seed=11, max_iter=50, tol=1e-6)
    r = oracle.kmeans(X, 8, "afk-mc2", n_init=3, afk_samples=7, seed=11, max_iter=50, tol=1e-6)
    np.testing.assert_array_equal(km.labels, r["labels"])
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("algorithm", ["lloyd", "elkan"])
def test_mixed_precision_vs_oracle(oracle, dtype, algorithm):
    """Mixed-precision refinement on larger data: the low-precision phase rounds the data and
    every centre update to float (double data) or half (float data)."""
    X, rng = _blobs(21, 3000, 5, 12, spread=2.0)
    X = X.astype(dtype).astype(np.float64)
    C0 = X[rng.choice(len(X), 12, replace=False)]
    lp = "float" if dtype == np.float64 else "half"
    km = _fit(X, 12, dtype, "row-major", C=C0, max_iter=100, tol=1e-7, mixed_precision=True,
              low_precision_tol=1e-3, low_precision_max_iter=50, algorithm=algorithm)
    r = oracle.kmeans(X, 12, "supplied", C0, max_iter=100, tol=1e-7, mixed_precision=lp,
                      lp_tol=1e-3, lp_max_iter=50, algorithm=algorithm)
    assert r["lp_n_iter"] > 1
    assert (km.lp_n_iter, km.n_iter) == (r["lp_n_iter"], r["n_iter"])
    np.testing.assert_array_equal(km.labels, r["labels"])
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kmeans_many_features(oracle, dtype):
    """Centres wider than the assignment's LDS stage (d = 5000 doubles) are read from global
    memory; k-means++ candidates wider than the stage likewise."""
    X, rng = _blobs(31, 120, 5000, 3, spread=0.5)
    X = X.astype(dtype).astype(np.float64)
    km = _fit(X, 3, dtype, "row-major", n_init=1, seed=5, max_iter=30, tol=1e-8)
    r = oracle.kmeans(X, 3, "k-means++", n_init=1, seed=5, max_iter=30, tol=1e-8)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)
    np.testing.assert_array_equal(km.predict(X[:50].astype(dtype)), r["labels"][:50])


def test_kmeanspp_wide_candidates(oracle):
    """k = 256 at d = 1500: seven k-means++ trials of 1500 doubles exceed 64 KB of LDS."""
    X, rng = _blobs(41, 600, 1500, 16, spread=0.3)
    km = _fit(X, 256, np.float64, "row-major", n_init=1, seed=9, max_iter=3, tol=0.0)
    r = oracle.kmeans(X, 256, "k-means++", n_init=1, seed=9, max_iter=3, tol=0.0)
    np.testing.assert_array_equal(km.labels, r["labels"])
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("distance", ["euclidean", "cosine"])
def test_kmeans_features_beyond_one_sum_pass(oracle, distance):
    """d = 20000: the centroid sums take two feature passes per cluster (at most 16,384
    (cluster, feature) pairs per pass), the spherical normalisation follows the last."""
    X, rng = _blobs(51, 60, 20000, 2, spread=0.5)
    if distance == "cosine":
        km = aoclda.clustering.kmeans(n_clusters=2, n_init=1, seed=3, max_iter=10, tol=1e-8,
                                      distance="cosine", normalize_data=True).fit(X)
        r = oracle.kmeans_sph(X, 2, "k-means++", n_init=1, seed=3, max_iter=10, tol=1e-8,
                              normalize=True)
    else:
        km = _fit(X, 2, np.float64, "row-major", n_init=1, seed=3, max_iter=10, tol=1e-8)
        r = oracle.kmeans(X, 2, "k-means++", n_init=1, seed=3, max_iter=10, tol=1e-8)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-12, atol=1e-12)
