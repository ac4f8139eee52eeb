"""Pins the oracle's spherical (cosine) k-means: the reference's own SphericalKMeans checks
(tests/unit_tests/kmeans/kmeans_public.cpp:721-1220: unit centres, invariance to the scale of
the data, normalised and pre-normalised data agreeing, supplied centres normalised, empty
clusters split into unit centres), the fixed point that defines spherical k-means (Dhillon and
Modha 2001: every label is the largest x.c and every centre is the normalised sum of its
members' directions), the Euclidean oracle on unit rows (where 1 - x.y = |x - y|^2 / 2) and
an independent restatement of the k-means++ and AFK-MC2 draws with cosine weights. CPU only."""
import itertools

import numpy as np
import pytest

# The 20 x 3 data set of the reference's SphericalKMeans test, column-major there.
SPH_A_COL = [
    1.2, -0.5, 0.3, 2.1, -1.8, 0.7, 0.4, -0.2, 1.5, -0.9, 1.7, 0.6,
    -1.1, 0.8, -0.3, 2.0, -1.4, 0.1, 0.9, -0.7, 0.5, 1.3, -0.8, 0.2,
    1.6, -1.0, 0.4, -0.6, 1.1, -1.3, 0.7, -0.4, 1.8, -0.1, 0.3, -1.5,
    0.6, 1.0, -0.2, 0.8, -0.3, 0.9, 1.4, -0.7, 0.1, 0.5, -1.2, 0.3,
    1.6, -0.4, 0.8, -0.1, 0.6, -1.0, 1.3, 0.2, -0.5, 0.7, -0.8, 1.1]
SPH_A = np.array(SPH_A_COL).reshape(3, 20).T
# Unnormalised supplied centres of the same test (column-major 3 x 3 there).
SPH_C = np.array([10.0, -5.0, 3.0, 7.0, -2.0, 8.0, -4.0, 6.0, 1.0]).reshape(3, 3).T
INITS = ["random", "k-means++", "afk-mc2", "random partitions"]


def _unit(X):
    n = np.linalg.norm(X, axis=1, keepdims=True)
    return np.where(n > 0, X / np.where(n > 0, n, 1), X)


@pytest.mark.parametrize("init", INITS)
def test_reference_sphericalkmeans_checks(oracle, init):
    """Unit centres, identical centres for A and 2A, and normalize yes on pre-normalised
    data equal to normalize no (the reference's checks, Lloyd)."""
    r1 = oracle.kmeans_sph(SPH_A, 3, init, seed=42)
    r2 = oracle.kmeans_sph(2.0 * SPH_A, 3, init, seed=42)
    assert r1["status"] == 0 and r2["status"] == 0
    np.testing.assert_allclose((r1["centres"] ** 2).sum(1), 1.0, atol=1e-12)
    np.testing.assert_allclose(r1["centres"], r2["centres"], atol=1e-12)
    np.testing.assert_array_equal(r1["labels"], r2["labels"])
    An = _unit(SPH_A)
    on = oracle.kmeans_sph(An, 3, init, seed=42, normalize=True)
    off = oracle.kmeans_sph(An, 3, init, seed=42, normalize=False)
    np.testing.assert_array_equal(on["labels"], off["labels"])
    np.testing.assert_allclose(on["centres"], off["centres"], atol=1e-12)


def test_reference_supplied_centres_normalised(oracle):
    r = oracle.kmeans_sph(SPH_A, 3, "supplied", SPH_C)
    assert r["status"] == 0
    np.testing.assert_allclose((r["centres"] ** 2).sum(1), 1.0, atol=1e-12)


def test_reference_empty_cluster_split(oracle):
    """All five rows point near (1, 1); the supplied centre (-1, 1) is orthogonal, so its
    cluster empties and the split mode refills it with a unit centre."""
    A = np.array([2.13, 2.11, 2.12, 2.13, 2.14, 2.11, 2.14, 2.13, 2.12, 2.13]).reshape(2, 5).T
    C = np.array([[1.0, 1.0], [-1.0, 1.0]])
    r = oracle.kmeans_sph(A, 2, "supplied", C, empty_clusters="split")
    assert r["status"] == 0
    assert (np.bincount(r["labels"], minlength=2) > 0).all()
    np.testing.assert_allclose((r["centres"] ** 2).sum(1), 1.0, atol=1e-12)
    # the error mode gives up instead
    assert oracle.kmeans_sph(A, 2, "supplied", C, empty_clusters="error")["status"] == 900


@pytest.mark.parametrize("normalize", [True, False])
def test_fixed_point_and_inertia(oracle, normalize):
    """A run that stops because no label changed is at the fixed point of spherical
    k-means; the reported inertia is sum(1 - x.c/|x|)."""
    rng = np.random.default_rng(5)
    dirs = _unit(rng.standard_normal((6, 5)))
    comp = rng.integers(0, 6, 400)
    X = dirs[comp] + 0.25 * rng.standard_normal((400, 5))
    X *= rng.uniform(0.5, 3.0, (400, 1))
    if not normalize:
        X = _unit(X)
    r = oracle.kmeans_sph(X, 6, "k-means++", seed=3, max_iter=200, tol=0.0,
                          normalize=normalize)
    assert r["status"] == 0 and r["n_iter"] < 200
    C, lab = r["centres"], r["labels"]
    G = X @ C.T
    assert (G[np.arange(len(X)), lab] >= G.max(1) - 1e-12).all()
    Xd = _unit(X)
    for c in range(6):
        s = Xd[lab == c].sum(0)
        np.testing.assert_allclose(C[c], s / np.linalg.norm(s), atol=1e-12)
    cos = (X * C[lab]).sum(1) / np.linalg.norm(X, axis=1)
    assert abs(r["inertia"] - (1.0 - cos).sum()) <= 1e-10 * len(X)


def test_unit_rows_follow_the_euclidean_update(oracle):
    """On unit rows with unit centres the first labels are the Euclidean ones and the new
    centres are the Euclidean means scaled to unit length."""
    rng = np.random.default_rng(9)
    X = _unit(rng.standard_normal((300, 4)) + np.array([2.0, 0, 0, 0]))
    C0 = _unit(X[rng.choice(300, 5, replace=False)] + 0.1)
    s = oracle.kmeans_sph(X, 5, "supplied", C0, max_iter=1, tol=0.0, normalize=False)
    e = oracle.kmeans(X, 5, "supplied", C0, max_iter=1, tol=0.0)
    np.testing.assert_allclose(s["centres"], _unit(e["centres"]), atol=1e-12)


def _exact_unit_rows(seed, n, scale):
    """Rows from the 24 unit vectors of R^4 with components in {0, +-1/2, +-1}: every dot
    product and cosine is exact in binary, so weights of either restatement agree bitwise."""
    vecs = [np.eye(4)[i] * s for i in range(4) for s in (1.0, -1.0)]
    vecs += [np.array(p) * 0.5 for p in itertools.product((1.0, -1.0), repeat=4)]
    V = np.array(vecs)
    rng = np.random.default_rng(seed)
    X = V[rng.integers(0, len(V), n)]
    if scale:
        X = X * 2.0 ** rng.integers(-2, 3, (n, 1))
    return X


def _cos_dist(X, i, j):
    x, y = X[i], X[j]
    inv = lambda v: 1.0 / float(np.sqrt(sum(t * t for t in v)))  # noqa: E731
    return max(0.0, 1.0 - float(sum(a * b for a, b in zip(x, y))) * inv(x) * inv(y))


@pytest.mark.parametrize("seed", [0, 7])
@pytest.mark.parametrize("k", [3, 6])
@pytest.mark.parametrize("scale", [False, True])
def test_kmeanspp_cosine_draws_equal_independent_restatement(oracle, seed, k, scale):
    import _stdrng
    X = _exact_unit_rows(seed + k, 40, scale)
    idx = _stdrng.kmeanspp(X, k, seed, dist=_cos_dist)
    r = oracle.kmeans_sph(X, k, "k-means++", seed=seed, max_iter=0, normalize=True)
    np.testing.assert_array_equal(r["centres"], _unit(_stdrng.rows(X, idx)))


@pytest.mark.parametrize("seed", [0, 78])
@pytest.mark.parametrize("k,m", [(3, 1), (5, 20)])
@pytest.mark.parametrize("scale", [False, True])
def test_afkmc2_cosine_draws_equal_independent_restatement(oracle, seed, k, m, scale):
    import _stdrng
    X = _exact_unit_rows(seed + 3 * k + m, 40, scale)
    idx = _stdrng.afkmc2(X, k, seed, m, dist=_cos_dist)
    r = oracle.kmeans_sph(X, k, "afk-mc2", seed=seed, max_iter=0, normalize=True,
                          afk_samples=m)
    np.testing.assert_array_equal(r["centres"], _unit(_stdrng.rows(X, idx)))
