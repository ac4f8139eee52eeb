"""GPU parity of spherical (cosine) k-means, the "distance" option "cosine" of da_kmeans:
the reference's own SphericalKMeans checks (tests/unit_tests/kmeans/kmeans_public.cpp:721-1220)
through the C API, and every initialisation, both data normalisations, the empty-cluster
modes and the cross-term precisions against the oracle on the same seeded inputs. Labels
must be identical; centres and inertia within 1e-12 relative (fp64) or 1e-5 (fp32)."""
import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu

SPH_A = np.array([
    1.2, -0.5, 0.3, 2.1, -1.8, 0.7, 0.4, -0.2, 1.5, -0.9, 1.7, 0.6,
    -1.1, 0.8, -0.3, 2.0, -1.4, 0.1, 0.9, -0.7, 0.5, 1.3, -0.8, 0.2,
    1.6, -1.0, 0.4, -0.6, 1.1, -1.3, 0.7, -0.4, 1.8, -0.1, 0.3, -1.5,
    0.6, 1.0, -0.2, 0.8, -0.3, 0.9, 1.4, -0.7, 0.1, 0.5, -1.2, 0.3,
    1.6, -0.4, 0.8, -0.1, 0.6, -1.0, 1.3, 0.2, -0.5, 0.7, -0.8, 1.1]).reshape(3, 20).T
SPH_C = np.array([10.0, -5.0, 3.0, 7.0, -2.0, 8.0, -4.0, 6.0, 1.0]).reshape(3, 3).T


def _unit(X):
    n = np.linalg.norm(X, axis=1, keepdims=True)
    return np.where(n > 0, X / np.where(n > 0, n, 1), X)


def _fit(X, k, order="C", **kw):
    X = np.asarray(X)
    X = np.asfortranarray(X) if order == "F" else np.ascontiguousarray(X)
    kw.setdefault("n_init", 1)
    return aoclda.clustering.kmeans(n_clusters=k, distance="cosine", **kw).fit(X)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("algorithm", ["lloyd", "elkan"])
@pytest.mark.parametrize("init", ["random", "k-means++", "afk-mc2"])
def test_reference_sphericalkmeans(dtype, order, algorithm, init):
    tol = 1e-6 if dtype == np.float32 else 1e-12
    A = SPH_A.astype(dtype)
    a = _fit(A, 3, order, algorithm=algorithm, initialization_method=init, seed=42,
             normalize_data=True)
    b = _fit(2 * A, 3, order, algorithm=algorithm, initialization_method=init, seed=42,
             normalize_data=True)
    Ca, Cb = a.cluster_centres.astype(np.float64), b.cluster_centres.astype(np.float64)
    np.testing.assert_allclose((Ca ** 2).sum(1), 1.0, atol=tol)
    np.testing.assert_allclose((Cb ** 2).sum(1), 1.0, atol=tol)
    np.testing.assert_allclose(Ca, Cb, atol=tol)
    # rows collinear with the centres: predicted as their own centre, cosine distance 0
    Xt = (2.0 * b.cluster_centres).astype(dtype)
    np.testing.assert_array_equal(b.predict(Xt), [0, 1, 2])
    np.testing.assert_allclose(np.diag(b.transform(Xt)), 0.0, atol=tol)
    # pre-normalised data: normalize yes and no agree
    An = _unit(SPH_A).astype(dtype)
    on = _fit(An, 3, order, algorithm=algorithm, initialization_method=init, seed=42,
              normalize_data=True)
    off = _fit(An, 3, order, algorithm=algorithm, initialization_method=init, seed=42,
               normalize_data=False)
    np.testing.assert_array_equal(on.labels, off.labels)
    np.testing.assert_allclose(on.cluster_centres, off.cluster_centres, atol=tol)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["F", "C"])
def test_reference_supplied_and_split(oracle, dtype, order):
    """The reference's checks run with the C API's default "normalize data" = yes."""
    tol = 1e-6 if dtype == np.float32 else 1e-12
    km = _fit(SPH_A.astype(dtype), 3, order, C=SPH_C.astype(dtype), normalize_data=True)
    np.testing.assert_allclose((km.cluster_centres.astype(np.float64) ** 2).sum(1), 1.0, atol=tol)
    A = np.array([2.13, 2.11, 2.12, 2.13, 2.14, 2.11, 2.14, 2.13, 2.12, 2.13]).reshape(2, 5).T
    C = np.array([[1.0, 1.0], [-1.0, 1.0]])
    for alg in ("lloyd", "elkan"):
        km = _fit(A.astype(dtype), 2, order, C=C.astype(dtype), algorithm=alg,
                  empty_clusters="split", normalize_data=True)
        assert (np.bincount(km.labels, minlength=2) > 0).all()
        np.testing.assert_allclose((km.cluster_centres.astype(np.float64) ** 2).sum(1), 1.0,
                                   atol=tol)
    # both normalisations against the oracle (rows of norm 3 taken as unit rows leave the
    # relocated cluster empty again, as in the reference)
    for norm in (True, False):
        km = _fit(A.astype(dtype), 2, order, C=C.astype(dtype), empty_clusters="split",
                  normalize_data=norm)
        r = oracle.kmeans_sph(A.astype(dtype).astype(np.float64), 2, "supplied", C,
                              empty_clusters="split", normalize=norm)
        np.testing.assert_array_equal(km.labels, r["labels"])
        assert km.n_iter == r["n_iter"]
        np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=tol, atol=tol)
    with pytest.raises(aoclda.DAError) as e:
        _fit(A, 2, order, C=C, empty_clusters="error", normalize_data=True)
    assert e.value.status == L.CODE["empty_clusters"]


def test_incompatible_and_unimplemented_options():
    km = _fit(SPH_A, 3, algorithm="hartigan-wong")
    assert km.status == L.CODE["incompatible_options"]
    with pytest.raises(aoclda.DAError) as e:
        _fit(SPH_A.astype(np.float32), 3, mixed_precision=True)
    assert e.value.status == L.CODE["not_implemented"]


def _directions(seed, n, d, k, scaled=True):
    rng = np.random.default_rng(seed)
    dirs = _unit(rng.standard_normal((k, d)))
    comp = rng.integers(0, k, n)
    X = dirs[comp] + 0.35 * rng.standard_normal((n, d))
    if scaled:
        X *= rng.uniform(0.2, 5.0, (n, 1))
    X[7] = 0.0  # a zero row: no direction, label 0, weight 0 when normalised
    return X


CASES = [(init, norm, empty) for init in ("random", "k-means++", "supplied",
                                          "random partitions", "afk-mc2")
         for norm in (True, False) for empty in ("ignore", "split")]


@pytest.mark.parametrize("init,normalize,empty", CASES)
@pytest.mark.parametrize("dtype,precision", [(np.float64, "working"), (np.float32, "working"),
                                             (np.float32, "bf16")])
def test_against_oracle(oracle, init, normalize, empty, dtype, precision):
    n, d, k = 3000, 12, 9
    X = _directions(11, n, d, k, scaled=normalize)
    if not normalize:
        X = _unit(X)
    X = X.astype(dtype)
    C0 = _unit(X[np.random.default_rng(1).choice(n, k, replace=False)] * 3.0)
    kw = dict(initialization_method=init, seed=17, max_iter=60, tol=1e-9,
              normalize_data=normalize, empty_clusters=empty, precision=precision)
    if init == "supplied":
        kw["C"] = C0.astype(dtype)
    km = _fit(X, k, "C", **kw)
    r = oracle.kmeans_sph(X.astype(np.float64), k, init, C0 if init == "supplied" else None,
                          max_iter=60, tol=1e-9, seed=17, empty_clusters=empty,
                          normalize=normalize)
    assert r["status"] in (0, 21)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=rtol, atol=rtol)
    assert abs(km.inertia - r["inertia"]) <= rtol * max(1.0, abs(r["inertia"]))
    # predict and transform on new rows
    Y = _directions(12, 500, d, k).astype(dtype)
    C = km.cluster_centres.astype(np.float64)
    G = Y.astype(np.float64) @ C.T
    np.testing.assert_array_equal(km.predict(Y), np.argmax(G, axis=1))
    inv = 1.0 / np.maximum(np.linalg.norm(Y.astype(np.float64), axis=1), 1e-300)
    ref = 1.0 - np.clip(G * (inv[:, None] if normalize else 1.0), -1, 1)
    ref[(np.linalg.norm(Y, axis=1) == 0)] = 1.0
    np.testing.assert_allclose(km.transform(Y), ref, rtol=0, atol=10 * rtol)


@pytest.mark.parametrize("precision", ["working", "bf16"])
def test_60000x32_k64_fp32(oracle, precision):
    """A larger problem on the MFMA assignment with its exact re-check: the same labels,
    iterations and centres as the oracle."""
    X = _directions(21, 60_000, 32, 64).astype(np.float32)
    km = _fit(X, 64, "C", initialization_method="k-means++", seed=5, max_iter=20, tol=0.0,
              normalize_data=True, precision=precision)
    r = oracle.kmeans_sph(X.astype(np.float64), 64, "k-means++", max_iter=20, tol=0.0, seed=5,
                          normalize=True)
    np.testing.assert_array_equal(km.labels, r["labels"])
    assert km.n_iter == r["n_iter"]
    np.testing.assert_allclose(km.cluster_centres, r["centres"], rtol=1e-5, atol=1e-6)
