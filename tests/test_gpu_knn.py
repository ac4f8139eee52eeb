"""GPU parity of brute-force kNN (da_nn_set_data / da_nn_kneighbors) against the reference's
fixtures and the oracle. Distances: fp64 1e-12 relative, fp32 1e-5 relative; indices must
be identical except where two candidates are at the same distance (ties)."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu


def assert_knn(ind, dist, exp_ind, exp_dist, Xt, Xq, rtol, metric="euclidean"):
    scale = np.maximum(np.abs(exp_dist), 1.0)
    assert np.all(np.abs(dist - exp_dist) <= rtol * scale), np.abs(dist - exp_dist).max()
    for q, j in np.argwhere(ind != exp_ind):
        a = np.linalg.norm(Xt[ind[q, j]] - Xq[q]) if metric == "euclidean" else dist[q, j]
        b = np.linalg.norm(Xt[exp_ind[q, j]] - Xq[q]) if metric == "euclidean" else exp_dist[q, j]
        assert abs(a - b) <= rtol * max(1.0, abs(b)), (q, j, ind[q, j], exp_ind[q, j])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_example(known, dtype):
    c = known["knn_example"]
    Xt = np.array(c["X_train_colmajor"], dtype=dtype).reshape((3, 6)).T.copy(order="F")
    Xq = np.array(c["X_test_colmajor"], dtype=dtype).reshape((3, 3)).T.copy(order="F")
    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=3).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=4)
    exp_i = np.array(c["k4_ind_colmajor"]).reshape((4, 3)).T
    exp_d = np.array(c["k4_dist_colmajor"]).reshape((4, 3)).T
    assert_knn(ind, dist, exp_i, exp_d, Xt.astype(np.float64), Xq.astype(np.float64),
               1e-12 if dtype == np.float64 else 1e-5)
    # the default number of neighbours comes from the option
    d3, i3 = nn.kneighbors(Xq)
    assert i3.shape == (3, 3)


@pytest.mark.parametrize("name", ["2092x5", "2125x20"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
def test_csv_fixtures(nn_csv, name, dtype, order):
    z = nn_csv[name]
    Xt = np.asarray(z["X_train"].astype(dtype), order=order)
    Xq = np.asarray(z["X_test"].astype(dtype), order=order)
    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=5).fit(Xt)
    dist, ind = nn.kneighbors(Xq)
    assert_knn(ind, dist, z["k_ind"], z["k_dist"], z["X_train"], z["X_test"],
               1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "manhattan", "cosine",
                                    "minkowski", "inner_product", "euclidean_gemm",
                                    "sqeuclidean_gemm"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vs_oracle_metrics(oracle, metric, dtype):
    rng = np.random.default_rng(3)
    Xt = rng.standard_normal((1500, 24)).astype(dtype)
    Xq = rng.standard_normal((300, 24)).astype(dtype)
    nn = aoclda.neighbors.nearest_neighbors(metric=metric, p=3.0).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=7)
    om = {"sqeuclidean_gemm": "sqeuclidean", "euclidean_gemm": "euclidean"}.get(metric, metric)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), 7, om, p=3.0,
                        eps=np.finfo(dtype).eps)
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(dist, ed, rtol=rtol, atol=rtol)
    np.testing.assert_array_equal(ind, ei)


@pytest.mark.parametrize("k", [1, 2, 64, 300])
def test_k_edges(oracle, k):
    rng = np.random.default_rng(4)
    Xt = rng.standard_normal((300, 5))
    Xq = rng.standard_normal((17, 5))
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=k)
    ei, ed = oracle.knn(Xt, Xq, k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12)


def test_k_errors():
    X = np.random.default_rng(0).standard_normal((10, 3))
    nn = aoclda.neighbors.nearest_neighbors().fit(X)
    with pytest.raises(aoclda.DAError) as e:
        nn.kneighbors(X, n_neighbors=11)
    assert e.value.status == L.CODE["invalid_input"]
    with pytest.raises(aoclda.DAError) as e:
        nn.kneighbors(np.ones((2, 4)))
    assert e.value.status == L.CODE["invalid_array_dimension"]


def test_duplicates_and_ties(oracle):
    """Many identical training rows: ties are resolved by index, deterministically."""
    base = np.random.default_rng(5).standard_normal((50, 4))
    Xt = np.concatenate([base, base, base])
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    d1, i1 = nn.kneighbors(base, n_neighbors=6)
    d2, i2 = nn.kneighbors(base, n_neighbors=6)
    np.testing.assert_array_equal(i1, i2)
    ei, ed = oracle.knn(Xt, base, 6)
    np.testing.assert_array_equal(i1, ei)
    assert np.all(d1[:, :3] == 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_device_api(oracle, dtype):
    """da_nn_kneighbors_device_?: queries and results in device memory."""
    rng = np.random.default_rng(6)
    Xt = rng.standard_normal((2000, 16)).astype(dtype)
    Xq = rng.standard_normal((100, 16)).astype(dtype)
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    h = nn.handle
    k = 10
    dq, di, dd = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(L.lib.da_amd_malloc(Xq.nbytes, C.byref(dq)))
    L.check(L.lib.da_amd_malloc(100 * k * 4, C.byref(di)))
    L.check(L.lib.da_amd_malloc(100 * k * Xq.itemsize, C.byref(dd)))
    try:
        L.check(L.lib.da_amd_memcpy_htod(dq, L.ptr(Xq), Xq.nbytes))
        h.call("da_nn_kneighbors_device", 100, 16, dq, 16, di, dd, k, 1)
        L.check(L.lib.da_amd_synchronize(h.h))
        ind = np.empty((100, k), dtype=np.int32)
        dist = np.empty((100, k), dtype=dtype)
        L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(ind), di, ind.nbytes))
        L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(dist), dd, dist.nbytes))
    finally:
        for p in (dq, di, dd):
            L.lib.da_amd_free(p)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12 if dtype == np.float64 else 1e-5)


def test_config_plain_small_knn(oracle):
    """X 2000x64 as training and Y 1500x64 as queries, fp64 U[0,1) seed 0, k = 5."""
    rng = np.random.default_rng(0)
    X = rng.random((2000, 64))
    Y = rng.random((1500, 64))
    dist, ind = aoclda.neighbors.nearest_neighbors(n_neighbors=5).fit(X).kneighbors(Y)
    ei, ed = oracle.knn(X, Y, 5)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12)


def gaussian_mixture(rng, n, means, sigma=1.0):
    comp = rng.integers(0, len(means), n)
    return means[comp] + sigma * rng.standard_normal((n, means.shape[1]))


@pytest.mark.parametrize("precision", ["working", "fp32", "fp16", "bf16", "fp16x3", "bf16x3",
                                       "fp16c"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [20, 96, 128])
def test_precision_forms_exact(oracle, precision, dtype, d):
    """Every operand precision of the distance tiles must return the exact neighbours: the
    candidates are re-ranked in double and certified, uncertified queries are recomputed."""
    rng = np.random.default_rng(d)
    means = rng.uniform(-10, 10, (16, d))
    Xt = gaussian_mixture(rng, 6000, means).astype(dtype)
    Xq = gaussian_mixture(rng, 500, means).astype(dtype)
    nn = aoclda.neighbors.nearest_neighbors(precision=precision).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=10)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), 10)
    np.testing.assert_array_equal(ind, ei)
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(dist, ed, rtol=rtol)
    nfb, ncand = L.da_int(), L.da_int()
    L.check(L.lib.da_nn_last_query_stats(nn.handle.h, C.byref(nfb), C.byref(ncand)))
    assert ncand.value > 0  # the fused tile path ran


def test_fused_split_edges(oracle):
    """Training sizes that leave partial steps and empty splits; queries not a multiple of
    the 128-query block."""
    rng = np.random.default_rng(12)
    for n in (16, 17, 33, 1000, 1025, 4097):
        Xt = rng.standard_normal((n, 40)).astype(np.float32)
        Xq = rng.standard_normal((130, 40)).astype(np.float32)
        nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
        dist, ind = nn.kneighbors(Xq, n_neighbors=min(16, n))
        ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), min(16, n))
        np.testing.assert_array_equal(ind, ei)
        np.testing.assert_allclose(dist, ed, rtol=1e-5)


def test_sqeuclidean_fused(oracle):
    rng = np.random.default_rng(13)
    Xt = rng.standard_normal((3000, 64)).astype(np.float32)
    Xq = rng.standard_normal((200, 64)).astype(np.float32)
    dist, ind = aoclda.neighbors.nearest_neighbors(metric="sqeuclidean").fit(Xt).kneighbors(Xq, 8)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), 8, "sqeuclidean")
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-5)


@pytest.mark.parametrize("k", [17, 32, 33, 64, 100, 128])
@pytest.mark.parametrize("precision", ["working", "fp16x3", "bf16"])
def test_fused_large_k(oracle, k, precision):
    """Candidate lists of 32, 64 and 128 entries per split (k up to 128) stay exact."""
    rng = np.random.default_rng(100 + k)
    means = rng.uniform(-10, 10, (8, 48))
    Xt = gaussian_mixture(rng, 5000, means).astype(np.float32)
    Xq = gaussian_mixture(rng, 150, means).astype(np.float32)
    nn = aoclda.neighbors.nearest_neighbors(precision=precision).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=k)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-5)
    nfb, ncand = L.da_int(), L.da_int()
    L.check(L.lib.da_nn_last_query_stats(nn.handle.h, C.byref(nfb), C.byref(ncand)))
    assert ncand.value >= k  # the fused tile path ran


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("precision", ["working", "fp64", "fp16", "fp16x3", "bf16x3"])
@pytest.mark.parametrize("d", [20, 64, 128])
@pytest.mark.parametrize("k", [5, 12, 30])
def test_tile_variants(oracle, variant, precision, d, k):
    """Every sweep of the fused tile (two training tiles per step with the loads under the
    bookkeeping, load-then-compute, one tile per step, the rolling prefetch) gives the exact
    result, for short
    (16) and longer (32) candidate lists, including training sizes that end inside a step
    and splits of unequal length."""
    rng = np.random.default_rng(200 + d)
    means = rng.uniform(-5, 5, (8, d))
    for n in (1000, 4133):
        Xt = gaussian_mixture(rng, n, means)
        Xq = gaussian_mixture(rng, 140, means)
        dtype = np.float64 if precision == "fp64" else np.float32
        prec = "working" if precision == "fp64" else precision
        nn = aoclda.neighbors.nearest_neighbors(precision=prec, tile_variant=variant).fit(
            Xt.astype(dtype))
        dist, ind = nn.kneighbors(Xq.astype(dtype), n_neighbors=k)
        ei, ed = oracle.knn(Xt, Xq, k)
        np.testing.assert_array_equal(ind, ei)
        np.testing.assert_allclose(dist, ed, rtol=1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("variant", [0, 2, 3])
@pytest.mark.parametrize("k", [10, 100])
def test_long_splits_16bit_offsets(oracle, variant, k):
    """A training set of more than 65,536 rows: sweeps that buffer 16-bit row offsets (16
    and 128 candidates per split here) take at least two splits and still return the exact
    neighbours with the right indices."""
    rng = np.random.default_rng(77)
    d = 24
    means = rng.uniform(-5, 5, (16, d))
    Xt = gaussian_mixture(rng, 70001, means)
    Xq = gaussian_mixture(rng, 48, means)
    nn = aoclda.neighbors.nearest_neighbors(precision="bf16x3", tile_variant=variant).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=k)
    ei, ed = oracle.knn(Xt, Xq, k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-5)

@pytest.mark.parametrize("precision", ["working", "fp16", "bf16x3", "fp16x3", "fp16c"])
@pytest.mark.parametrize("k", [1, 10, 16, 30])
def test_seeded_thresholds(oracle, precision, k):
    """Seeding the thresholds (k <= 16, training sets of 131,072 rows or more) from the
    group minima of an 8,192-row sample returns exactly the neighbours of the unseeded sweep
    and of the oracle, also for queries far from every training row (whose seeds are loose),
    for k = 16 (the last of the 16 groups) and for plain fp16, whose wider error bound sends
    queries to the exact fallback. A seed is an upper bound on the k-th key, so for the
    forms that certify every query unseeded it sends none to the fallback either; k = 30
    runs unseeded."""
    rng = np.random.default_rng(91)
    d = 40
    means = rng.uniform(-5, 5, (32, d))
    Xt = gaussian_mixture(rng, 140000, means)
    Xq = np.concatenate([gaussian_mixture(rng, 290, means),
                         rng.uniform(-40, 40, (10, d)).astype(np.float32)])
    ei, ed = oracle.knn(Xt, Xq, k)
    fallbacks = {}
    for seed in (True, False):
        nn = aoclda.neighbors.nearest_neighbors(precision=precision,
                                                seed_thresholds=seed).fit(Xt)
        dist, ind = nn.kneighbors(Xq, n_neighbors=k)
        np.testing.assert_array_equal(ind, ei, err_msg=f"seed={seed}")
        np.testing.assert_allclose(dist, ed, rtol=1e-5)
        nfb, ncand = L.da_int(), L.da_int()
        L.check(L.lib.da_nn_last_query_stats(nn.handle.h, C.byref(nfb), C.byref(ncand)))
        fallbacks[seed] = nfb.value
    if precision in ("working", "bf16x3", "fp16x3") and fallbacks[False] == 0:
        assert fallbacks[True] == 0


@pytest.mark.parametrize("precision", ["working", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("k,kc", [(100, 64), (40, 16), (60, 32), (128, 64)])
def test_lists_shorter_than_k(oracle, precision, k, kc):
    """Candidate lists shorter than k per split ("candidates per split" < k) return the exact
    neighbours: a split's rejected rows lie above its threshold, so only queries whose k
    nearest rows were all kept are certified; the others take the exact fallback."""
    rng = np.random.default_rng(71)
    means = rng.uniform(-5, 5, (10, 24))
    Xt = gaussian_mixture(rng, 60000, means).astype(np.float32)
    Xq = gaussian_mixture(rng, 300, means).astype(np.float32)
    nn = aoclda.neighbors.nearest_neighbors(precision=precision, candidates_per_split=kc).fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=k)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-5)


@pytest.mark.parametrize("n,d", [(16, 8), (383, 40), (1000, 20), (4133, 64), (52000, 128),
                                 (70001, 24)])
@pytest.mark.parametrize("k", [1, 5, 16, 30, 100])
def test_centred_fp16_form(oracle, n, d, k):
    """fp16 training rows stored about the centres of k-means groups (one group per 384
    rows, padded to whole 32-row steps, in a random order): the exact neighbours with the
    original indices, on data far from the origin (so that the centring matters), with
    duplicated rows (ties go to the lower index) and sizes from a single group to
    several 16-bit-offset splits; most queries are certified by the fused tile."""
    if k > n:
        pytest.skip("k > n")
    rng = np.random.default_rng(n + d)
    means = rng.uniform(-30, 30, (24, d))
    Xt = gaussian_mixture(rng, n, means).astype(np.float32)
    if n > 100:
        Xt[n // 2: n // 2 + 20] = Xt[:20]  # exact duplicates
    Xq = gaussian_mixture(rng, 300, means).astype(np.float32)
    Xq[:5] = Xt[:5]
    nn = aoclda.neighbors.nearest_neighbors(precision="fp16c").fit(Xt)
    dist, ind = nn.kneighbors(Xq, n_neighbors=k)
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq.astype(np.float64), k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-5, atol=1e-6)
    nfb, ncand = L.da_int(), L.da_int()
    L.check(L.lib.da_nn_last_query_stats(nn.handle.h, C.byref(nfb), C.byref(ncand)))
    assert ncand.value > 0
    # a handle with the split form returns the same indices
    nn2 = aoclda.neighbors.nearest_neighbors(precision="bf16x3").fit(Xt)
    d2, i2 = nn2.kneighbors(Xq, n_neighbors=k)
    np.testing.assert_array_equal(i2, ind)
