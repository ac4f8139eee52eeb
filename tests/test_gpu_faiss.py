"""GPU tests of the faiss front end (aoclda.faiss), after the reference's
python_interface/tests/faissAPI/{flat_l2,ivfflat,kmeans}_faiss_test.py, with the oracle as
the checker: IndexFlatL2 returns the exact squared-euclidean neighbours (indices identical),
IndexIVFFlat the oracle's probed-list search over the centroids it trained, and Kmeans the
same centres as the k-means front end with assign = nearest centre. faiss is not
installed, so a stand-in module plays its part for the patch."""
import pickle
import sys
import types

import numpy as np
import pytest

import aoclda
from aoclda.faiss import IndexFlatL2, IndexIVFFlat, Kmeans, faiss_patch, undo_faiss_patch
from aoclda.faiss import patch_faiss as pf

pytestmark = pytest.mark.gpu

D, N, NQ, K = 16, 500, 20, 5
_rng = np.random.default_rng(0)
X_CORPUS = _rng.random((N, D)).astype(np.float32)
X_QUERY = _rng.random((NQ, D)).astype(np.float32)
DTYPES_ORDERS = [(np.float32, "C"), (np.float32, "F"), (np.float64, "C"), (np.float64, "F")]


@pytest.fixture
def faiss(monkeypatch):
    mod = types.ModuleType("faiss")
    mod.IndexFlatL2, mod.IndexIVFFlat, mod.Kmeans = object, object, object
    mod.index_factory = lambda d, desc, metric=1: None
    mod.METRIC_L2, mod.METRIC_INNER_PRODUCT = 1, 0
    monkeypatch.setitem(sys.modules, "faiss", mod)
    monkeypatch.setattr(pf, "_saved", {})
    faiss_patch(print_patched=False)
    yield mod
    undo_faiss_patch(print_patched=False)


def _rtol(dtype):
    return 1e-12 if dtype == np.float64 else 1e-5


@pytest.mark.parametrize("dtype,order", DTYPES_ORDERS)
def test_flat_two_adds_match_oracle(faiss, oracle, dtype, order):
    corpus = np.array(X_CORPUS, dtype=dtype, order=order)
    query = np.array(X_QUERY, dtype=dtype, order=order)
    idx = faiss.IndexFlatL2(D)
    assert isinstance(idx, IndexFlatL2)
    idx.add(corpus[:250])
    assert idx.ntotal == 250
    idx.add(corpus[250:])
    assert idx.ntotal == N
    dist, ind = idx.search(query, K)
    assert dist.shape == (NQ, K) and ind.shape == (NQ, K)
    assert dist.dtype == dtype and ind.dtype == np.int64
    oi, od = oracle.knn(corpus, query, K, metric="sqeuclidean")
    np.testing.assert_array_equal(ind, oi)
    np.testing.assert_allclose(dist, od, rtol=_rtol(dtype), atol=1e-12)
    # The same as the k-NN front end directly.
    nn = aoclda.neighbors.nearest_neighbors(metric="sqeuclidean").fit(corpus)
    d2, i2 = nn.kneighbors(query, n_neighbors=K)
    np.testing.assert_array_equal(ind, i2)
    np.testing.assert_array_equal(dist, d2)


def test_flat_reset_assign_reconstruct(faiss):
    idx = faiss.IndexFlatL2(D)
    idx.add(X_CORPUS)
    dist, ind = idx.search(X_QUERY, K)
    np.testing.assert_array_equal(idx.assign(X_QUERY, K), ind)
    np.testing.assert_array_equal(idx.reconstruct(7), X_CORPUS[7])
    with pytest.raises(IndexError):
        idx.reconstruct(N)
    np.testing.assert_array_equal(idx.reconstruct_n(3, 4), X_CORPUS[3:7])
    np.testing.assert_array_equal(idx.reconstruct_batch([9, 1, 4]), X_CORPUS[[9, 1, 4]])
    d2, i2, rec = idx.search_and_reconstruct(X_QUERY, K)
    np.testing.assert_array_equal(i2, ind)
    np.testing.assert_array_equal(d2, dist)
    assert rec.shape == (NQ, K, D) and rec.dtype == np.float32
    np.testing.assert_array_equal(rec, X_CORPUS[ind])
    idx.reset()
    assert idx.ntotal == 0
    with pytest.raises(RuntimeError):
        idx.search(X_QUERY, K)


def test_flat_pickle_round_trip(faiss, tmp_path):
    idx = faiss.IndexFlatL2(D)
    idx.add(X_CORPUS)
    dist, ind = idx.search(X_QUERY, K)
    path = tmp_path / "flat.pkl"
    path.write_bytes(pickle.dumps(idx))
    del idx
    undo_faiss_patch(print_patched=False)
    loaded = pickle.loads(path.read_bytes())
    assert loaded.aocl is True and loaded.ntotal == N
    d2, i2 = loaded.search(X_QUERY, K)
    np.testing.assert_array_equal(i2, ind)
    np.testing.assert_array_equal(d2, dist)
    np.testing.assert_array_equal(loaded.reconstruct_n(0, 3), X_CORPUS[:3])
    faiss_patch(print_patched=False)


@pytest.mark.parametrize("dtype,order", DTYPES_ORDERS)
def test_ivf_train_add_search_match_oracle(faiss, oracle, dtype, order):
    corpus = np.array(X_CORPUS, dtype=dtype, order=order)
    query = np.array(X_QUERY, dtype=dtype, order=order)
    idx = faiss.IndexIVFFlat(faiss.IndexFlatL2(D), D, 16, faiss.METRIC_L2)
    assert isinstance(idx, IndexIVFFlat) and not idx.is_trained
    idx.nprobe = 4
    idx.train(corpus)
    assert idx.is_trained and idx.ntotal == 0
    idx.add(corpus)
    assert idx.ntotal == N
    dist, ind = idx.search(query, K)
    assert dist.shape == (NQ, K) and dist.dtype == dtype and ind.dtype == np.int64
    assert (dist >= 0).all()
    C = idx._ann.cluster_centroids
    oi, od, sizes = oracle.ivf_search(C, corpus, query, K, 4, "sqeuclidean",
                                      single=dtype == np.float32)
    np.testing.assert_array_equal(ind, oi)
    np.testing.assert_allclose(dist, od, rtol=_rtol(dtype) * 10, atol=1e-12)
    np.testing.assert_array_equal(idx._ann.list_sizes, sizes)


@pytest.mark.parametrize("dtype,order", DTYPES_ORDERS)
def test_ivf_inner_product_matches_oracle(faiss, oracle, dtype, order):
    """METRIC_INNER_PRODUCT: spherical k-means lists, the largest inner products first,
    reported as inner products (faiss's convention)."""
    corpus = np.array(X_CORPUS - 0.5, dtype=dtype, order=order)
    query = np.array(X_QUERY - 0.5, dtype=dtype, order=order)
    idx = faiss.IndexIVFFlat(faiss.IndexFlatL2(D), D, 8, faiss.METRIC_INNER_PRODUCT)
    assert idx.metric_type == faiss.METRIC_INNER_PRODUCT
    idx.nprobe = 3
    idx.train(corpus)
    idx.add(corpus)
    dist, ind = idx.search(query, K)
    assert (np.diff(dist.astype(np.float64), axis=1) <= 0).all()
    C = idx._ann.cluster_centroids
    oi, od, sizes = oracle.ivf_search(C, corpus, query, K, 3, "inner product",
                                      single=dtype == np.float32)
    np.testing.assert_array_equal(ind, oi)
    np.testing.assert_allclose(dist, od, rtol=_rtol(dtype) * 10, atol=1e-6)
    np.testing.assert_array_equal(idx._ann.list_sizes, sizes)


def _recall(ind, exact):
    return np.mean([len(set(a) & set(b)) / K for a, b in zip(ind, exact)])


def test_ivf_recall_fixed_seed(faiss, oracle):
    rng = np.random.default_rng(123)
    X = rng.random((1300, D)).astype(np.float32)
    Q = rng.random((50, D)).astype(np.float32)
    exact, _ = oracle.knn(X, Q, K, metric="sqeuclidean")
    idx = faiss.IndexIVFFlat(faiss.IndexFlatL2(D), D, 32, faiss.METRIC_L2)
    idx.train(X)
    idx.add(X)
    idx.nprobe = 8
    assert _recall(idx.search(Q, K)[1], exact) >= 0.8


def test_ivf_nprobe_monotonic_recall(faiss, oracle):
    rng = np.random.default_rng(456)
    X = rng.random((1300, D)).astype(np.float32)
    Q = rng.random((50, D)).astype(np.float32)
    exact, _ = oracle.knn(X, Q, K, metric="sqeuclidean")
    idx = faiss.IndexIVFFlat(faiss.IndexFlatL2(D), D, 32, faiss.METRIC_L2)
    idx.train(X)
    idx.add(X)
    recalls, means = [], []
    for p in (1, 2, 4, 8, 16, 32):
        idx.nprobe = p
        assert idx.nprobe == p
        dist, ind = idx.search(Q, K)
        recalls.append(_recall(ind, exact))
        means.append(float(dist.mean()))
    # Every list probed: the exact neighbours.
    assert recalls[-1] == 1.0
    assert all(b >= a - 1e-9 for a, b in zip(recalls, recalls[1:]))
    assert all(b <= a + 1e-5 for a, b in zip(means, means[1:]))


def test_ivf_cp_lifecycle(faiss):
    idx = faiss.IndexIVFFlat(faiss.IndexFlatL2(D), D, 16)
    idx.cp.niter, idx.cp.seed, idx.cp.max_points_per_centroid = 5, 42, 20
    idx.nprobe = 4
    idx.train(X_CORPUS)
    assert idx.is_trained and idx.ntotal == 0 and idx.nprobe == 4
    # 20 points per list of 16 lists: a 320-point training sample; 5 iterations at most.
    assert idx._ann.rinfo[3] <= 5
    idx.add(X_CORPUS)
    assert idx.ntotal == N
    idx.reset()
    assert not idx.is_trained and idx.ntotal == 0
    assert idx.cp.niter == 5 and idx.cp.seed == 42 and idx.nprobe == 4
    with pytest.raises(TypeError):
        pickle.dumps(idx)


def test_kmeans_train_assign(faiss, oracle):
    rng = np.random.default_rng(2)
    centres = rng.uniform(-10, 10, size=(4, 8))
    X = (centres[rng.integers(0, 4, 400)] + rng.normal(size=(400, 8))).astype(np.float64)
    km = faiss.Kmeans(8, 4, niter=20, seed=3)
    assert isinstance(km, Kmeans)
    km.train(X)
    assert km.is_trained and km.centroids.shape == (4, 8)
    ref = aoclda.clustering.kmeans(n_clusters=4, max_iter=20, n_init=1, seed=3).fit(X)
    np.testing.assert_array_equal(km.centroids, ref.cluster_centres)
    dist, ind = km.assign(X)
    assert ind.dtype == np.int64 and dist.shape == (400,)
    oi, od = oracle.knn(km.centroids, X, 1, metric="euclidean")
    np.testing.assert_array_equal(ind, oi[:, 0])
    np.testing.assert_allclose(dist, od[:, 0], rtol=1e-10)
    np.testing.assert_array_equal(ind, ref.labels)


def test_kmeans_spherical(faiss):
    """spherical=True maps to the k-means option distance = "cosine": unit centroids, the
    same as the k-means front end."""
    X = np.random.default_rng(0).standard_normal((200, 8))
    km = faiss.Kmeans(8, 3, spherical=True, seed=1, niter=15)
    km.train(X)
    np.testing.assert_allclose(np.linalg.norm(km.centroids, axis=1), 1.0, atol=1e-12)
    ref = aoclda.clustering.kmeans(n_clusters=3, max_iter=15, n_init=1, seed=1,
                                   distance="cosine").fit(X)
    np.testing.assert_array_equal(km.centroids, ref.cluster_centres)
