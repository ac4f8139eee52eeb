"""GPU parity of radius neighbours (da_nn_radius_neighbors_? and the radius results of
da_handle_get_result_?) against the reference's known answers and the oracle. Indices are
bit-exact (each query's list in ascending index, or by (distance, index) when sorted);
distances within 1e-12 (fp64) and 1e-6 (fp32) relative."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu


def split(counts, flat):
    b = np.concatenate(([0], np.cumsum(counts)))
    return [flat[b[q]:b[q + 1]] for q in range(len(counts))]


def check_against_oracle(dists, inds, counts, eind, edist, rtol):
    assert [len(v) for v in inds] == list(counts)
    for q, (gi, ei) in enumerate(zip(inds, split(counts, eind))):
        np.testing.assert_array_equal(gi, ei, err_msg=f"query {q}")
    if dists is not None:
        for q, (gd, ed) in enumerate(zip(dists, split(counts, edist))):
            np.testing.assert_allclose(gd, ed, rtol=rtol, atol=rtol, err_msg=f"query {q}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_known_answers(known, dtype, metric):
    """The reference's answers on its 6x3 example at radius 5, sorted and unsorted."""
    c = known["knn_example"]
    Xt = np.array(c["X_train_colmajor"], dtype=dtype).reshape((3, 6)).T.copy(order="F")
    Xq = np.array(c["X_test_colmajor"], dtype=dtype).reshape((3, 3)).T.copy(order="F")
    nn = aoclda.neighbors.nearest_neighbors(metric=metric, radius=c["radius"]).fit(Xt)
    dist, ind = nn.radius_neighbors(Xq, sort_results=True)
    rtol = 1e-14 if dtype == np.float64 else 1e-6
    for q in range(3):
        np.testing.assert_array_equal(ind[q], c[f"radius_{metric}_ind"][q])
        np.testing.assert_allclose(dist[q], c[f"radius_{metric}_dist"][q], rtol=rtol)
    ind = nn.radius_neighbors(Xq, return_distance=False)
    for q in range(3):
        np.testing.assert_array_equal(ind[q], sorted(c[f"radius_{metric}_ind"][q]))


@pytest.mark.parametrize("metric,p", [("euclidean", 2.0), ("sqeuclidean", 2.0),
                                      ("manhattan", 2.0), ("cosine", 2.0), ("minkowski", 3.0),
                                      ("minkowski", 2.0), ("euclidean_gemm", 2.0),
                                      ("sqeuclidean_gemm", 2.0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sort", [False, True])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle(oracle, metric, p, dtype, sort, order):
    rng = np.random.default_rng(41)
    Xt = rng.standard_normal((1500, 24)).astype(dtype)
    Xq = rng.standard_normal((300, 24)).astype(dtype)
    # a radius that keeps about 2% of the training rows per query
    Xt64, Xq64 = Xt.astype(np.float64), Xq.astype(np.float64)
    _, _, d_all = oracle.radius(Xt64[:200], Xq64[:20], 1e300, metric, p)
    r = float(np.quantile(d_all, 0.02))
    single = dtype == np.float32
    counts, eind, edist = oracle.radius(Xt64, Xq64, r, metric, p, single=single, sort=sort)
    assert 0 < counts.sum() < Xt.shape[0] * Xq.shape[0]
    nn = aoclda.neighbors.nearest_neighbors(metric=metric, p=p).fit(
        np.asarray(Xt, order=order))
    dist, ind = nn.radius_neighbors(np.asarray(Xq, order=order), radius=r, sort_results=sort)
    check_against_oracle(dist, ind, counts, eind, edist, 1e-12 if not single else 1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_query_blocks(oracle, dtype):
    """Key tiles are processed in query blocks of at most 256 MB: 600 queries against 150,000
    training rows take several blocks (fp64: 223 rows a block, fp32: 447)."""
    rng = np.random.default_rng(5)
    Xt = rng.standard_normal((150000, 6)).astype(dtype)
    Xq = rng.standard_normal((600, 6)).astype(dtype)
    r = 1.1
    counts, eind, edist = oracle.radius(Xt.astype(np.float64), Xq.astype(np.float64), r,
                                        single=dtype == np.float32, sort=True)
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    dist, ind = nn.radius_neighbors(Xq, radius=r, sort_results=True)
    check_against_oracle(dist, ind, counts, eind, edist,
                         1e-12 if dtype == np.float64 else 1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_edge_cases(oracle, dtype):
    rng = np.random.default_rng(8)
    Xt = rng.standard_normal((301, 5)).astype(dtype)
    Xq = np.vstack([Xt[:7], rng.standard_normal((9, 5))]).astype(dtype)
    nn = aoclda.neighbors.nearest_neighbors(radius=0.0).fit(Xt)
    # radius 0 from the constructor: only exact copies of training rows match
    dist, ind = nn.radius_neighbors(Xq)
    for q in range(16):
        assert list(ind[q]) == ([q] if q < 7 else [])
        assert np.all(dist[q] == 0)
    # no neighbour at all
    far = np.full((3, 5), 1e3, dtype=dtype)
    dist, ind = nn.radius_neighbors(far, radius=1.0)
    assert all(len(v) == 0 for v in ind) and all(len(v) == 0 for v in dist)
    # every training row a neighbour
    dist, ind = nn.radius_neighbors(Xq[:2], radius=1e6, sort_results=True)
    counts, eind, edist = oracle.radius(Xt.astype(np.float64), Xq[:2].astype(np.float64), 1e6,
                                        single=dtype == np.float32, sort=True)
    check_against_oracle(dist, ind, counts, eind, edist,
                         1e-12 if dtype == np.float64 else 1e-6)
    # a single query and a single training row
    nn1 = aoclda.neighbors.nearest_neighbors().fit(Xt[:1])
    dist, ind = nn1.radius_neighbors(Xt[:1], radius=0.5)
    assert list(ind[0]) == [0]


def test_results_interface_and_errors():
    """da_handle_get_result_? for the radius queries and the reference's error statuses."""
    rng = np.random.default_rng(2)
    Xt = rng.standard_normal((64, 3))
    Xq = rng.standard_normal((5, 3))
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    h = nn.handle.h
    n = L.da_int(6)
    out = np.zeros(6, dtype=np.int32)
    # nothing computed yet
    assert L.lib.da_handle_get_result_int(h, L.RNN_COUNT, C.byref(n), L.ptr(out)) == L.CODE["no_data"]
    dist, ind = nn.radius_neighbors(Xq, radius=1.2)
    counts = np.array([len(v) for v in ind])
    total = int(counts.sum())
    assert total > 0
    # counts: n_queries + 1 entries, the total last
    n = L.da_int(6)
    assert L.lib.da_handle_get_result_int(h, L.RNN_COUNT, C.byref(n), L.ptr(out)) == 0
    assert list(out[:5]) == list(counts) and out[5] == total
    # offsets: -1 for a query without neighbours, the total last
    n = L.da_int(6)
    assert L.lib.da_handle_get_result_int(h, L.RNN_OFFSETS, C.byref(n), L.ptr(out)) == 0
    off = 0
    for q in range(5):
        assert out[q] == (off if counts[q] > 0 else -1)
        off += counts[q]
    assert out[5] == total
    # too small an array: the needed size comes back in dim
    small = L.da_int(2)
    assert (L.lib.da_handle_get_result_int(h, L.RNN_COUNT, C.byref(small), L.ptr(out)) ==
            L.CODE["invalid_array_dimension"]) and small.value == 6
    # one query's indices and distances, its index passed in result[0]
    q = int(np.argmax(counts))
    buf = np.zeros(64, dtype=np.int32)
    buf[0] = q
    n = L.da_int(64)
    assert L.lib.da_handle_get_result_int(h, L.RNN_INDICES_INDEX, C.byref(n), L.ptr(buf)) == 0
    np.testing.assert_array_equal(buf[:counts[q]], ind[q])
    dbuf = np.zeros(64)
    dbuf[0] = q
    n = L.da_int(64)
    assert L.lib.da_handle_get_result_d(h, L.RNN_DISTANCES_INDEX, C.byref(n), L.ptr(dbuf)) == 0
    np.testing.assert_array_equal(dbuf[:counts[q]], dist[q])
    buf[0] = 5
    n = L.da_int(64)
    assert (L.lib.da_handle_get_result_int(h, L.RNN_INDICES_INDEX, C.byref(n), L.ptr(buf)) ==
            L.CODE["invalid_input"])
    # distances not requested
    nn.radius_neighbors(Xq, radius=1.2, return_distance=False)
    n = L.da_int(64)
    assert L.lib.da_handle_get_result_d(h, L.RNN_DISTANCES, C.byref(n), L.ptr(dbuf)) == L.CODE["no_data"]
    # sorting without distances
    Xc = np.ascontiguousarray(Xq)
    st = L.lib.da_nn_radius_neighbors_d(h, 5, 3, L.ptr(Xc), 3, 1.0, 0, 1)
    assert st == L.CODE["invalid_input"]
    with pytest.raises(ValueError):
        nn.radius_neighbors(Xq, return_distance=False, sort_results=True)
    with pytest.raises(ValueError):
        nn.radius_neighbors(Xq, radius=-1.0)
    # feature count mismatch
    Xw = np.zeros((5, 4))
    st = L.lib.da_nn_radius_neighbors_d(h, 5, 4, L.ptr(Xw), 4, 1.0, 1, 0)
    assert st == L.CODE["invalid_array_dimension"]
    # the inner product metric has no radius search
    ip = aoclda.neighbors.nearest_neighbors(metric="inner_product").fit(Xt)
    st = L.lib.da_nn_radius_neighbors_d(ip.handle.h, 5, 3, L.ptr(Xc), 3, 1.0, 1, 0)
    assert st == L.CODE["incompatible_options"]
    # no training data
    fresh = L.Handle(L.HANDLE_NN, np.float64)
    st = L.lib.da_nn_radius_neighbors_d(fresh.h, 5, 3, L.ptr(Xc), 3, 1.0, 1, 0)
    assert st == L.CODE["no_data"]


def test_kneighbors_after_radius(oracle):
    """A radius search leaves the k-nearest-neighbour path of the same handle intact."""
    rng = np.random.default_rng(9)
    Xt = rng.standard_normal((2000, 16))
    Xq = rng.standard_normal((50, 16))
    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=7).fit(Xt)
    nn.radius_neighbors(Xq, radius=3.0)
    dist, ind = nn.kneighbors(Xq)
    ei, ed = oracle.knn(Xt, Xq, 7)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12)
