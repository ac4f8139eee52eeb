"""GPU parity of DBSCAN (da_dbscan_set_data / da_dbscan_compute and its results) against the
reference's test cases and the oracle. Labels, core samples and cluster counts are
integers and must be identical (the reference's cases up to renaming, as its own tests
compare them; the oracle exactly, numbering included)."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu


def first_seen(labels):
    m, out = {}, []
    for v in labels:
        out.append(-1 if v == -1 else m.setdefault(v, len(m)))
    return out


def run_capi(h, dtype, n, d, A, lda, opts):
    for k, v in opts.items():
        if isinstance(v, str):
            h.set_string(k, v)
        elif isinstance(v, int):
            h.set_int(k, v)
        else:
            h.set_real(k, v)
    A = np.ascontiguousarray(A, dtype=dtype)
    h.call("da_dbscan_set_data", n, d, L.ptr(A), lda)
    h.call("da_dbscan_compute")
    one = L.da_int(1)
    nc, ncore = np.zeros(1, np.int32), np.zeros(1, np.int32)
    L.check(L.lib.da_handle_get_result_int(h.h, L.DBSCAN_N_CLUSTERS, C.byref(one), L.ptr(nc)))
    L.check(L.lib.da_handle_get_result_int(h.h, L.DBSCAN_N_CORE, C.byref(one), L.ptr(ncore)))
    labels = np.zeros(n, np.int32)
    dim = L.da_int(n)
    L.check(L.lib.da_handle_get_result_int(h.h, L.DBSCAN_LABELS, C.byref(dim), L.ptr(labels)))
    core = np.zeros(max(int(ncore[0]), 1), np.int32)
    if ncore[0] > 0:
        dim = L.da_int(int(ncore[0]))
        L.check(L.lib.da_handle_get_result_int(h.h, L.DBSCAN_CORE_INDICES, C.byref(dim),
                                               L.ptr(core)))
    rinfo = np.zeros(9, dtype=dtype)
    dim = L.da_int(9)
    h.call("da_handle_get_result", L.RINFO, C.byref(dim), L.ptr(rinfo))
    return int(nc[0]), labels, core[:int(ncore[0])], rinfo


def case_opts(c):
    return {"algorithm": c["algorithm"], "metric": c["metric"], "storage order": c["order"],
            "min samples": c["min_samples"], "leaf size": c["leaf_size"], "eps": c["eps"],
            "power": c["power"]}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_cases(known, dtype):
    """Every case of dbscan_test_data.hpp through the C API: status, cluster and core counts,
    rinfo, labels (up to renaming) and the sorted core sample indices."""
    for c in known["dbscan"]:
        h = L.Handle(L.HANDLE_DBSCAN, dtype)
        tr = np.zeros(1, np.int32)
        one = L.da_int(1)
        nc, labels, core, rinfo = run_capi(h, dtype, c["n_samples"], c["n_features"], c["A"],
                                           c["lda"], case_opts(c))
        L.check(L.lib.da_handle_get_result_int(h.h, L.TRAINED, C.byref(one), L.ptr(tr)))
        assert tr[0] == 1
        assert nc == c["n_clusters"], c["name"]
        assert len(core) == c["n_core"], c["name"]
        assert first_seen(labels) == first_seen(c["labels"]), c["name"]
        assert sorted(core) == sorted(c["core"]), c["name"]
        np.testing.assert_array_equal(rinfo, np.array(c["rinfo"], dtype=dtype), err_msg=c["name"])


def test_multiple_calls_one_handle(known):
    """One handle computes the 30 x 3 cases and then the 25 x 2 cases (MultipleCalls of
    dbscan_public.cpp)."""
    for dtype in (np.float32, np.float64):
        h = L.Handle(L.HANDLE_DBSCAN, dtype)
        for c in known["dbscan"]:
            if not (c["name"].startswith("30 by 3") or c["name"].startswith("25 by 2")):
                continue
            nc, labels, core, rinfo = run_capi(h, dtype, c["n_samples"], c["n_features"],
                                               c["A"], c["lda"], case_opts(c))
            assert nc == c["n_clusters"] and first_seen(labels) == first_seen(c["labels"])


def blobs(rng, n, d, k, spread, noise):
    centres = rng.uniform(-8, 8, (k, d))
    X = np.vstack([c + rng.normal(0, spread, (n // k, d)) for c in centres] +
                  [rng.uniform(-10, 10, (noise, d))])
    return X[rng.permutation(len(X))]


@pytest.mark.parametrize("metric,eps,p", [("euclidean", 0.8, 2.0), ("l2", 0.8, 2.0),
                                          ("sqeuclidean", 0.64, 2.0), ("manhattan", 1.7, 2.0),
                                          ("cosine", 0.01, 2.0), ("minkowski", 0.75, 3.0),
                                          ("minkowski", 0.64, 2.0), ("minkowski", 1.7, 1.0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle(oracle, metric, eps, p, dtype, order):
    rng = np.random.default_rng(5)
    X = blobs(rng, 2400, 6, 8, 0.5, 200).astype(dtype)
    for ms in (1, 4, 15):
        db = aoclda.clustering.DBSCAN(min_samples=ms, metric=metric, eps=eps, power=p,
                                      algorithm="brute").fit(np.asarray(X, order=order))
        el, ecore, enc = oracle.dbscan(X.astype(np.float64), eps, ms, metric, p,
                                       single=dtype == np.float32)
        np.testing.assert_array_equal(db.labels, el, err_msg=f"min_samples {ms}")
        np.testing.assert_array_equal(db.core_sample_indices, np.nonzero(ecore)[0])
        assert db.n_clusters == enc and db.n_core_samples == int(ecore.sum())
        assert db.n_samples == len(X) and db.n_features == 6


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_row_blocks_and_long_chains(oracle, dtype):
    """13,000 samples take two 1 GiB key blocks in fp64 and a neighbour list past its first
    allocation; a chain of points spaced just under eps forms one cluster through thousands
    of links."""
    rng = np.random.default_rng(17)
    X = blobs(rng, 11000, 3, 5, 0.35, 500)
    chain = np.zeros((1500, 3))
    chain[:, 0] = np.arange(1500) * 0.29 - 200.0
    X = np.vstack([X, chain]).astype(dtype)
    db = aoclda.clustering.DBSCAN(min_samples=3, eps=0.3, algorithm="brute").fit(X)
    el, ecore, enc = oracle.dbscan(X.astype(np.float64), 0.3, 3, single=dtype == np.float32)
    np.testing.assert_array_equal(db.labels, el)
    assert db.n_clusters == enc
    assert len(set(db.labels[-1500:])) == 1 and db.labels[-1] != -1


def test_error_exits():
    """ErrorExits, BadHandleTests and IncorrectHandlePrecision of dbscan_public.cpp."""
    A = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12], dtype=np.float64)
    h = L.Handle(L.HANDLE_DBSCAN, np.float64)
    dim = L.da_int(1)
    r = np.zeros(1)
    ri = np.zeros(1, np.int32)
    assert L.lib.da_dbscan_compute_d(h.h) == L.CODE["no_data"]
    assert L.lib.da_handle_get_result_d(h.h, L.RINFO, C.byref(dim), L.ptr(r)) == L.CODE["no_data"]
    assert L.lib.da_handle_get_result_int(h.h, L.RINFO, C.byref(dim), L.ptr(ri)) == L.CODE["no_data"]
    h.set_string("algorithm", "kd tree")
    h.set_string("metric", "cosine")
    assert L.lib.da_dbscan_set_data_d(h.h, 4, 3, L.ptr(A), 4) == 0
    assert L.lib.da_dbscan_compute_d(h.h) == L.CODE["incompatible_options"]
    h.set_string("metric", "minkowski")
    h.set_real("power", 0.5)
    assert L.lib.da_dbscan_set_data_d(h.h, 4, 3, L.ptr(A), 4) == 0
    assert L.lib.da_dbscan_compute_d(h.h) == L.CODE["incompatible_options"]
    h.set_string("algorithm", "brute")
    h.set_string("metric", "Euclidean")
    h.set_int("min samples", 2)
    h.set_real("eps", 20.0)
    assert L.lib.da_dbscan_set_data_d(h.h, 4, 3, L.ptr(A), 4) == 0
    assert L.lib.da_dbscan_compute_d(h.h) == 0
    assert L.lib.da_handle_get_result_d(h.h, L.RINFO, C.byref(dim), None) == L.CODE["invalid_input"]
    assert L.lib.da_handle_get_result_int(h.h, L.RINFO, None, L.ptr(ri)) == L.CODE["invalid_input"]
    assert L.lib.da_handle_get_result_int(h.h, L.RINFO, C.byref(dim), L.ptr(ri)) == L.CODE["unknown_query"]
    assert L.lib.da_handle_get_result_d(h.h, 1000, C.byref(dim), L.ptr(r)) == L.CODE["unknown_query"]
    dim = L.da_int(1)
    assert (L.lib.da_handle_get_result_d(h.h, L.RINFO, C.byref(dim), L.ptr(r)) ==
            L.CODE["invalid_array_dimension"]) and dim.value == 9
    dim = L.da_int(0)
    assert (L.lib.da_handle_get_result_int(h.h, L.DBSCAN_LABELS, C.byref(dim), L.ptr(ri)) ==
            L.CODE["invalid_array_dimension"]) and dim.value == 4
    dim = L.da_int(1)
    assert (L.lib.da_handle_get_result_int(h.h, L.DBSCAN_CORE_INDICES, C.byref(dim), L.ptr(ri)) ==
            L.CODE["invalid_array_dimension"]) and dim.value == 4
    # an option change drops the results
    h.set_int("min samples", 3)
    dim = L.da_int(4)
    out = np.zeros(4, np.int32)
    assert L.lib.da_handle_get_result_int(h.h, L.DBSCAN_LABELS, C.byref(dim), L.ptr(out)) == L.CODE["no_data"]
    # argument checks of set_data
    assert L.lib.da_dbscan_set_data_d(h.h, 0, 3, L.ptr(A), 4) == L.CODE["invalid_array_dimension"]
    assert L.lib.da_dbscan_set_data_d(h.h, 4, 3, None, 4) == L.CODE["invalid_pointer"]
    assert L.lib.da_dbscan_set_data_d(h.h, 4, 3, L.ptr(A), 3) == L.CODE["invalid_leading_dimension"]
    # handle not initialised, wrong handle type, wrong precision
    assert L.lib.da_dbscan_compute_d(None) == L.CODE["handle_not_initialized"]
    nn = L.Handle(L.HANDLE_NN, np.float64)
    assert L.lib.da_dbscan_set_data_d(nn.h, 4, 3, L.ptr(A), 4) == L.CODE["invalid_handle_type"]
    assert L.lib.da_dbscan_compute_d(nn.h) == L.CODE["invalid_handle_type"]
    hs = L.Handle(L.HANDLE_DBSCAN, np.float32)
    assert L.lib.da_dbscan_set_data_d(hs.h, 4, 3, L.ptr(A), 4) == L.CODE["wrong_type"]
    assert L.lib.da_dbscan_compute_d(hs.h) == L.CODE["wrong_type"]
    assert L.lib.da_dbscan_compute_s(h.h) == L.CODE["wrong_type"]
