"""Serving loop (da_nn_server_*): continuous batching must not change any result. Each
request's neighbours are compared alone, mixed with others, and against the oracle."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L
from aoclda.neighbors import nn_server

pytestmark = pytest.mark.gpu


def unit_rows(rng, n, means, sigma=0.1):
    comp = rng.integers(0, len(means), n)
    X = means[comp] + sigma * rng.standard_normal((n, means.shape[1]))
    return X / np.linalg.norm(X, axis=1, keepdims=True)


@pytest.mark.parametrize("precision", ["working", "fp16", "bf16"])
def test_alone_equals_mixed(oracle, precision):
    rng = np.random.default_rng(31)
    means = rng.standard_normal((64, 96))
    Xt = unit_rows(rng, 20000, means).astype(np.float32)
    nn = aoclda.neighbors.nearest_neighbors(precision=precision).fit(Xt)
    sizes = [1, 3, 17, 250, 700, 2049, 5]
    ks = [1, 10, 16, 10, 1, 16, 10]
    reqs = [unit_rows(rng, s, means).astype(np.float32) for s in sizes]
    srv = nn_server(nn, max_batch_queries=1024, max_device_bytes=64 << 20)
    alone = []
    for X, k in zip(reqs, ks):
        rid = srv.submit(X, k)
        alone.append(srv.wait(rid))
    ids = [srv.submit(X, k) for X, k in zip(reqs, ks)]
    mixed = [srv.wait(r) for r in ids]
    for (da, ia, la), (dm, im, lm), X, k in zip(alone, mixed, reqs, ks):
        np.testing.assert_array_equal(ia, im)
        np.testing.assert_array_equal(da, dm)
        ei, ed = oracle.knn(Xt.astype(np.float64), X.astype(np.float64), k)
        np.testing.assert_array_equal(ia, ei)
        np.testing.assert_allclose(da, ed, rtol=1e-5, atol=1e-6)
        assert la > 0 and lm > 0
    st = srv.stats()
    assert st["requests"] == 2 * len(sizes)
    assert st["queries"] == 2 * sum(sizes)
    # the 2049-query request exceeds a 1024-query batch and is served in slices
    assert st["batches"] >= 4
    srv.close()


def test_argument_errors():
    rng = np.random.default_rng(32)
    Xt = rng.standard_normal((100, 8)).astype(np.float32)
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    srv = nn_server(nn, max_batch_queries=64)
    with pytest.raises(aoclda.DAError):
        srv.submit(np.ones((2, 7), dtype=np.float32), 3)  # wrong feature count
    with pytest.raises(aoclda.DAError):
        srv.submit(np.ones((2, 8), dtype=np.float32), 101)  # k > n_samples
    srv.close()


@pytest.mark.parametrize("precision,dtype", [("fp16x3", np.float32), ("bf16x3", np.float32),
                                             ("working", np.float64)])
def test_mixed_k_and_short_lists(oracle, precision, dtype):
    """Requests with k above 32 (candidate lists shorter than k, the batch searched at its
    largest k and the smaller requests taking a prefix) submitted together with k = 1 and
    k = 50: every result equals the request served alone and the oracle, also for the split
    16-bit forms and an fp64 server."""
    rng = np.random.default_rng(33)
    means = rng.standard_normal((64, 48))
    Xt = unit_rows(rng, 30000, means).astype(dtype)
    nn = aoclda.neighbors.nearest_neighbors(precision=precision).fit(Xt)
    sizes = [40, 7, 300, 129, 1]
    ks = [33, 128, 50, 1, 100]
    reqs = [unit_rows(rng, s, means).astype(dtype) for s in sizes]
    srv = nn_server(nn, max_batch_queries=256, max_device_bytes=64 << 20)
    alone = [srv.wait(srv.submit(X, k)) for X, k in zip(reqs, ks)]
    ids = [srv.submit(X, k) for X, k in zip(reqs, ks)]
    mixed = [srv.wait(r) for r in ids]
    rtol = 1e-12 if dtype == np.float64 else 1e-5
    for (da, ia, _), (dm, im, _), X, k in zip(alone, mixed, reqs, ks):
        np.testing.assert_array_equal(ia, im)
        np.testing.assert_array_equal(da, dm)
        ei, ed = oracle.knn(Xt.astype(np.float64), X.astype(np.float64), k)
        np.testing.assert_array_equal(ia, ei)
        np.testing.assert_allclose(da, ed, rtol=rtol, atol=1e-6 if dtype == np.float32 else 0)
    srv.close()


def test_second_wait_on_a_request_is_refused():
    """A result goes to one waiter: waiting again on a finished request, or from a second
    thread while the first still waits, returns invalid_input instead of touching freed
    state."""
    import threading
    rng = np.random.default_rng(34)
    Xt = rng.standard_normal((5000, 16)).astype(np.float32)
    nn = aoclda.neighbors.nearest_neighbors().fit(Xt)
    srv = nn_server(nn, max_batch_queries=64)
    rid = srv.submit(Xt[:500], 10)
    nq, k = srv._k[rid]
    results = []

    def waiter():
        ind = np.empty((nq, k), dtype=np.int32)
        dist = np.empty((nq, k), dtype=np.float32)
        lat = C.c_double()
        results.append(L.lib.da_nn_server_wait_s(srv.h, rid, L.ptr(ind), L.ptr(dist),
                                                 C.byref(lat)))

    th = [threading.Thread(target=waiter) for _ in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert sorted(results) == [L.CODE["success"], L.CODE["invalid_input"]]
    srv._k.pop(rid)
    srv.close()
