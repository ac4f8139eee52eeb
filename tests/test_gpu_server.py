"""Serving loop (da_nn_server_*): continuous batching must not change any result. Each
request's neighbours are compared alone, mixed with others, and against the oracle."""
import numpy as np
import pytest

import aoclda
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
