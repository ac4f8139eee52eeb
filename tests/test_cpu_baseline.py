"""The CPU baseline that bench.py times beside the GPU path (oracle/cpu_knn.cpp, the
restated blocked host kNN of the reference) returns the oracle's neighbours: a baseline that
computed something else would time the wrong work."""
import numpy as np
import pytest


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,d,nq,k,threads", [(2000, 64, 1500, 5, 4), (1237, 37, 301, 10, 3),
                                              (700, 130, 90, 16, 1), (300, 3, 50, 7, 2),
                                              (1030, 513, 1025, 3, 4)])
def test_matches_oracle(oracle, dtype, n, d, nq, k, threads):
    from oracle import cpu_baseline as cb
    rng = np.random.default_rng(n + d)
    X = rng.random((n, d)).astype(dtype)
    Y = rng.random((nq, d)).astype(dtype)
    ind, dist = cb.knn(X, Y, k, threads, native=False)
    ei, ed = oracle.knn(X.astype(np.float64), Y.astype(np.float64), k)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12 if dtype == np.float64 else 1e-5)


def test_replace_max_ties_take_later_rows():
    """Equal distances: a later row replaces the first maximum (<=, first index of the
    largest), so with every distance equal slot 0 takes each later row in turn while slots 1
    and 2 keep rows 1 and 2; the stable sort keeps the slot order."""
    from oracle import cpu_baseline as cb
    X = np.zeros((40, 2))
    ind, dist = cb.knn(X, np.zeros((1, 2)), 3, 1, native=False)
    assert (dist == 0).all()
    assert ind[0].tolist() == [39, 1, 2]
