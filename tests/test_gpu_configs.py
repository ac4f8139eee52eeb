"""The benchmark configurations at their full sizes on the GPU:

  * brute-force kNN, 200,000 x 128 training rows and 20,000 queries from a 64-component
    Gaussian mixture, k = 10, fp32 (plain and split 16-bit cross terms) and fp64;
  * k-means, 2,000,000 x 32 from 256 blobs, k = 256, k-means++ seed, 25 iterations, fp32 data
    with the bf16 cross term;
  * batched k-means, 1,024 uneven instances, 15 iterations;
  * the serving loop over a resident 2,000,000 x 96 index of unit rows, k in {1, 10, 100},
    with the plain fp16 and the split fp16x3 cross terms.

At these sizes the oracle checks a sample (the same seeded inputs), and every result is
checked in full through properties that do not depend on the size: ascending distances that
equal a recomputation over the returned indices, no sampled training row nearer than the
k-th neighbour unless it is one of them; labels that are the argmin over the returned
centres, an inertia that equals its recomputation; results that do not depend on batching."""
import numpy as np
import pytest

import aoclda
from aoclda.neighbors import nn_server

pytestmark = pytest.mark.gpu


def _mixture(rng, n, means, sigma=1.0):
    comp = rng.integers(0, len(means), n)
    return (means[comp] + sigma * rng.standard_normal((n, means.shape[1]))).astype(np.float32)


@pytest.fixture(scope="module")
def knn_data(oracle):
    rng = np.random.default_rng(0)
    means = rng.uniform(-10, 10, (64, 128))
    Xt = _mixture(rng, 200_000, means)
    Xq = _mixture(rng, 20_000, means)
    s = 300
    ei, ed = oracle.knn(Xt.astype(np.float64), Xq[:s].astype(np.float64), 10)
    # A fixed random sample of training rows for the full-size property check.
    sample = np.sort(rng.choice(len(Xt), 2000, replace=False))
    return Xt, Xq, ei, ed, sample


def _check_knn(Xt, Xq, ind, dist, ei, ed, sample, rtol):
    s = len(ei)
    np.testing.assert_array_equal(ind[:s], ei)
    np.testing.assert_allclose(dist[:s], ed, rtol=rtol)
    # ascending distances
    assert (np.diff(dist.astype(np.float64), axis=1) >= 0).all()
    # distances equal the distance to the returned rows (recomputed in double)
    X64, Q64 = Xt.astype(np.float64), Xq.astype(np.float64)
    rec = np.sqrt(((X64[ind] - Q64[:, None, :]) ** 2).sum(axis=2))
    np.testing.assert_allclose(dist, rec, rtol=rtol)
    # no sampled training row outside the result is nearer than the k-th neighbour
    Ys = X64[sample]
    for q0 in range(0, len(Xq), 4000):
        Qb = Q64[q0:q0 + 4000]
        Ds = np.sqrt(np.maximum(0.0, (Qb ** 2).sum(1)[:, None] + (Ys ** 2).sum(1)[None, :]
                                - 2.0 * Qb @ Ys.T))
        kth = rec[q0:q0 + 4000, -1:]
        inres = (ind[q0:q0 + 4000, :, None] == sample[None, None, :]).any(axis=1)
        assert ((Ds >= kth * (1 - 1e-9) - 1e-9) | inres).all()


@pytest.mark.parametrize("precision", ["working", "bf16x3", "fp16x3", "fp16c"])
def test_knn_200000x128_fp32(knn_data, precision):
    Xt, Xq, ei, ed, sample = knn_data
    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=10, precision=precision).fit(Xt)
    dist, ind = nn.kneighbors(Xq)
    _check_knn(Xt, Xq, ind, dist, ei, ed, sample, 1e-5)


def test_knn_200000x128_fp64(knn_data):
    Xt, Xq, ei, ed, sample = knn_data
    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=10).fit(Xt.astype(np.float64))
    dist, ind = nn.kneighbors(Xq.astype(np.float64))
    _check_knn(Xt, Xq, ind, dist, ei, ed, sample, 1e-12)


def test_kmeans_2000000x32_k256_bf16():
    rng = np.random.default_rng(3)
    n, d, k = 2_000_000, 32, 256
    means = rng.uniform(-20, 20, (k, d))
    sig = rng.uniform(0.5, 2.0, k)
    comp = rng.integers(0, k, n)
    X = (means[comp] + sig[comp, None] * rng.standard_normal((n, d))).astype(np.float32)
    del comp
    runs = {}
    for prec in ("bf16", "working"):
        runs[prec] = aoclda.clustering.kmeans(n_clusters=k, n_init=1, seed=42, max_iter=25,
                                              tol=0.0, precision=prec).fit(X)
    km, ref = runs["bf16"], runs["working"]
    # the bf16 cross term with its exact re-check assigns exactly as the fp32 one
    np.testing.assert_array_equal(km.labels, ref.labels)
    np.testing.assert_array_equal(km.cluster_centres, ref.cluster_centres)
    assert km.n_iter == ref.n_iter
    C = km.cluster_centres.astype(np.float64)
    lab = km.labels
    # inertia equals its recomputation over the returned labels and centres
    inert = 0.0
    for b in range(0, n, 250_000):
        Xb = X[b:b + 250_000].astype(np.float64)
        inert += ((Xb - C[lab[b:b + 250_000]]) ** 2).sum()
    assert abs(inert - km.inertia) <= 1e-5 * km.inertia
    # labels are the argmin over the centres on a sample (a near tie between centres that
    # the float rounding of the returned centres can swap is allowed)
    idx = rng.choice(n, 20_000, replace=False)
    Xs = X[idx].astype(np.float64)
    D = (Xs ** 2).sum(1)[:, None] + (C ** 2).sum(1)[None, :] - 2.0 * Xs @ C.T
    best = D.min(axis=1)
    got = D[np.arange(len(idx)), lab[idx]]
    assert (got <= best + 1e-5 * np.maximum(1.0, np.abs(best))).all()
    # when the run stopped because no label changed, the centres are the means of the
    # members (empty clusters at the origin)
    if km.n_iter < 25:
        sums = np.zeros((k, d))
        np.add.at(sums, lab, X.astype(np.float64))
        cnt = np.bincount(lab, minlength=k)
        mean = sums / np.maximum(cnt, 1)[:, None]
        np.testing.assert_allclose(C, mean, rtol=1e-6, atol=1e-5)


def test_kmeans_batch_1024_instances(oracle):
    rng = np.random.default_rng(5)
    ni = 1024
    n = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(2e5), ni))).astype(np.int64)
    d = rng.choice(np.array([8, 16, 32, 64]), ni)
    k = rng.integers(4, 65, ni)
    data, inits = [], []
    for i in range(ni):
        means = rng.uniform(-10, 10, (int(k[i]), int(d[i]))).astype(np.float32)
        lab = rng.integers(0, k[i], n[i])
        Xi = means[lab] + rng.standard_normal((int(n[i]), int(d[i])), dtype=np.float32)
        data.append(Xi)
        inits.append(Xi[:k[i]].copy())
    cs, labs, n_iter, inertia = aoclda.clustering.kmeans_batch(data, inits, 15, 0.0, "bf16")
    # the smallest instances against the oracle
    for i in np.argsort(n * k * d)[:6]:
        r = oracle.kmeans(data[i].astype(np.float64), int(k[i]), "supplied",
                          inits[i].astype(np.float64), max_iter=15, tol=0.0)
        np.testing.assert_array_equal(labs[i], r["labels"], err_msg=f"instance {i}")
        assert n_iter[i] == r["n_iter"]
        np.testing.assert_allclose(cs[i], r["centres"], rtol=1e-5, atol=1e-5)
    # every instance: inertia recomputed, labels the argmin on a sample
    for i in range(ni):
        Xi, Ci = data[i].astype(np.float64), cs[i].astype(np.float64)
        rec = ((Xi - Ci[labs[i]]) ** 2).sum()
        assert abs(rec - inertia[i]) <= 1e-4 * max(rec, 1.0), i
        s = Xi[:: max(1, len(Xi) // 200)]
        D = ((s[:, None, :] - Ci[None, :, :]) ** 2).sum(2)
        got = D[np.arange(len(s)), labs[i][:: max(1, len(Xi) // 200)]]
        assert (got <= D.min(1) + 1e-4 * np.maximum(1.0, D.min(1))).all(), i


@pytest.mark.parametrize("precision", ["fp16", "fp16x3"])
def test_serving_2000000x96_alone_equals_mixed(oracle, precision):
    rng = np.random.default_rng(21)
    means = rng.standard_normal((1024, 96))
    comp = rng.integers(0, 1024, 2_000_000)
    Xt = (means[comp] + 0.3 * rng.standard_normal((2_000_000, 96))).astype(np.float32)
    Xt /= np.linalg.norm(Xt, axis=1, keepdims=True)
    del comp

    def req(m):
        c = rng.integers(0, 1024, m)
        Q = (means[c] + 0.3 * rng.standard_normal((m, 96))).astype(np.float32)
        return Q / np.linalg.norm(Q, axis=1, keepdims=True)

    nn = aoclda.neighbors.nearest_neighbors(precision=precision).fit(Xt)
    sizes = [1, 37, 1000, 4096, 5, 300]
    ks = [1, 100, 10, 100, 10, 1]
    reqs = [req(s) for s in sizes]
    srv = nn_server(nn, max_batch_queries=4096, max_device_bytes=1 << 30)
    alone = []
    for X, k in zip(reqs, ks):
        alone.append(srv.wait(srv.submit(X, k)))
    ids = [srv.submit(X, k) for X, k in zip(reqs, ks)]
    mixed = [srv.wait(r) for r in ids]
    X64 = Xt.astype(np.float64)
    for (da, ia, _), (dm, im, _), X, k in zip(alone, mixed, reqs, ks):
        np.testing.assert_array_equal(ia, im)
        np.testing.assert_array_equal(da, dm)
        ei, ed = oracle.knn(X64, X[:4].astype(np.float64), k)
        np.testing.assert_array_equal(ia[:4], ei)
        np.testing.assert_allclose(da[:4], ed, rtol=1e-5, atol=1e-6)
        assert (np.diff(da.astype(np.float64), axis=1) >= 0).all()
    srv.close()
