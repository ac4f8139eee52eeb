"""ctypes loader of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product library never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

METRICS = {"euclidean": 0, "sqeuclidean": 1, "minkowski": 2, "manhattan": 3, "cosine": 4,
           "euclidean_gemm": 5, "sqeuclidean_gemm": 6, "inner_product": 7}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


if not os.path.exists(LIB):
    build()
_lib = C.CDLL(LIB)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_lib.oracle_pairwise.restype = C.c_int
_lib.oracle_pairwise.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _d, C.c_int, _d, C.c_int,
                                 _d, C.c_int, C.c_double, C.c_int, C.c_double]
_lib.oracle_knn.restype = C.c_int
_lib.oracle_knn.argtypes = [C.c_int, C.c_int, _d, C.c_int, _d, C.c_int, C.c_int, C.c_double,
                            _i, _d, C.c_double]
_lib.oracle_kmeans.restype = C.c_int
_lib.oracle_kmeans.argtypes = [C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_int, C.c_int,
                               C.c_double, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_int, _d, _i, _d]
_lib.oracle_kmeans_init.restype = C.c_int
_lib.oracle_kmeans_init.argtypes = [C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_longlong,
                                    C.c_int, _d]
_lib.oracle_kmeans_sph.restype = C.c_int
_lib.oracle_kmeans_sph.argtypes = [C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_int, C.c_int,
                                   C.c_double, C.c_longlong, C.c_int, C.c_int, C.c_int, _d, _i,
                                   _d]
_lib.oracle_round_lp.restype = C.c_double
_lib.oracle_round_lp.argtypes = [C.c_double, C.c_int]


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


def pairwise(X, Y=None, metric="euclidean", p=2.0, eps=np.finfo(np.float64).eps):
    """Distance matrix (row-major), restating da_pairwise_distances in double."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    m, k = X.shape
    Yc = None if Y is None else np.ascontiguousarray(Y, dtype=np.float64)
    n = m if Yc is None else Yc.shape[0]
    D = np.empty((m, n))
    _lib.oracle_pairwise(0, m, n, k, _p(X), k, _p(Yc), k, _p(D), n, float(p), METRICS[metric],
                         float(eps))
    return D


def knn(Xtrain, Xq, k, metric="euclidean", p=2.0, eps=np.finfo(np.float64).eps):
    """(indices, distances) of the k nearest training rows per query, ordered by
    (distance, index)."""
    Xt = np.ascontiguousarray(Xtrain, dtype=np.float64)
    Q = np.ascontiguousarray(Xq, dtype=np.float64)
    ind = np.empty((Q.shape[0], k), dtype=np.int32)
    dist = np.empty((Q.shape[0], k))
    _lib.oracle_knn(Xt.shape[0], Xt.shape[1], _p(Xt), Q.shape[0], _p(Q), k, METRICS[metric],
                    float(p), _p(ind, _i), _p(dist), float(eps))
    return ind, dist


_lib.oracle_radius.restype = C.c_longlong
_lib.oracle_radius.argtypes = [C.c_int, C.c_int, _d, C.c_int, _d, C.c_int, C.c_double, C.c_double,
                               C.c_int, C.c_int, _i, _i, _d, C.c_longlong, C.c_double]


def radius(Xtrain, Xq, r, metric="euclidean", p=2.0, single=False, sort=False,
           eps=np.finfo(np.float64).eps):
    """(counts, indices, distances) of the radius neighbours of every query: indices and
    distances flattened in query order, each query's list in ascending index or, with sort,
    by distance. single: distances and threshold compared in float."""
    Xt = np.ascontiguousarray(Xtrain, dtype=np.float64)
    Q = np.ascontiguousarray(Xq, dtype=np.float64)
    nq, n = Q.shape[0], Xt.shape[0]
    counts = np.empty(nq, dtype=np.int32)
    cap = max(1, nq * min(n, 64))
    while True:
        ind = np.empty(cap, dtype=np.int32)
        dist = np.empty(cap)
        total = _lib.oracle_radius(n, Xt.shape[1], _p(Xt), nq, _p(Q), METRICS[metric],
                                   float(p), float(r), int(single), int(sort), _p(counts, _i),
                                   _p(ind, _i), _p(dist), cap, float(eps))
        if total <= cap:
            return counts, ind[:total].copy(), dist[:total].copy()
        cap = total  # the first pass only counted: run again with room for every neighbour


_lib.oracle_dbscan.restype = C.c_int
_lib.oracle_dbscan.argtypes = [C.c_int, C.c_int, _d, C.c_int, C.c_double, C.c_double, C.c_int,
                               C.c_int, _i, _i]
DBSCAN_METRICS = {"euclidean": 5, "l2": 0, "sqeuclidean": 6, "manhattan": 3, "l1": 3,
                  "cityblock": 3, "cosine": 4, "minkowski": 2}


def dbscan(X, eps, min_samples, metric="euclidean", p=2.0, single=False):
    """(labels, core flags, n_clusters) of the restated serial DBSCAN; single: keys and
    threshold compared in float."""
    A = np.ascontiguousarray(X, dtype=np.float64)
    n, d = A.shape
    labels = np.empty(n, dtype=np.int32)
    core = np.empty(n, dtype=np.int32)
    nc = _lib.oracle_dbscan(n, d, _p(A), DBSCAN_METRICS[metric], float(p), float(eps),
                            int(min_samples), int(single), _p(labels, _i), _p(core, _i))
    return labels, core.astype(bool), nc


def _lists(Xt, Xq, k, r, metric, p, single):
    """Neighbour lists (indices, distances) per query: the k nearest by (distance, index), or
    those within r in ascending index, with distances rounded to float when single."""
    if r is None:
        ind, dist = knn(Xt, Xq, k, metric, p)
        if single:
            dist = dist.astype(np.float32)
        return [(ind[q], dist[q]) for q in range(len(Xq))]
    counts, ind, dist = radius(Xt, Xq, r, metric, p, single=single)
    b = np.concatenate(([0], np.cumsum(counts)))
    return [(ind[b[q]:b[q + 1]], dist[b[q]:b[q + 1]]) for q in range(len(Xq))]


def _weights(d, weights, T):
    # get_weights (nearest_neighbors.cpp:1021-1033): 1, or 1/d with d <= eps(T) counted as 1
    if weights == "uniform":
        return [T(1)] * len(d)
    eps = np.finfo(T).eps
    return [T(1) if T(v) <= eps else T(1) / T(v) for v in d]


def nn_classify(Xt, y, Xq, k=None, r=None, metric="euclidean", p=2.0, weights="uniform",
                outlier="none", outlier_label=0, single=False):
    """(proba, labels, first query without neighbours or -1) of the k-NN / radius classifier
    restated from nearest_neighbors.cpp:1166-1600: per class the weights of the neighbours in
    list order, normalised by the sum of the class totals; the label of the first largest
    probability; radius queries without neighbours handled by outlier ('none', 'manual',
    'most_frequent'). Arithmetic in float when single."""
    T = np.float32 if single else np.float64
    y = np.asarray(y)
    classes = np.unique(y)
    pos = {c: i for i, c in enumerate(classes)}
    vals, cnts = np.unique(y, return_counts=True)
    mf = vals[np.argmax(cnts)]  # smallest label among the most frequent
    proba = np.zeros((len(Xq), len(classes)), dtype=T)
    labels = np.zeros(len(Xq), dtype=np.int64)
    first = -1
    for q, (ind, d) in enumerate(_lists(Xt, Xq, k, r, metric, p, single)):
        row = [T(0)] * len(classes)
        if len(ind) == 0:
            first = q if first < 0 else first
            fill = pos.get(outlier_label, -1) if outlier == "manual" else (
                pos[mf] if outlier == "most_frequent" else -1)
            if fill >= 0:
                row[fill] = T(1)
        else:
            w = _weights(d, weights, T)
            for j, wj in zip(ind, w):
                row[pos[y[j]]] = T(row[pos[y[j]]] + wj)
            den = T(0)
            for v in row:
                den = T(den + v)
            row = [T(v / den) for v in row]
        proba[q] = row
        best = int(np.argmax(proba[q]))
        if proba[q, best] == 0:
            labels[q] = outlier_label if outlier == "manual" else mf
        else:
            labels[q] = classes[best]
    return proba, labels, first


def nn_regress(Xt, t, Xq, k=None, r=None, metric="euclidean", p=2.0, weights="uniform",
               outlier="none", outlier_target=0.0, single=False):
    """(targets, first query without neighbours or -1) of the k-NN / radius regressor
    restated from nearest_neighbors.cpp:1610-1910: the mean of the neighbours' targets, or
    sum(t w) / sum(w); radius outliers get outlier_target ('manual') or the mean target
    ('most_frequent')."""
    T = np.float32 if single else np.float64
    t = np.asarray(t, dtype=T)
    out = np.zeros(len(Xq), dtype=T)
    first = -1
    mean = T(0)
    for v in t:
        mean = T(mean + v)
    mean = T(mean / T(len(t)))
    for q, (ind, d) in enumerate(_lists(Xt, Xq, k, r, metric, p, single)):
        if len(ind) == 0:
            first = q if first < 0 else first
            out[q] = T(outlier_target) if outlier == "manual" else (
                mean if outlier == "most_frequent" else T(0))
            continue
        w = _weights(d, weights, T)
        num, den = T(0), T(0)
        for j, wj in zip(ind, w):
            num = T(num + T(t[j] * wj))
            den = T(den + wj)
        out[q] = T(num / den) if weights == "distance" else T(num / T(len(ind)))
    return out, first


def _unit_rows(A, T):
    """normalize_rows (da_utils.cpp:857-905) as the GPU stores it: 1/|x| in double, the
    product rounded to the data type, a zero row left at zero."""
    A = np.asarray(A, dtype=np.float64)
    n = np.sqrt((A * A).sum(axis=1))
    inv = np.where(n > 0, 1.0 / np.where(n > 0, n, 1.0), 1.0)
    return (A * inv[:, None]).astype(T).astype(np.float64)


def ivf_search(C, X, Xq, k, n_probe, metric="sqeuclidean", single=False):
    """IVF-Flat add and search restated (approximate_neighbors.cpp:565-842 add, :944-1315
    search) for given centroids C: every row of X (in order of addition) joins the list of
    its nearest centroid (ties to the lower centroid); a query scans the n_probe lists of its
    nearest centroids (by (distance, centroid)) in list order and keeps its k nearest
    candidates, equal distances resolved for the lower list and the earlier position in it as
    the reference's first-seen heap does; padded with index -1 and distance +inf. Distances:
    squared euclidean (square-rooted for "euclidean"); for "inner product" the key -x.y,
    reported as x.y (padding -inf); for "cosine" the rows and queries normalised first, the
    key -x.y reported as 1 - x.y. Exact keys, rounded to float when single. Returns (ind,
    dist, list_sizes)."""
    T = np.float32 if single else np.float64
    metric = metric.replace("_", " ")
    ip = metric in ("inner product", "cosine")
    C = np.asarray(C, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    Q = np.asarray(Xq, dtype=np.float64)
    if metric == "cosine":
        X, Q = _unit_rows(X, T), _unit_rows(Q, T)

    def sq(A, B):
        if ip:
            return (-(A @ B.T)).astype(T)
        return ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1).astype(T)

    lab = np.argmin(-(X @ C.T) if ip else sq(X, C), axis=1)
    sizes = np.bincount(lab, minlength=len(C))
    cd = sq(Q, C)
    ind = np.full((len(Q), k), -1, dtype=np.int64)
    dist = np.full((len(Q), k), np.inf, dtype=T)
    for q in range(len(Q)):
        probes = sorted(range(len(C)), key=lambda c: (cd[q, c], c))[:n_probe]
        cand = np.nonzero(np.isin(lab, probes))[0]
        dq = sq(Q[q:q + 1], X[cand])[0]
        order = sorted(range(len(cand)), key=lambda t: (dq[t], lab[cand[t]], cand[t]))[:k]
        ind[q, :len(order)] = cand[order]
        dist[q, :len(order)] = dq[order]
    if metric == "euclidean":
        dist = np.where(np.isfinite(dist), np.sqrt(dist), dist).astype(T)
    elif metric == "cosine":
        dist = (T(1) + dist).astype(T)
    elif metric == "inner product":
        dist = (-dist).astype(T)
    return ind, dist, sizes


INIT = {"random": 0, "k-means++": 1, "supplied": 2, "random partitions": 3, "afk-mc2": 4}
EMPTY = {"ignore": 0, "error": 1, "split": 2}
ALGORITHM = {"lloyd": 0, "elkan": 1}


def kmeans(X, n_clusters, init="k-means++", C=None, n_init=1, max_iter=300, tol=1e-4, seed=0,
           empty_clusters="ignore", algorithm="lloyd", afk_samples=50, mixed_precision=None,
           lp_tol=1e-2, lp_max_iter=200):
    """Returns dict(status, centres, labels, n_iter, inertia, lp_n_iter) of the restated
    k-means. mixed_precision: None, "float" (the lower-precision phase of double data) or
    "half" (that of float data)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    Ci = None if C is None else np.ascontiguousarray(C, dtype=np.float64)
    Co = np.zeros((n_clusters, d))
    lab = np.zeros(n, dtype=np.int32)
    info = np.zeros(3)
    lp = {None: 0, "float": 1, "half": 2}[mixed_precision]
    st = _lib.oracle_kmeans(n, d, _p(X), n_clusters, INIT[init], _p(Ci), n_init, max_iter,
                            float(tol), int(seed), EMPTY[empty_clusters], ALGORITHM[algorithm],
                            int(afk_samples), lp, float(lp_tol), int(lp_max_iter), _p(Co),
                            _p(lab, _i), _p(info))
    return {"status": st, "centres": Co, "labels": lab, "n_iter": int(info[0]),
            "inertia": float(info[1]), "lp_n_iter": int(info[2])}


def kmeans_sph(X, n_clusters, init="k-means++", C=None, n_init=1, max_iter=300, tol=1e-4,
               seed=0, empty_clusters="ignore", normalize=True, afk_samples=50):
    """Spherical (cosine) k-means, Lloyd: dict(status, centres, labels, n_iter, inertia)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    Ci = None if C is None else np.ascontiguousarray(C, dtype=np.float64)
    Co = np.zeros((n_clusters, d))
    lab = np.zeros(n, dtype=np.int32)
    info = np.zeros(3)
    st = _lib.oracle_kmeans_sph(n, d, _p(X), n_clusters, INIT[init], _p(Ci), n_init, max_iter,
                                float(tol), int(seed), EMPTY[empty_clusters], int(afk_samples),
                                int(bool(normalize)), _p(Co), _p(lab, _i), _p(info))
    return {"status": st, "centres": Co, "labels": lab, "n_iter": int(info[0]),
            "inertia": float(info[1])}


def kmeans_init(X, n_clusters, init="k-means++", seed=0, afk_samples=50):
    """Initial centres (n_clusters x d) of the first run, k-means++ or AFK-MC2."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Co = np.zeros((n_clusters, X.shape[1]))
    st = _lib.oracle_kmeans_init(X.shape[0], X.shape[1], _p(X), n_clusters, INIT[init],
                                 int(seed), int(afk_samples), _p(Co))
    assert st == 0
    return Co


def round_lp(x, mode):
    """The oracle's rounding to the lower precision of mixed-precision refinement (mode 1
    float, 2 half through float), elementwise."""
    f = np.vectorize(lambda v: _lib.oracle_round_lp(float(v), int(mode)))
    return f(np.asarray(x, dtype=np.float64))
