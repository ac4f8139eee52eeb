"""Stored models on the GPU, after the reference's tests/unit_tests/model_persistence/
algorithms/*_persistence_public.cpp: a model is computed, its results taken, it is saved and
the handle destroyed; the model loaded from the file gives the same results, bit for bit, and
the same option values. Cases: k-means (Lloyd / Elkan, random / k-means++), k-NN
classification and regression (several algorithms, metrics, weights, storage orders, radius
search mode), radius-search results stored with the model, and IVF-Flat for every metric."""
import ctypes as C

import numpy as np
import pytest

from aoclda import _lib as L

pytestmark = pytest.mark.gpu
lib = L.lib
st = L.CODE


def sfx(dt):
    return "d" if np.dtype(dt) == np.float64 else "s"


def fn(name, dt):
    return getattr(lib, f"{name}_{sfx(dt)}")


def ok(status, h=None):
    assert status == 0, (status, L.error_message(h) if h else "")


def new(htype, dt):
    h = L.da_handle()
    ok(fn("da_handle_init", dt)(C.byref(h), htype))
    return h


def destroy(h):
    lib.da_handle_destroy(C.byref(h))


def save_load(h, path):
    ok(lib.da_handle_save_model(h, str(path).encode()), h)
    destroy(h)
    h2 = L.da_handle()
    ok(lib.da_handle_load_model(C.byref(h2), str(path).encode()), h2)
    ok(lib.da_handle_print_model_versions(h2), h2)
    return h2


def opt_str(h, name):
    buf = C.create_string_buffer(64)
    n = L.da_int(64)
    ok(lib.da_options_get_string(h, name.encode(), buf, C.byref(n)), h)
    return buf.value.decode()


def opt_int(h, name):
    v = L.da_int()
    ok(lib.da_options_get_int(h, name.encode(), C.byref(v)), h)
    return v.value


def result(h, query, n, dt):
    out = np.zeros(n, dt)
    dim = L.da_int(n)
    ok(fn("da_handle_get_result", dt)(h, query, C.byref(dim), L.ptr(out)), h)
    return out


def result_int(h, query, n):
    out = np.zeros(n, np.int32)
    dim = L.da_int(n)
    ok(lib.da_handle_get_result_int(h, query, C.byref(dim), L.ptr(out)), h)
    return out


def blobs(rng, n, d, k, dt):
    means = rng.uniform(-5, 5, (k, d))
    return (means[rng.integers(0, k, n)] + 0.6 * rng.standard_normal((n, d))).astype(dt)


# kmeans_persistence_public.cpp:47-53 (MacQueen and Hartigan-Wong are not provided here)
KMEANS = [("elkan", "k-means++", 3, 150, 77), ("lloyd", "random", 2, 200, 42),
          ("lloyd", "k-means++", 5, 300, 7), ("elkan", "random", 4, 100, 3)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("algo,init,k,max_iter,seed", KMEANS)
def test_kmeans(tmp_path, dt, algo, init, k, max_iter, seed):
    rng = np.random.default_rng(seed)
    A = blobs(rng, 400, 6, k, dt)
    Xt = blobs(rng, 37, 6, k, dt)
    h = new(L.HANDLE_KMEANS, dt)
    ok(lib.da_options_set_string(h, b"algorithm", algo.encode()), h)
    ok(lib.da_options_set_string(h, b"initialization method", init.encode()), h)
    ok(lib.da_options_set_string(h, b"storage order", b"row-major"), h)
    for name, v in (("n_clusters", k), ("max_iter", max_iter), ("seed", seed), ("n_init", 1)):
        ok(lib.da_options_set_int(h, name.encode(), v), h)
    ok(fn("da_kmeans_set_data", dt)(h, 400, 6, L.ptr(A), 6), h)
    s = fn("da_kmeans_compute", dt)(h)
    assert s in (0, st["maxit"])

    def outputs(h):
        c = result(h, L.KMEANS_CENTRES, k * 6, dt)
        lab = result_int(h, L.KMEANS_LABELS, 400)
        info = result(h, L.RINFO, 6, dt)[:5]
        pred = np.zeros(37, np.int32)
        ok(fn("da_kmeans_predict", dt)(h, 37, 6, L.ptr(Xt), 6, L.ptr(pred)), h)
        tr = np.zeros((37, k), dt)
        ok(fn("da_kmeans_transform", dt)(h, 37, 6, L.ptr(Xt), 6, L.ptr(tr), k), h)
        return c, lab, info, pred, tr

    before = outputs(h)
    h = save_load(h, tmp_path / "kmeans.bin")
    after = outputs(h)
    for x, y in zip(before, after):
        np.testing.assert_array_equal(x, y)
    assert opt_str(h, "algorithm") == algo
    assert opt_str(h, "initialization method") == init
    assert opt_int(h, "max_iter") == max_iter and opt_int(h, "seed") == seed
    destroy(h)


# knn_class_persistence_public.cpp:60-80 (trees are searched by brute force here)
KNN = [("brute", "manhattan", "distance", 4, 30, 2.0, 0, 0.0, "column-major", 2),
       ("brute", "euclidean", "uniform", 3, 30, 2.0, 0, 0.0, "row-major", 1),
       ("kd tree", "euclidean", "uniform", 3, 40, 2.0, 0, 0.0, "column-major", 0),
       ("kd tree", "minkowski", "uniform", 3, 30, 3.0, 0, 0.0, "row-major", 0),
       ("ball tree", "euclidean", "distance", 3, 30, 2.0, 0, 0.0, "row-major", 2),
       ("brute", "euclidean", "uniform", 0, 30, 2.0, 1, 8.0, "column-major", 3),
       ("kd tree", "manhattan", "distance", 0, 35, 2.0, 1, 12.0, "row-major", 1)]


def layout(X, order, pad):
    n, d = X.shape
    if order == "row-major":
        buf = np.zeros((n, d + pad), X.dtype)
        buf[:, :d] = X
        return np.ascontiguousarray(buf), d + pad
    buf = np.zeros((d, n + pad), X.dtype)
    buf[:, :n] = X.T
    return np.ascontiguousarray(buf), n + pad


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", KNN)
def test_knn_classifier_regressor(tmp_path, dt, case):
    algo, metric, weights, k, leaf, p, mode, radius, order, pad = case
    rng = np.random.default_rng(5)
    n, d, nq = 60, 4, 17
    X = blobs(rng, n, d, 3, dt)
    Q = blobs(rng, nq, d, 3, dt)
    y = rng.integers(0, 3, n).astype(np.int32)
    yr = rng.standard_normal(n).astype(dt)
    Xb, ldx = layout(X, order, pad)
    Qb, ldq = layout(Q, order, pad)
    h = new(L.HANDLE_NN, dt)
    ok(lib.da_options_set_string(h, b"algorithm", algo.encode()), h)
    ok(lib.da_options_set_string(h, b"metric", metric.encode()), h)
    ok(lib.da_options_set_string(h, b"storage order", order.encode()), h)
    ok(lib.da_options_set_string(h, b"weights", weights.encode()), h)
    if k:
        ok(lib.da_options_set_int(h, b"number of neighbors", k), h)
    else:
        ok(fn("da_options_set_real", dt)(h, b"radius", radius), h)
    ok(lib.da_options_set_int(h, b"leaf size", leaf), h)
    ok(fn("da_options_set_real", dt)(h, b"minkowski parameter", p), h)
    ok(fn("da_nn_set_data", dt)(h, n, d, L.ptr(Xb), ldx), h)
    ok(fn("da_nn_set_labels", dt)(h, n, L.ptr(y)), h)

    def classify(h):
        out = []
        if k:
            ind = np.zeros(nq * k, np.int32)
            dist = np.zeros(nq * k, dt)
            ok(fn("da_nn_kneighbors", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(ind), L.ptr(dist), k, 1), h)
            out += [ind, dist]
        nc = L.da_int(0)
        ok(fn("da_nn_classes", dt)(h, C.byref(nc), None), h)
        proba = np.zeros(nq * nc.value, dt)
        ok(fn("da_nn_classifier_predict_proba", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(proba), mode), h)
        pred = np.zeros(nq, np.int32)
        ok(fn("da_nn_classifier_predict", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(pred), mode), h)
        return out + [np.array([nc.value]), proba, pred]

    before = classify(h)
    h = save_load(h, tmp_path / "knn_class.bin")
    after = classify(h)
    for x, z in zip(before, after):
        np.testing.assert_array_equal(x, z)
    assert opt_str(h, "storage order") == order and opt_str(h, "algorithm") == algo
    assert opt_str(h, "metric") == metric and opt_int(h, "leaf size") == leaf
    # the loaded model still takes targets and regresses like a fresh one
    ok(fn("da_nn_set_targets", dt)(h, n, L.ptr(yr)), h)
    reg = np.zeros(nq, dt)
    ok(fn("da_nn_regressor_predict", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(reg), mode), h)
    h = save_load(h, tmp_path / "knn_reg.bin")
    reg2 = np.zeros(nq, dt)
    ok(fn("da_nn_regressor_predict", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(reg2), mode), h)
    np.testing.assert_array_equal(reg, reg2)
    destroy(h)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("sort_results", [0, 1])
def test_radius_results_stored(tmp_path, dt, sort_results):
    """radius_neighbors_persistence_public.cpp: the radius results of the saved handle are
    results of the loaded one."""
    rng = np.random.default_rng(9)
    X = blobs(rng, 80, 3, 4, dt)
    Q = blobs(rng, 21, 3, 4, dt)
    h = new(L.HANDLE_NN, dt)
    ok(lib.da_options_set_string(h, b"storage order", b"row-major"), h)
    ok(fn("da_nn_set_data", dt)(h, 80, 3, L.ptr(X), 3), h)
    ok(fn("da_nn_radius_neighbors", dt)(h, 21, 3, L.ptr(Q), 3, 1.5, 1, sort_results), h)

    def radius_out(h):
        cnt = result_int(h, L.RNN_COUNT, 22)
        off = result_int(h, L.RNN_OFFSETS, 22)
        ind = result_int(h, L.RNN_INDICES, max(1, int(cnt[-1])))
        dist = result(h, L.RNN_DISTANCES, max(1, int(cnt[-1])), dt)
        return cnt, off, ind, dist

    before = radius_out(h)
    assert before[0][-1] > 0
    h = save_load(h, tmp_path / "radius.bin")
    for x, z in zip(before, radius_out(h)):
        np.testing.assert_array_equal(x, z)
    destroy(h)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "inner product", "cosine"])
@pytest.mark.parametrize("order", ["row-major", "column-major"])
def test_ivf(tmp_path, dt, metric, order):
    """approximate_neighbors_persistence_public.cpp: a trained and filled index searches the
    same after a save and load, including vectors added in two calls."""
    rng = np.random.default_rng(11)
    n, d, nq = 500, 8, 40
    X = blobs(rng, n, d, 6, dt)
    X2 = blobs(rng, 120, d, 6, dt)
    Q = blobs(rng, nq, d, 6, dt)
    Xb, ldx = layout(X, order, 0)
    X2b, ld2 = layout(X2, order, 1)
    Qb, ldq = layout(Q, order, 0)
    h = new(L.HANDLE_APPROX_NN, dt)
    ok(lib.da_options_set_string(h, b"storage order", order.encode()), h)
    ok(lib.da_options_set_string(h, b"metric", metric.encode()), h)
    ok(lib.da_options_set_int(h, b"n_list", 8), h)
    ok(lib.da_options_set_int(h, b"n_probe", 3), h)
    ok(fn("da_approx_nn_set_training_data", dt)(h, n, d, L.ptr(Xb), ldx), h)
    ok(fn("da_approx_nn_train_and_add", dt)(h), h)
    ok(fn("da_approx_nn_add", dt)(h, 120, d, L.ptr(X2b), ld2), h)

    def search(h):
        ind = np.zeros(nq * 7, np.int32)
        dist = np.zeros(nq * 7, dt)
        ok(fn("da_approx_nn_kneighbors", dt)(h, nq, d, L.ptr(Qb), ldq, L.ptr(ind), L.ptr(dist), 7, 1), h)
        cen = result(h, L.ANN_CENTROIDS, 8 * d, dt)
        sizes = result_int(h, L.ANN_LIST_SIZES, 8)
        return ind, dist, cen, sizes

    before = search(h)
    h = save_load(h, tmp_path / "ivf.bin")
    after = search(h)
    for x, z in zip(before, after):
        np.testing.assert_array_equal(x, z)
    assert opt_int(h, "n_probe") == 3 and opt_str(h, "metric") == metric
    # a loaded index keeps growing
    ok(fn("da_approx_nn_add", dt)(h, 120, d, L.ptr(X2b), ld2), h)
    assert result_int(h, L.ANN_LIST_SIZES, 8).sum() == n + 240
    destroy(h)
