"""GPU parity of k-NN / radius classification and regression (da_nn_set_labels / targets,
da_nn_classes, da_nn_classifier_predict_proba, da_nn_classifier_predict,
da_nn_regressor_predict) against the reference's answers and the oracle. Labels are
identical; probabilities and targets within 1e-12 (fp64) and 2e-6 (fp32) relative."""
import ctypes as C

import numpy as np
import pytest

import aoclda
from aoclda import _lib as L

pytestmark = pytest.mark.gpu

FAMILY = {"euclidean": "euclidean", "l2": "euclidean", "minkowski": "euclidean",
          "euclidean_gemm": "euclidean", "sqeuclidean": "sqeuclidean",
          "sqeuclidean_gemm": "sqeuclidean"}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("metric", list(FAMILY))
def test_reference_answers(known, dtype, metric):
    c = known["knn_predict"]
    Xt = np.array(c["X_train_colmajor"], dtype=dtype).reshape(3, 6).T.copy(order="F")
    Xq = np.array(c["X_test_colmajor"], dtype=dtype).reshape(3, 3).T.copy(order="F")
    tol = 1e-12 if dtype == np.float64 else 2e-6
    fam = FAMILY[metric]
    for w in ("uniform", "distance"):
        for k in ("3", "5"):
            nn = aoclda.neighbors.nearest_neighbors(n_neighbors=int(k), weights=w, metric=metric,
                                                    radius=c["radius"])
            nn.fit(Xt, np.array(c["y_class"]))
            proba = nn.classifier_predict_proba(Xq)
            np.testing.assert_allclose(proba, np.array(c["proba_knn"][k][w][fam]).reshape(3, 3).T,
                                       rtol=tol, atol=tol)
            assert list(nn.classifier_predict(Xq)) == c["labels_knn"][k][w]
            assert list(nn.classes) == [0, 1, 2]
            nn.y = np.array(c["y_reg"])
            np.testing.assert_allclose(nn.regressor_predict(Xq), c["targets_knn"][k][w][fam],
                                       rtol=tol)
        nn = aoclda.neighbors.nearest_neighbors(n_neighbors=3, weights=w, metric=metric,
                                                radius=c["radius"]).fit(Xt, np.array(c["y_class"]))
        proba = nn.classifier_predict_proba(Xq, search_mode="radius_neighbors")
        np.testing.assert_allclose(proba, np.array(c["proba_radius"][w][fam]).reshape(3, 3).T,
                                   rtol=tol, atol=tol)
        assert list(nn.classifier_predict(Xq, search_mode="radius_neighbors")) == \
            c["labels_radius"][fam]
        nn.y = np.array(c["y_reg"])
        np.testing.assert_allclose(nn.regressor_predict(Xq, search_mode="radius_neighbors"),
                                   c["targets_radius"][w][fam], rtol=tol)


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "manhattan", "cosine",
                                    "minkowski"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("weights", ["uniform", "distance"])
@pytest.mark.parametrize("order", ["C", "F"])
def test_vs_oracle(oracle, metric, dtype, weights, order):
    rng = np.random.default_rng(13)
    means = rng.uniform(-4, 4, (5, 7))
    lab = rng.integers(0, 5, 3000)
    Xt = (means[lab] + rng.standard_normal((3000, 7))).astype(dtype)
    y = np.array([-3, 7, 10, 42, 5])[lab]
    t = rng.standard_normal(3000)
    Xq = (means[rng.integers(0, 5, 400)] + rng.standard_normal((400, 7))).astype(dtype)
    single = dtype == np.float32
    tol = 1e-12 if not single else 2e-6
    p = 3.0 if metric == "minkowski" else 2.0
    r = {"euclidean": 3.0, "sqeuclidean": 9.0, "manhattan": 6.5, "cosine": 0.05,
         "minkowski": 2.4}[metric]
    Xt64, Xq64 = Xt.astype(np.float64), Xq.astype(np.float64)
    for k in (1, 7, 40):
        nn = aoclda.neighbors.nearest_neighbors(n_neighbors=k, weights=weights, metric=metric,
                                                p=p, radius=r).fit(np.asarray(Xt, order=order), y)
        eproba, elab, _ = oracle.nn_classify(Xt64, y, Xq64, k=k, metric=metric, p=p,
                                             weights=weights, single=single)
        np.testing.assert_allclose(nn.classifier_predict_proba(np.asarray(Xq, order=order)),
                                   eproba, rtol=tol, atol=tol)
        np.testing.assert_array_equal(nn.classifier_predict(np.asarray(Xq, order=order)), elab)
        nn.y = t
        et, _ = oracle.nn_regress(Xt64, t.astype(dtype), Xq64, k=k, metric=metric, p=p,
                                  weights=weights, single=single)
        np.testing.assert_allclose(nn.regressor_predict(np.asarray(Xq, order=order)), et,
                                   rtol=tol * 10, atol=tol)
    # radius mode with every outlier handling
    Xo = np.vstack([Xq[:50], np.full((3, 7), 40.0, dtype=dtype)])
    Xo64 = Xo.astype(np.float64)
    for outlier, val in (("most_frequent", "most_frequent"), ("manual", 7), ("manual", 99)):
        nn = aoclda.neighbors.nearest_neighbors(weights=weights, metric=metric, p=p, radius=r,
                                                outlier_handling=val).fit(Xt, y)
        eproba, elab, first = oracle.nn_classify(
            Xt64, y, Xo64, r=r, metric=metric, p=p, weights=weights, outlier=outlier,
            outlier_label=val if outlier == "manual" else 0, single=single)
        proba = nn.classifier_predict_proba(Xo, search_mode="radius_neighbors")
        np.testing.assert_allclose(proba, eproba, rtol=tol, atol=tol)
        if metric != "cosine":  # the far rows have cosine neighbours
            assert 0 <= first <= 50
        expect_warn = first >= 0 and outlier == "manual" and val == 99
        assert nn.status == (L.CODE["outlier_warning"] if expect_warn else 0)
        np.testing.assert_array_equal(nn.classifier_predict(Xo, search_mode="radius_neighbors"),
                                      elab)
        nn.y = t
        tv = 2.5 if outlier == "manual" else "most_frequent"
        nn.outlier_handling = tv
        et, _ = oracle.nn_regress(Xt64, t.astype(dtype), Xo64, r=r, metric=metric, p=p,
                                  weights=weights, outlier=outlier, outlier_target=2.5,
                                  single=single)
        np.testing.assert_allclose(nn.regressor_predict(Xo, search_mode="radius_neighbors"), et,
                                   rtol=tol * 10, atol=tol)


def test_errors():
    rng = np.random.default_rng(4)
    Xt = rng.standard_normal((50, 3))
    Xq = np.ascontiguousarray(rng.standard_normal((5, 3)))
    y = np.ascontiguousarray(rng.integers(0, 3, 50), dtype=np.int32)
    h = L.Handle(L.HANDLE_NN, np.float64)
    h.set_string("storage order", "row-major")
    out = np.zeros(5, np.int32)
    proba = np.zeros(15)
    # nothing set
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(out), 0) == L.CODE["no_data"]
    h.call("da_nn_set_data", 50, 3, L.ptr(np.ascontiguousarray(Xt)), 3)
    assert L.lib.da_nn_classifier_predict_proba_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(proba), 0) == L.CODE["no_data"]
    assert L.lib.da_nn_regressor_predict_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(proba), 0) == L.CODE["no_data"]
    assert L.lib.da_nn_set_labels_d(h.h, 49, L.ptr(y)) == L.CODE["invalid_array_dimension"]
    assert L.lib.da_nn_set_labels_d(h.h, 50, None) == L.CODE["invalid_pointer"]
    assert L.lib.da_nn_set_labels_d(h.h, 50, L.ptr(y)) == 0
    n = L.da_int(0)
    assert L.lib.da_nn_classes_d(h.h, C.byref(n), None) == 0 and n.value == len(set(y.tolist()))
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 3, L.ptr(Xq), 3, None, 0) == L.CODE["invalid_pointer"]
    assert L.lib.da_nn_classifier_predict_d(h.h, 0, 3, L.ptr(Xq), 3, L.ptr(out), 0) == L.CODE["invalid_array_dimension"]
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 4, L.ptr(Xq), 3, L.ptr(out), 0) == L.CODE["invalid_array_dimension"]
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(out), 7) == L.CODE["invalid_input"]
    # radius search with a query without neighbours and "outlier handling" = none
    h.set_real("radius", 0.01)
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(out), 1) == L.CODE["operation_failed"]
    assert L.lib.da_nn_classifier_predict_d(h.h, 5, 3, L.ptr(Xq), 3, L.ptr(out), 0) == 0
    # wrong precision
    assert L.lib.da_nn_set_labels_s(h.h, 50, L.ptr(y)) == L.CODE["wrong_type"]
    # inner product with 1/d weights
    hi = L.Handle(L.HANDLE_NN, np.float64)
    hi.set_string("storage order", "row-major")
    hi.set_string("metric", "inner product")
    hi.set_string("weights", "distance")
    hi.call("da_nn_set_data", 50, 3, L.ptr(np.ascontiguousarray(Xt)), 3)
    hi.call("da_nn_set_labels", 50, L.ptr(y))
    assert L.lib.da_nn_classifier_predict_d(hi.h, 5, 3, L.ptr(Xq), 3, L.ptr(out), 0) == L.CODE["incompatible_options"]
