"""Writes the golden fixtures used by the tests from the reference's own test data.

Run once in the build container (the reference checkout is not available where the tests
run):  python tests/golden/make_fixtures.py /root/reference

Outputs (data only: inputs and expected outputs):
  nn_csv_2092x5.npz, nn_csv_2125x20.npz  - tests/data/nn_data/{X_train,X_test,k_ind,k_dist}
                                           CSV files of the reference, k = 5, 0-based indices.
  known_answers.json                     - small known-answer cases transcribed from the
                                           reference's unit tests (file:line in each entry).
"""
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def csv(name, dtype=np.float64):
    return np.loadtxt(os.path.join(REF, "tests", "data", "nn_data", name), delimiter=",",
                      dtype=dtype, ndmin=2)


for ntr, nte, d in ((2092, 2071, 5), (2125, 2068, 20)):
    np.savez_compressed(
        os.path.join(HERE, f"nn_csv_{ntr}x{d}.npz"),
        X_train=csv(f"X_train_{ntr}x{d}.csv"),
        X_test=csv(f"X_test_{nte}x{d}.csv"),
        k_ind=csv(f"k_ind_{ntr}x{nte}x{d}_5.csv", np.int64).astype(np.int32),
        k_dist=csv(f"k_dist_{ntr}x{nte}x{d}_5.csv"),
        source=np.array(f"tests/data/nn_data/*_{ntr}x{nte}x{d}_5.csv (k=5)"))

known = {
    "pairwise_fixed_xy": {
        "source": "tests/unit_tests/metrics/metric_public.cpp:297-364",
        "X_row": [[1.0, 2.0, -3.0], [1.0, -1.0, 2.0], [0.0, -2.0, 4.0]],
        "Y_row": [[4.0, -2.0, 3.0], [3.0, -1.0, 2.0]],
        "p": 4.5,
        "expected": {
            "sqeuclidean": [[61., 38.], [11., 4.], [17., 14.]],
            "manhattan": [[13., 10.], [5., 2.], [5., 6.]],
            "cosine": [[1.4466625002869187, 1.3571428571428572],
                       [0.0902823477053158, 0.1271284390560302],
                       [0.3356361611700803, 0.4023856953328032]],
            "minkowski": [[6.254413691636758, 5.124034794569624],
                          [3.009451461921273, 2.0],
                          [4.001734793974489, 3.1055785043917186]],
            "inner_product": [[-9., -5.], [12., 8.], [16., 10.]],
        },
    },
    "pairwise_fixed_xx": {
        "source": "tests/unit_tests/metrics/metric_public.cpp:366-459",
        "X_row": [[1.0, 4.0, -3.0], [2.0, -1.0, 3.0], [1.0, -2.0, 5.0], [-3.0, 1.0, 3.0]],
        "p": 1.5,
        "expected": {
            "sqeuclidean": [[0., 62., 100., 61.], [62., 0., 6., 29.], [100., 6., 0., 29.],
                            [61., 29., 29., 0.]],
            "manhattan": [[0., 12., 14., 13.], [12., 0., 4., 7.], [14., 4., 0., 9.],
                          [13., 7., 9., 0.]],
            "cosine": [[0., 1.576556660197055, 1.7877263614433763, 1.359937016532678],
                       [1.576556660197055, 0., 0.0728949306988934, 0.8773721321030069],
                       [1.7877263614433763, 0.0728949306988934, 0., 0.5811460917083046],
                       [1.359937016532678, 0.8773721321030069, 0.5811460917083046, 0.]],
            "minkowski": [[0., 8.972707819822075, 11.168500752960059, 9.19738630386655],
                          [8.972707819822075, 0., 2.8567382778502783, 5.811210513661822],
                          [11.168500752960059, 2.8567382778502783, 0., 6.356105477264618],
                          [9.19738630386655, 5.811210513661822, 6.356105477264618, 0.]],
            "inner_product": [[26., -11., -22., -8.], [-11., 14., 19., 2.],
                              [-22., 19., 30., 10.], [-8., 2., 10., 19.]],
        },
    },
    "knn_example": {
        "source": "tests/unit_tests/nearest_neighbors/nearest_neighbors_tests.hpp:97-144,443-470",
        "order": "column-major",
        "X_train_colmajor": [-1., -2., -3., 1., 2., 3., -1., -1., -2., 3., 5., -1.,
                             2., 3., -1., 1., 1., 2.],
        "n_samples": 6, "n_features": 3, "n_queries": 3,
        "X_test_colmajor": [-2., -1., 2., 2., -2., 1., 3., -1., -3.],
        "k3_ind_colmajor": [1, 2, 3, 0, 0, 5, 3, 1, 4],
        "k4_ind_colmajor": [1, 2, 3, 0, 0, 5, 3, 1, 4, 4, 5, 2],
        "k4_dist_colmajor": [3., 2., 4.58257569495584, 3.3166247903554, 3.1622776601683795,
                             5.477225575051661, 3.7416573867739413, 4.242640687119285,
                             5.656854249492381, 5.385164807134504, 5.0990195135927845,
                             6.164414002968976],
        "radius_source": "tests/unit_tests/nearest_neighbors/nearest_neighbors_tests.hpp:"
                         "147-159,431-432 (radius 5, results sorted by distance)",
        "radius": 5.0,
        "radius_euclidean_ind": [[1, 0, 3], [2, 0, 1], [3]],
        "radius_euclidean_dist": [[3., 3.3166247903554, 3.7416573867739413],
                                  [2., 3.1622776601683795, 4.242640687119285],
                                  [4.58257569495584]],
        "radius_sqeuclidean_ind": [[], [2], []],
        "radius_sqeuclidean_dist": [[], [4.], []],
    },
    "kmeans": [
        {
            "name": "three clusters, lloyd",
            "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:250-308,586-595",
            "n_samples": 10, "n_features": 2, "n_clusters": 3, "max_iter": 50,
            "A_colmajor": [1.0, 1.1, 0.5, 0.49, -2.0, -2.0, 0.53, 0.9, 1.2, -1.8,
                           1.0, 1.2, -2.0, -1.9, 0.5, 0.51, -2.1, 0.95, 0.8, 0.6],
            "C_colmajor": [0.5, 0.7, -1.3, 0.5, -1.7, 0.2],
            "tol": 1e-4, "empty_clusters": "ignore",
            "rinfo": [10.0, 2.0, 3.0, 1.0, 0.185475, 0.0],
            "centres_colmajor": [1.05, 0.506667, -1.93333333333, 0.9875, -2.0, 0.53666666666],
            "labels": [0, 0, 1, 1, 2, 2, 1, 0, 0, 2],
            "X_colmajor": [0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, -1.0], "m_samples": 4,
            "X_transform_colmajor": [1.4414077320453085, 1.0500744021258683, 0.9887650125282548,
                                     2.2478114355968564, 2.0631798542810347, 3.042484364973978,
                                     2.0599460618612757, 1.1210312712458608, 2.0064368639179446,
                                     1.9880783574989047, 2.9820220581939925, 2.469640099735629],
            "Y_colmajor": [0.5, 0.5, -1.0, 0.5, -1.0, 0.0], "k_samples": 3,
            "Y_labels": [0, 1, 2],
            "tol_check": "100*sqrt(eps)",
        },
        {
            "name": "larger data matrix, lloyd",
            "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:444-528",
            "n_samples": 80, "n_features": 2, "n_clusters": 16, "max_iter": 30,
            "A_colmajor": [
                2.1, -1.0, 3.0, 2.0, -3.0, -2.2, -2.4, 0.9, 1.0, -2.0, 2.0, 1.0, -2.0,
                -1.0, -3.0, 2.3, 2.1, -1.5, 3.0, 2.0, -3.0, -2.8, -2.8, 1.0, 1.2, -2.0,
                2.1, 3.7, -2.0, -1.0, -3.8, 2.0, 2.0, -3.9, 3.0, 2.0, -3.9, -2.7, -2.0,
                1.0, 1.0, -2.4, 2.1, 3.0, -2.6, -1.5, -3.4, 2.0, 2.0, -1.0, 3.0, 2.6,
                -3.0, -2.6, -2.0, 1.0, 1.3, -2.5, 2.8, 3.5, -2.0, -1.0, -3.0, 2.2, 2.0,
                -1.0, 3.4, 2.0, -3.8, -2.0, -2.0, 1.0, 1.9, -2.1, 2.1, 3.0, -2.0, -1.0,
                -3.4, 2.0, 2.0, -8.0, 3.7, 2.0, -3.2, -2.6, -2.2, 1.0, 1.0, -2.5, 6.0,
                -5.5, -2.0, -1.4, -3.7, 2.0, 2.0, -1.2, 3.4, 2.0, -4.0, -2.0, -2.0, 1.0,
                1.0, -2.0, 2.0, 3.3, -2.0, -1.4, -3.0, 2.3, 2.0, -9.0, 3.0, 2.0, -3.0,
                -2.0, -2.0, 1.7, 1.3, -2.0, 2.3, 3.3, -2.0, -6.0, -3.0, 2.7, 2.0, -1.5,
                3.0, 2.6, -3.3, -2.0, -2.0, 1.0, 1.7, -2.0, 2.0, 8.5, -2.0, -1.7, -3.0,
                2.0, 2.2, -1.2, 3.6, 2.9, -6.0, -2.0, -2.1, 1.5, 9.0, -2.0, 2.0, 3.8,
                -2.0, -1.0, -8.1, 2.9],
            "C_colmajor": [1.9, -5.1, 1.0, -3.0, 1.0, -3.2, 6.0, -3.7, 8.5, -3.0, 1.7, -3.0,
                           1.0, -7.0, 1.0, -3.4, 4.8, -3.0, 1.0, -3.4, 1.9, -3.0, 1.3, -3.8,
                           5.4, -3.2, -8.0, -3.0, -6.1, -3.3, 1.6, -3.0],
            "tol": 1e-4, "empty_clusters": "ignore",
            "rinfo": [80.0, 2.0, 16.0, 4.0, 45.78715, 0.0],
            "centres_colmajor": [2.4666666667, 0.0, 1.02000, -3.0, 2.42593, 0.0, 0.0, -3.7,
                                 0.0, -3.0, -1.0, -1.94815, -0.25, 0.0, 1.075, -3.7,
                                 7.833333333333, 0.0, 1.0, -3.85, 2.55555556, 0.0, 0.0, -7.7,
                                 0.0, -3.166666667, -8.0, -1.88148, -5.75, 0.0, 1.55, -3.0],
            "labels": [4, 10, 4, 4, 9, 11, 11, 2, 2, 11, 0, 12, 11, 11, 3, 4, 4, 11, 4, 4,
                       3, 11, 11, 2, 2, 11, 4, 4, 11, 11, 15, 4, 4, 7, 4, 4, 15, 11, 11, 14,
                       14, 11, 4, 4, 11, 12, 15, 4, 4, 11, 4, 4, 9, 11, 11, 2, 14, 11, 4, 0,
                       11, 11, 9, 4, 4, 11, 4, 4, 7, 11, 11, 14, 0, 11, 4, 4, 11, 11, 7, 4],
            "X_colmajor": [-3.700001, -7.7], "m_samples": 1,
            "X_transform_colmajor": [16.712636602948745, 8.5428332536694178, 9.8978987669100764,
                                     3.9131189606246322, 11.945852343094563, 8.5428332536694178,
                                     8.5428332536694178, 0.0, 8.5428332536694178,
                                     4.5870590917396212, 2.7166155414412292, 6.0765238962074113,
                                     3.9629534440868719, 8.5428332536694178, 10.409761044327579,
                                     4.7000000000000002],
            "Y_colmajor": [-3.7, -7.7], "k_samples": 1, "Y_labels": [7],
            "tol_check": "10000*sqrt(eps)",
        },
        {
            "name": "empty cluster, ignore",
            "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:141-193",
            "n_samples": 5, "n_features": 2, "n_clusters": 2, "max_iter": 30,
            "A_colmajor": [2.13, 2.11, 2.12, 2.13, 2.14, 2.11, 2.14, 2.13, 2.12, 2.13],
            "C_colmajor": [2.12, 100.0, 2.13, 100.0],
            "tol": 1e-4, "empty_clusters": "ignore",
            "rinfo": [5.0, 2.0, 2.0, 1.0, 0.00104, 0.0],
            "centres_colmajor": [2.126, 0.0, 2.126, 0.0],
            "labels": [0, 0, 0, 0, 0],
            "X_colmajor": [3.3, 3.3], "m_samples": 1,
            "X_transform_colmajor": [1.6602867222260123, 4.6669047558312133],
            "Y_colmajor": [2.3, 2.4], "k_samples": 1, "Y_labels": [0],
            "tol_check": "10*eps",
        },
        {
            "name": "empty cluster, split",
            "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:784-797",
            "n_samples": 5, "n_features": 2, "n_clusters": 2, "max_iter": 30,
            "A_colmajor": [2.13, 2.11, 2.12, 2.13, 2.14, 2.11, 2.14, 2.13, 2.12, 2.13],
            "C_colmajor": [2.12, 100.0, 2.13, 100.0],
            "tol": 1e-4, "empty_clusters": "split",
            "rinfo": [5.0, 2.0, 2.0, 1.0, 0.000475, 0.0],
            "centres_colmajor": [2.13, 2.11, 2.1225, 2.14],
            "labels": [0, 1, 0, 0, 0],
            "X_colmajor": [3.3, 3.3], "m_samples": 1,
            "X_transform_colmajor": [1.6599416405404137, 1.6618363336983584],
            "Y_colmajor": [2.3, 2.4], "k_samples": 1, "Y_labels": [1],
            "tol_check": "10*eps",
        },
        {
            "name": "1 by 1 data matrix, random initialization, lloyd",
            "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:88-139,558-567",
            "init": "random",
            "n_samples": 1, "n_features": 1, "n_clusters": 1, "max_iter": 30,
            "A_colmajor": [2.1], "C_colmajor": [2.3],
            "tol": 1e-4, "empty_clusters": "ignore",
            "rinfo": [1.0, 1.0, 1.0, 0.0, 0.0, 0.0],
            "centres_colmajor": [2.1], "labels": [0],
            "X_colmajor": [3.3], "m_samples": 1, "X_transform_colmajor": [1.2],
            "Y_colmajor": [1.3], "k_samples": 1, "Y_labels": [0],
            "tol_check": "10*eps",
        },
    ],
    "kmeans_error_empty": {
        "source": "tests/unit_tests/kmeans/kmeans_test_data.hpp:799-805",
        "comment": "the empty-cluster case with 'empty clusters'='error' returns da_status_empty_clusters",
    },
}


def _three_blobs(n, d, seed):
    """The reference's 30 x d three-cluster data (dbscan_test_data.hpp:99-131): entry f of
    the column-major array is U[-0.1, 0.1] plus 2, -2 or 0 by f % 3, so sample i sits near
    (c, ..., c) with c chosen by i % 3. The draw is seeded here (the reference seeds from
    std::random_device); the expected labels hold for every draw."""
    r = np.random.default_rng(seed).uniform(-0.1, 0.1, n * d)
    f = np.arange(n * d)
    return (r + np.where(f % 3 == 0, 2.0, np.where(f % 3 == 1, -2.0, 0.0))).tolist()


_A25 = [0, 1.0, 1.0, -1.0, -1.0, 10.0, -7.0, -6.0, -8.0, -6.5, -5.0, -8.0, -6.5, -5.0, -12.0,
        -12.1, -11.9, 15.0, -5.0, -5.1, -4.9, -5.0, -5.1, -5.1, -5.0, 0, 1.0, -1.0, 1.0, -1.0,
        10.0, 2.0, 2.0, 3.0, 3.1, 3.0, 1.0, 1.0, 1.0, -12.0, -12.1, -11.9, 0.0, -5.0, -5.1,
        -4.9, -5.0, -5.1, -5.1, -5.0]
_L25 = [0, 0, 0, 0, 0, -1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 2, 2, 2, 2, 2, 2, 2]
_C25 = [0, 6, 7, 12, 18, 19, 20, 21, 22, 23, 24]
_ABC = [0, 1, 2] * 10


def _db(name, src, n, d, A, lda, order, min_samples, eps, labels, core, rinfo, n_clusters,
        algorithm="brute", metric="euclidean", power=2.0, leaf_size=30):
    return {"name": name, "source": "tests/unit_tests/dbscan/dbscan_test_data.hpp:" + src,
            "n_samples": n, "n_features": d, "A": A, "lda": lda, "order": order,
            "min_samples": min_samples, "eps": eps, "algorithm": algorithm, "metric": metric,
            "power": power, "leaf_size": leaf_size, "labels": labels, "core": core,
            "rinfo": rinfo, "n_clusters": n_clusters, "n_core": len(core)}


_a25_row = np.array(_A25).reshape(2, 25).T.ravel().tolist()
_a25_sub = _A25[:25] + [0, 0] + _A25[25:] + [0, 0]
known["dbscan"] = [
    _db("1 by 1", "40-55,260-268", 1, 1, [2.1], 1, "column-major", 1, 0.5, [0], [0],
        [1, 1, 1, 0.5, 1, 30, 2, 1, 1], 1),
    _db("2 by 25, k-d tree", "57-78,280-288", 2, 25, [float(i % 2) for i in range(50)], 2,
        "column-major", 1, 100.0, [0, 0], [0, 1], [2, 25, 2, 100, 1, 1, 2, 2, 1], 1,
        algorithm="kd tree", leaf_size=1),
    _db("2 by 25, auto", "57-78,280-288", 2, 25, [float(i % 2) for i in range(50)], 2,
        "column-major", 1, 100.0, [0, 0], [0, 1], [2, 25, 2, 100, 1, 1, 2, 2, 1], 1,
        algorithm="auto", leaf_size=1),
    _db("zeros", "80-97,270-278", 25, 5, [0.0] * 125, 25, "column-major", 2, 0.5, [0] * 25,
        list(range(25)), [25, 5, 25, 0.5, 2, 30, 2, 25, 1], 1),
    _db("25 by 2", "166-190,316-321", 25, 2, _A25, 25, "column-major", 4, 1.5, _L25, _C25,
        [25, 2, 25, 1.5, 4, 30, 2, 11, 3], 3),
    _db("25 by 2 row-major", "248-258", 25, 2, _a25_row, 2, "row-major", 4, 1.5, _L25, _C25,
        [25, 2, 2, 1.5, 4, 30, 2, 11, 3], 3),
    _db("25 by 2 tiny eps", "322-334", 25, 2, _A25, 25, "column-major", 4, 0.0001, [-1] * 25,
        [], [25, 2, 25, 0.0001, 4, 30, 2, 0, 0], 0),
    _db("25 by 2 subarray", "335-348", 25, 2, _a25_sub, 27, "column-major", 4, 1.5, _L25, _C25,
        [25, 2, 27, 1.5, 4, 30, 2, 11, 3], 3),
]
for alg in ("brute", "kd tree", "auto", "ball tree"):
    known["dbscan"].append(
        _db(f"30 by 3, {alg}", "99-131,290-306", 30, 3, _three_blobs(30, 3, 11), 30,
            "column-major", 3, 0.5, _ABC, list(range(30)), [30, 3, 30, 0.5, 3, 30, 2, 30, 3], 3,
            algorithm=alg))
for metric, alg, power, eps in (("euclidean", "brute", 2.0, 0.5), ("manhattan", "brute", 2.0, 0.5),
                                ("minkowski", "brute", 2.0, 0.5),
                                ("minkowski", "brute", 2.0000001, 0.5),
                                ("euclidean", "kd tree", 2.0000001, 0.5),
                                ("euclidean", "ball tree", 2.0000001, 0.5),
                                ("sqeuclidean", "brute", 2.0000001, 0.25)):
    known["dbscan"].append(
        _db(f"30 by 1, {metric}, {alg}, p={power}", "133-164,207-246", 30, 1,
            _three_blobs(30, 1, 12), 30, "column-major", 3, eps, _ABC, list(range(30)),
            [30, 1, 30, eps, 3, 30, power, 30, 3], 3, algorithm=alg, metric=metric, power=power))
# k-NN / radius classification and regression of the 6 x 3 example
# (nearest_neighbors_tests.hpp:164-427, data :446-466): proba column-major n_queries x
# n_classes; "euclidean" stands for euclidean, l2, minkowski and euclidean_gemm,
# "sqeuclidean" for sqeuclidean and sqeuclidean_gemm; radius searches use the "radius"
# option 21.
known["knn_predict"] = {
    "source": "tests/unit_tests/nearest_neighbors/nearest_neighbors_tests.hpp:164-466",
    "X_train_colmajor": [-1., -2., -3., 1., 2., 3., -1., -1., -2., 3., 5., -1., 2., 3., -1., 1.,
                         1., 2.],
    "X_test_colmajor": [-2., -1., 2., 2., -2., 1., 3., -1., -3.],
    "y_class": [1, 2, 0, 1, 2, 2], "y_reg": [1.5, 2.2, 0.5, 1, 2.8, 5.2], "radius": 21.0,
    "proba_knn": {
        "5": {"uniform": {"euclidean": [0.2, 0.2, 0.2, 0.4, 0.4, 0.4, 0.4, 0.4, 0.4],
                          "sqeuclidean": [0.2, 0.2, 0.2, 0.4, 0.4, 0.4, 0.4, 0.4, 0.4]},
              "distance": {
                  "euclidean": [0.1379511568268668, 0.3515868265794006, 0.1798440493222374,
                                0.4507346784224799, 0.3447698547319956, 0.4217676420329797,
                                0.4113141647506533, 0.3036433186886039, 0.3983883086447829],
                  "sqeuclidean": [0.0895917616770872, 0.5270701941190925, 0.1596502898413,
                                  0.4799558661272527, 0.2747153739044967, 0.4485412905065095,
                                  0.4304523721956602, 0.1982144319764108, 0.3918084196521904]}},
        "3": {"uniform": {"euclidean": [0., 1 / 3, 0., 2 / 3, 1 / 3, 1 / 3, 1 / 3, 1 / 3, 2 / 3],
                          "sqeuclidean": [0., 1 / 3, 0., 2 / 3, 1 / 3, 1 / 3, 1 / 3, 1 / 3,
                                          2 / 3]},
              "distance": {
                  "euclidean": [0., 0.47531678671182, 0., 0.6304942401901729, 0.3006167312243614,
                                0.3778214838715254, 0.3695057598098272, 0.2240664820638185,
                                0.6221785161284746],
                  "sqeuclidean": [0., 0.6164383561643836, 0., 0.5936675461741425,
                                  0.2465753424657534, 0.4244031830238727, 0.4063324538258575,
                                  0.136986301369863, 0.5755968169761273]}},
    },
    "labels_knn": {"5": {"uniform": [1, 1, 1], "distance": [1, 0, 1]},
                   "3": {"uniform": [1, 0, 2], "distance": [1, 0, 2]}},
    "targets_knn": {
        "5": {"uniform": {"euclidean": [1.6, 2.08, 2.2], "sqeuclidean": [1.6, 2.08, 2.2]},
              "distance": {"euclidean": [1.6323653119203947, 1.7134711834322838,
                                         2.202877674352422],
                           "sqeuclidean": [1.6673041559396837, 1.3229994266524698,
                                           2.190592036491495]}},
        "3": {"uniform": {"euclidean": [1.5666666666666667, 1.4000000000000001, 3.0],
                          "sqeuclidean": [1.5666666666666667, 1.4000000000000001, 3.0]},
              "distance": {"euclidean": [1.6105221794760263, 1.1815297507328528,
                                         2.8785808475618806],
                           "sqeuclidean": [1.6538258575197888, 0.9794520547945206,
                                           2.749071618037135]}},
    },
    "proba_radius": {
        "uniform": {"euclidean": [1 / 6, 1 / 6, 1 / 6, 1 / 3, 1 / 3, 1 / 3, 0.5, 0.5, 0.5],
                    "sqeuclidean": [0.0, 1 / 3, 0.0, 2 / 3, 1 / 3, 1.0, 1 / 3, 1 / 3, 0.0]},
        "distance": {
            "euclidean": [0.12165522957880101, 0.32276306991001125, 0.1566384758301224,
                          0.3974901845254738, 0.31650496637875825, 0.3673462694566996,
                          0.4808545858957253, 0.3607319637112305, 0.47601525471317796],
            "sqeuclidean": [0.0, 0.6164383561643836, 0.0, 0.5936675461741425,
                            0.24657534246575344, 1.0, 0.40633245382585753, 0.136986301369863,
                            0.0]},
    },
    "labels_radius": {"euclidean": [2, 2, 2], "sqeuclidean": [1, 0, 1]},
    "targets_radius": {
        "uniform": {"euclidean": [2.1999999999999997, 2.1999999999999997, 2.1999999999999997],
                    "sqeuclidean": [1.5666666666666667, 1.4000000000000001, 1.0]},
        "distance": {"euclidean": [2.053803715641692, 1.802546882108823, 2.2025063632975206],
                     "sqeuclidean": [1.6538258575197888, 0.9794520547945206, 1.0]},
    },
}
# IVF-Flat cases with fixed answers (ann_tests.hpp:393-575): squared euclidean, all-zero
# training data (n_list 2, n_probe 1, k 2: distances only, the indices tie) and the 1 x 1
# case (n_list 1, n_probe 1, k 1), in column- and row-major order.
known["ivf"] = [
    {"name": "col zero", "source": "tests/unit_tests/approximate_neighbors/ann_tests.hpp:393-412",
     "order": "column-major", "n_list": 2, "n_probe": 1, "k": 2, "n_samples": 8, "n_features": 4,
     "n_queries": 2, "X_train": [0.0] * 32, "X_test": [1., 2., 3., 0.5, 1.5, 2.5, -0.5, -1],
     "kdist": [12.5, 11.5, 12.5, 11.5], "kind": None},
    {"name": "row zero", "source": "tests/unit_tests/approximate_neighbors/ann_tests.hpp:413-431",
     "order": "row-major", "n_list": 2, "n_probe": 1, "k": 2, "n_samples": 8, "n_features": 4,
     "n_queries": 2, "X_train": [0.0] * 32, "X_test": [1., 3., 1.5, -0.5, 2., 0.5, 2.5, -1],
     "kdist": [12.5, 12.5, 11.5, 11.5], "kind": None},
    {"name": "col 1x1", "source": "tests/unit_tests/approximate_neighbors/ann_tests.hpp:499-521",
     "order": "column-major", "n_list": 1, "n_probe": 1, "k": 1, "n_samples": 1, "n_features": 1,
     "n_queries": 3, "X_train": [0.5], "X_test": [-1., 0., 1.], "kdist": [2.25, 0.25, 0.25],
     "kind": [0, 0, 0]},
    {"name": "row 1x1", "source": "tests/unit_tests/approximate_neighbors/ann_tests.hpp:522-547",
     "order": "row-major", "n_list": 1, "n_probe": 1, "k": 1, "n_samples": 1, "n_features": 1,
     "n_queries": 3, "X_train": [0.5], "X_test": [-1., 0., 1.], "kdist": [2.25, 0.25, 0.25],
     "kind": [0, 0, 0]},
]
known["dbscan_errors"] = {
    "source": "tests/unit_tests/dbscan/dbscan_public.cpp:309-395",
    "A_colmajor": [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12], "n_samples": 4, "n_features": 3,
    "lda": 4,
}
# k-means cases of the reference whose answers are bounds or iteration counts
# (kmeans_test_data.hpp). "base" names the entry of known["kmeans"] whose data, initial
# centres (for "supplied") and options are reused; the fields given here override them.
# Random initialisations are checked as the reference checks them (kmeans_public.cpp:272-276):
# inertia <= max_inertia + check_tol with check_tol = 100*sqrt(eps) of the test's type.
_3c = "three clusters, lloyd"
known["kmeans_random"] = [
    dict(base=_3c, source="kmeans_test_data.hpp:604-624", init=i, algorithm="elkan",
         n_init=10, max_iter=30, seed=78, n_clusters=3, max_inertia=0.185475)
    for i in ("k-means++", "random", "random partitions", "afk-mc2")
] + [
    dict(base=_3c, source="kmeans_test_data.hpp:625-636", init="k-means++", algorithm=a,
         n_init=1, max_iter=30, seed=78, n_clusters=1, max_inertia=34.89176)
    for a in ("elkan", "lloyd")
] + [
    dict(base=_3c, source="kmeans_test_data.hpp:643-649", init="k-means++", algorithm="lloyd",
         n_init=1, max_iter=30, seed=78, n_clusters=10, max_inertia=0.11),
    dict(base=_3c, source="kmeans_test_data.hpp:655-663", init="random partitions",
         algorithm="elkan", n_init=5, max_iter=30, seed=78, n_clusters=10, max_inertia=0.11),
    dict(base=_3c, source="kmeans_test_data.hpp:664-672", init="k-means++", algorithm="lloyd",
         n_init=10, max_iter=30, seed=78, n_clusters=2, max_inertia=15.7915238095238),
    dict(base=_3c, source="kmeans_test_data.hpp:673-676", init="afk-mc2", algorithm="lloyd",
         n_init=10, max_iter=30, seed=78, n_clusters=2, max_inertia=15.7915238095238),
    dict(base=_3c, source="kmeans_test_data.hpp:677-680", init="random", algorithm="elkan",
         n_init=10, max_iter=30, seed=78, n_clusters=2, max_inertia=15.7915238095238),
    dict(base=_3c, source="kmeans_test_data.hpp:690-695", init="k-means++", algorithm="lloyd",
         n_init=10, max_iter=30, seed=78, n_clusters=5, max_inertia=0.076),
    dict(base=_3c, source="kmeans_test_data.hpp:696-701", init="random", algorithm="elkan",
         n_init=10, max_iter=30, seed=78, n_clusters=5, max_inertia=0.11),
    dict(base="three clusters, row-major", source="kmeans_test_data.hpp:923-942",
         init="k-means++", algorithm="elkan", n_init=1, max_iter=50, seed=78, n_clusters=3,
         max_inertia=0.185475),
    dict(base="three clusters, row-major", source="kmeans_test_data.hpp:929-933",
         init="afk-mc2", algorithm="elkan", n_init=1, max_iter=50, seed=78, n_clusters=3,
         max_inertia=0.185475),
    dict(base="three clusters, row-major", source="kmeans_test_data.hpp:934-942",
         init="random", algorithm="elkan", n_init=1, max_iter=50, seed=12345, n_clusters=3,
         max_inertia=0.185475),
    dict(base="three clusters, row-major", source="kmeans_test_data.hpp:939-942",
         init="random partitions", algorithm="elkan", n_init=1, max_iter=50, seed=12345,
         n_clusters=3, max_inertia=0.185475),
    dict(base="pseudorandom 20 x 2", source="kmeans_test_data.hpp:530-556,746-755",
         init="random", algorithm="elkan", n_init=1, max_iter=30, seed=593228, n_clusters=3,
         max_inertia=4.8, check_tol="1000*sqrt(eps)"),
]
# The row-major 3-cluster data (kmeans_test_data.hpp:310-370) is the column-major data of
# known["kmeans"][0] transposed; its expected answers are the same.
known["kmeans_extra_data"] = {
    "three clusters, row-major": {
        "source": "kmeans_test_data.hpp:310-370", "order": "row-major",
        "A_rowmajor": [1.0, 1.0, 1.1, 1.2, 0.5, -2.0, 0.49, -1.9, -2.0, 0.5,
                       -2.0, 0.51, 0.53, -2.1, 0.9, 0.95, 1.2, 0.8, -1.8, 0.6],
        "C_rowmajor": [0.5, 0.5, 0.7, -1.7, -1.3, 0.2],
        "n_samples": 10, "n_features": 2},
    "pseudorandom 20 x 2": {
        "source": "kmeans_test_data.hpp:530-556", "order": "column-major",
        "A_colmajor": [0.31, 0.61, 0.65, -0.49, -0.7, -0.35, 0.53, 0.29, 0.23, -0.58, 1.0,
                       0.23, -0.04, -0.79, 0.25, 0.51, -0.41, 0.95, -0.81, -0.61, -0.41, 0.12,
                       0.75, 0.49, -0.47, -0.85, -0.53, 0.19, 0.25, -0.82, 0.52, -0.26, -0.01,
                       -0.49, 0.56, 0.51, -0.61, 0.95, 0.83, -0.76],
        "n_samples": 20, "n_features": 2},
}
# Deterministic cases with the Elkan algorithm: the reference expects the same answers as
# the base entry (kmeans_test_data.hpp:600-603 three clusters, :757-766 the larger matrix,
# :785-797 empty cluster split, :918-922 row-major), i.e. rinfo n_iter 1, 4, 1 and 1.
known["kmeans_elkan"] = [
    {"base": "three clusters, lloyd", "source": "kmeans_test_data.hpp:600-603"},
    {"base": "larger data matrix, lloyd", "source": "kmeans_test_data.hpp:757-766"},
    {"base": "empty cluster, split", "source": "kmeans_test_data.hpp:785-797"},
]
# Cases with every answer fixed (is_random false): all-zero 5 x 3 data, k = 2, max_iter 3
# (kmeans_test_data.hpp:195-248,718-744), expected rinfo {5, 3, 2, 2, 0, 0}, centres and
# transform zero, labels 0; and the 1 x 1 matrix (:88-139,558-584), rinfo {1, 1, 1, 0, 0, 0}.
known["kmeans_fixed_random_init"] = [
    dict(name="zero data", source="kmeans_test_data.hpp:195-248,718-728", init="k-means++",
         algorithm="elkan", n_init=3, seed=-1, n_samples=5, n_features=3, n_clusters=2,
         max_iter=3, tol=1e-4, A_colmajor=[0.0] * 15, rinfo=[5.0, 3.0, 2.0, 2.0, 0.0, 0.0],
         centres_colmajor=[0.0] * 6, labels=[0] * 5, X_colmajor=[0.0] * 9, m_samples=3,
         X_transform_colmajor=[0.0] * 6, Y_colmajor=[0.0] * 6, k_samples=2, Y_labels=[0, 0]),
] + [
    dict(name="zero data " + i, source="kmeans_test_data.hpp:729-737", init=i,
         algorithm="lloyd", n_init=3, seed=-1, n_samples=5, n_features=3, n_clusters=2,
         max_iter=3, tol=1e-4, A_colmajor=[0.0] * 15, rinfo=[5.0, 3.0, 2.0, 2.0, 0.0, 0.0],
         centres_colmajor=[0.0] * 6, labels=[0] * 5, X_colmajor=[0.0] * 9, m_samples=3,
         X_transform_colmajor=[0.0] * 6, Y_colmajor=[0.0] * 6, k_samples=2, Y_labels=[0, 0])
    for i in ("random", "afk-mc2")
] + [
    dict(name="1 by 1 " + i + " " + a, source="kmeans_test_data.hpp:88-139,558-584", init=i,
         algorithm=a, n_init=1, seed=78, n_samples=1, n_features=1, n_clusters=1,
         max_iter=30, tol=1e-4, A_colmajor=[2.1], rinfo=[1.0, 1.0, 1.0, 0.0, 0.0, 0.0],
         centres_colmajor=[2.1], labels=[0], X_colmajor=[3.3], m_samples=1,
         X_transform_colmajor=[1.2], Y_colmajor=[1.3], k_samples=1, Y_labels=[0])
    for i, a in (("k-means++", "elkan"), ("random", "elkan"), ("random", "lloyd"),
                 ("random partitions", "lloyd"))
]
# Mixed-precision iterative refinement (kmeans_test_data.hpp:944-969): rinfo[3] (iterations
# in the working precision) and rinfo[5] (iterations of the lower-precision phase).
known["kmeans_mixed"] = [
    dict(base="three clusters, lloyd", source="kmeans_test_data.hpp:947-957", lp_max_iter=1,
         lp_tol=1e-2, n_iter={"double": 0, "float": 1}, lp_n_iter=1),
    dict(base="larger data matrix, lloyd", source="kmeans_test_data.hpp:958-968",
         lp_max_iter=3, lp_tol=1e-3, n_iter={"double": 1, "float": 1}, lp_n_iter=3),
]


# Every IVF-Flat case of ann_tests.hpp:395-1850 with fixed expectations, read from the test
# source as data: the constructor (n_list, n_probe, k, metric, algorithm, order), the scalar
# fields it sets (sizes, leading dimensions, seed, k-means iterations, train fraction) and the
# vectors X_train, X_test, expected_kind and expected_kdist. "...Euclidean" cases are the
# squared-euclidean ones with the metric changed and the distances square-rooted
# (ConvertToEuclidean, :141-153).
def _ann_cases():
    import math
    import re
    src = open(os.path.join(REF, "tests", "unit_tests", "approximate_neighbors",
                            "ann_tests.hpp")).read()
    lines = src.split("\n")
    heads = [(m.start(), m.group(1)) for m in
             re.finditer(r"void (\w+)\(std::vector<ANNParamType<T>> &params\) \{", src)]
    bodies = {}
    for i, (pos, name) in enumerate(heads):
        end = heads[i + 1][0] if i + 1 < len(heads) else len(src)
        bodies[name] = (src.count("\n", 0, pos) + 1, src[pos:end])

    def num(t):
        return float(t.strip().rstrip("f"))

    def parse(name):
        line0, b = bodies[name]
        m = re.search(r'ANNParamType<T> test\((\d+), (\d+), (\d+), "([^"]+)", "([^"]+)", '
                      r'"([^"]+)"\)', b)
        c = {"name": name, "source": "tests/unit_tests/approximate_neighbors/ann_tests.hpp:%d"
             % line0, "n_list": int(m.group(1)), "n_probe": int(m.group(2)),
             "k": int(m.group(3)), "metric": m.group(4), "algorithm": m.group(5),
             "order": m.group(6), "seed": 0,
             "kmeans_iter": 10, "train_fraction": 1.0}
        for f, v in re.findall(r"test\.(n_features|n_samples|n_queries|ldx_train|ldx_test|"
                               r"seed|kmeans_iter|train_fraction) = ([^;]+);", b):
            v = v.strip()
            mm = re.fullmatch(r"test\.(\w+)", v)
            c[f] = c[mm.group(1)] if mm else (float(v) if "." in v else int(v))
        for f, typ in (("X_train", "T"), ("X_test", "T"), ("expected_kind", "da_int"),
                       ("expected_kdist", "T")):
            mm = re.search(r"std::vector<%s> %s\{([^}]*)\}" % (typ, f), b, re.S)
            c[f] = None if mm is None else [num(t) for t in mm.group(1).split(",")
                                             if t.strip()]
        if c["expected_kind"] is not None:
            c["expected_kind"] = [int(v) for v in c["expected_kind"]]
        return c

    out = []
    for name in bodies:
        b = bodies[name][1]
        if "ANNParamType<T> test(" in b and "csvname" not in b:
            c = parse(name)
            if c["X_train"] is None or c["expected_kdist"] is None:
                continue
            out.append(c)
        conv = re.search(r"(\w+SqEuclidean)\(params\);\s*ConvertToEuclidean", b)
        if conv:
            c = parse(conv.group(1))
            c["name"] = name
            c["metric"] = "euclidean"
            c["expected_kdist"] = [math.sqrt(v) for v in c["expected_kdist"]]
            out.append(c)
    return out


known["ivf_cases"] = _ann_cases()
with open(os.path.join(HERE, "known_answers.json"), "w") as f:
    json.dump(known, f, indent=1)
print("fixtures written to", HERE)
