"""Resolves the k-means fixtures of tests/golden/known_answers.json into data and options.

The reference's cases (tests/unit_tests/kmeans/kmeans_test_data.hpp) store matrices in the
order the case names; every case here is returned with row-major numpy arrays and the
storage order of the original, so that a test can pass either layout."""
import numpy as np

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def colmajor(v, rows, cols):
    return np.array(v, dtype=np.float64).reshape((cols, rows)).T


def base_data(known, name):
    """(A, C0, order, base entry) of a named case: the entries of known["kmeans"] and of
    known["kmeans_extra_data"]."""
    for c in known["kmeans"]:
        if c["name"] == name:
            n, d, k = c["n_samples"], c["n_features"], c["n_clusters"]
            return (colmajor(c["A_colmajor"], n, d), colmajor(c["C_colmajor"], k, d),
                    "column-major", c)
    c = known["kmeans_extra_data"][name]
    n, d = c["n_samples"], c["n_features"]
    if "A_rowmajor" in c:
        A = np.array(c["A_rowmajor"], dtype=np.float64).reshape(n, d)
        C0 = np.array(c["C_rowmajor"], dtype=np.float64).reshape(-1, d)
        return A, C0, "row-major", c
    return colmajor(c["A_colmajor"], n, d), None, "column-major", c


def check_tol(name, dtype):
    eps = EPS64 if dtype == np.float64 else EPS32
    return {"100*sqrt(eps)": 100 * np.sqrt(eps), "1000*sqrt(eps)": 1000 * np.sqrt(eps),
            "10000*sqrt(eps)": 1e4 * np.sqrt(eps), "10*eps": 10 * eps}[name]
