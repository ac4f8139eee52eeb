"""k-nearest neighbours and IVF-Flat approximate neighbours
(python_interface/python_package/aoclda/neighbors.py:39-780 of the reference): brute force
and probed inverted lists on the GPU, with the indexed vectors resident in GPU memory."""
import ctypes as C

import numpy as np

from . import _lib as L


class nearest_neighbors:
    def __init__(self, n_neighbors=5, weights="uniform", algorithm="auto", leaf_size=30,
                 metric="euclidean", p=2.0, radius=1.0, outlier_handling=None,
                 check_data=False, precision="working", tile_variant=-1,
                 candidates_per_split=0, seed_thresholds=None):
        # precision: operand precision of the MFMA distance tiles ("working", "fp32",
        # "fp16", "bf16", "fp16x3", "bf16x3", "fp16c": fp16 rows stored about k-means group
        # centres); results are re-ranked exactly whatever the
        # choice. candidates_per_split: candidate-list length per query and split of
        # the fused tile (0 = from k; shorter than k is allowed and stays exact).
        # seed_thresholds: bound the tile's thresholds (k <= 16) by the group
        # minima over a row sample first (None: the library's choice, which seeds the
        # 16-bit forms). All three are MI355X options that change speed only.
        if not isinstance(n_neighbors, int) and n_neighbors is not None:
            raise ValueError("Number of neighbors must be an integer or None. Got n_neighbors="
                             + str(n_neighbors))
        self.n_neighbors = n_neighbors
        self.weights, self.algorithm, self.leaf_size = weights, algorithm, leaf_size
        self.metric, self.p, self.radius = metric, p, radius
        self.outlier_handling = outlier_handling
        self.y = None
        self.check_data = check_data
        self.precision = precision
        self.tile_variant = tile_variant
        self.candidates_per_split = candidates_per_split
        self.seed_thresholds = seed_thresholds
        self._h = None
        self.order = None
        self.dtype = None

    def _options(self, h):
        h.set_int("number of neighbors", self.n_neighbors if self.n_neighbors else 1)
        h.set_string("weights", self.weights)
        h.set_string("algorithm", self.algorithm)
        h.set_int("leaf size", self.leaf_size)
        h.set_string("metric", "inner product" if self.metric == "inner_product" else self.metric)
        h.set_real("minkowski parameter", self.p)
        h.set_string("storage order", "column-major" if self.order == L.COLUMN_MAJOR else "row-major")
        h.set_string("check data", "yes" if self.check_data else "no")
        h.set_string("cross term precision", self.precision)
        h.set_int("tile variant", self.tile_variant)
        h.set_int("candidates per split", self.candidates_per_split)
        if self.seed_thresholds is not None:
            h.set_string("seed thresholds", "yes" if self.seed_thresholds else "no")
        if self.radius is not None:
            h.set_real("radius", self.radius)

    def fit(self, X, y=None):
        X, self.order, self.dtype = L.prepare(X)
        self._h = L.Handle(L.HANDLE_NN, self.dtype)
        self._options(self._h)
        n, d = X.shape
        ld = n if self.order == L.COLUMN_MAJOR else d
        self._h.call("da_nn_set_data", n, d, L.ptr(X), ld)
        self.n_samples, self.n_features = n, d
        self.y = None if y is None else np.asarray(y)
        return self

    def kneighbors(self, X, n_neighbors=None, return_distance=True):
        X, _, _ = L.prepare(X, dtype=self.dtype, order=self.order)
        nq, d = X.shape
        k = n_neighbors if n_neighbors is not None else (self.n_neighbors or 0)
        kk = k if k > 0 else 5
        ind = np.empty((nq, kk), dtype=np.int32, order="F" if self.order == L.COLUMN_MAJOR else "C")
        dist = np.empty((nq, kk), dtype=self.dtype, order=ind.flags.f_contiguous and "F" or "C")
        ld = nq if self.order == L.COLUMN_MAJOR else d
        self._h.call("da_nn_kneighbors", nq, d, L.ptr(X), ld, L.ptr(ind),
                     L.ptr(dist) if return_distance else None, k, 1 if return_distance else 0)
        return (dist, ind) if return_distance else ind

    def radius_neighbors(self, X, radius=None, return_distance=True, sort_results=False):
        """Training rows within `radius` of each query (neighbors.py:232-330 of the
        reference): object arrays of per-query index (and distance) arrays, in ascending
        training index, or by (distance, index) with sort_results."""
        if not return_distance and sort_results:
            raise ValueError("Cannot sort results if distances are not returned.")
        if X is None:
            raise ValueError("radius_neighbors() is not implemented for X=None")
        if radius is not None and not isinstance(radius, (int, float)):
            raise ValueError("Radius must be a number or None. Got radius=" + str(radius))
        if radius is None and self.radius is None:
            raise ValueError("radius cannot be None in radius_neighbors() if it was None in "
                             "the constructor.")
        if radius is not None and radius < 0.0:
            raise ValueError("Radius must be non-negative or None. Got radius=" + str(radius))
        X, _, _ = L.prepare(X, dtype=self.dtype, order=self.order)
        nq, d = X.shape
        ld = nq if self.order == L.COLUMN_MAJOR else d
        self._h.call("da_nn_radius_neighbors", nq, d, L.ptr(X), ld,
                     -1.0 if radius is None else float(radius), 1 if return_distance else 0,
                     1 if sort_results else 0)
        counts = self._int_result(L.RNN_COUNT, nq + 1)
        total = int(counts[nq])
        flat_ind = self._int_result(L.RNN_INDICES, max(total, 1))[:total]
        bounds = np.concatenate(([0], np.cumsum(counts[:nq])))
        inds = np.empty(nq, dtype=object)
        for q in range(nq):
            inds[q] = flat_ind[bounds[q]:bounds[q + 1]].copy()
        if not return_distance:
            return inds
        flat_d = np.zeros(max(total, 1), dtype=self.dtype)
        n = L.da_int(flat_d.size)
        L.check(getattr(L.lib, f"da_handle_get_result_{L.suffix(self.dtype)}")(
            self._h.h, L.RNN_DISTANCES, C.byref(n), L.ptr(flat_d)), self._h.h)
        dists = np.empty(nq, dtype=object)
        for q in range(nq):
            dists[q] = flat_d[bounds[q]:bounds[q + 1]].copy()
        return dists, inds

    _MODES = {"knn": 0, "radius_neighbors": 1}

    def _predict_setup(self, search_mode, classification):
        """Labels or targets and the outlier options of a prediction (neighbors.py:335-535 of
        the reference)."""
        if search_mode not in self._MODES:
            raise ValueError("search_mode must be 'knn' or 'radius_neighbors'")
        if self.y is None:
            raise ValueError("y was not provided during fit().")
        if search_mode == "radius_neighbors" and self.radius is None:
            raise ValueError("radius was None in the constructor.")
        h = self._h
        if classification:
            y = np.ascontiguousarray(self.y, dtype=np.int32)
            h.call("da_nn_set_labels", len(y), L.ptr(y))
        else:
            y = np.ascontiguousarray(self.y, dtype=self.dtype)
            h.call("da_nn_set_targets", len(y), L.ptr(y))
        self._y_dev = y
        mode = "none"
        if search_mode == "radius_neighbors" and self.outlier_handling is not None:
            if self.outlier_handling == "most_frequent":
                mode = "most frequent"
            else:
                mode = "manual"
                if classification:
                    if not isinstance(self.outlier_handling, (int, np.integer)):
                        raise ValueError("outlier_handling must be None, 'most_frequent', or "
                                         "an integer for classification.")
                    h.set_int("outlier label", int(self.outlier_handling))
                else:
                    h.set_real("outlier target", float(self.outlier_handling))
        h.set_string("outlier handling", mode)
        return self._MODES[search_mode]

    def _queries(self, X):
        X, _, _ = L.prepare(X, dtype=self.dtype, order=self.order)
        nq, d = X.shape
        return X, nq, d, (nq if self.order == L.COLUMN_MAJOR else d)

    def classifier_predict_proba(self, X, search_mode="knn"):
        """Class probabilities (n_queries x n_classes, classes ascending) from the k nearest
        neighbours or the neighbours within the radius."""
        m = self._predict_setup(search_mode, True)
        X, nq, d, ld = self._queries(X)
        n = L.da_int(0)
        self._h.call("da_nn_classes", C.byref(n), None)
        proba = np.empty((nq, n.value), dtype=self.dtype,
                         order="F" if self.order == L.COLUMN_MAJOR else "C")
        self.status = self._h.call("da_nn_classifier_predict_proba", nq, d, L.ptr(X), ld,
                                   L.ptr(proba), m, allow=(L.CODE["outlier_warning"],))
        return proba

    def classifier_predict(self, X, search_mode="knn"):
        m = self._predict_setup(search_mode, True)
        X, nq, d, ld = self._queries(X)
        out = np.empty(nq, dtype=np.int32)
        self.status = self._h.call("da_nn_classifier_predict", nq, d, L.ptr(X), ld, L.ptr(out),
                                   m, allow=(L.CODE["outlier_warning"],))
        return out

    def regressor_predict(self, X, search_mode="knn"):
        m = self._predict_setup(search_mode, False)
        X, nq, d, ld = self._queries(X)
        out = np.empty(nq, dtype=self.dtype)
        self.status = self._h.call("da_nn_regressor_predict", nq, d, L.ptr(X), ld, L.ptr(out), m,
                                   allow=(L.CODE["outlier_warning"],))
        return out

    @property
    def classes(self):
        n = L.da_int(0)
        self._h.call("da_nn_classes", C.byref(n), None)
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        self._h.call("da_nn_classes", C.byref(n), L.ptr(out))
        return out[:n.value]

    def _int_result(self, query, size):
        out = np.zeros(size, dtype=np.int32)
        n = L.da_int(size)
        L.check(L.lib.da_handle_get_result_int(self._h.h, query, C.byref(n), L.ptr(out)),
                self._h.h)
        return out

    @property
    def handle(self):
        return self._h


class nn_server:
    """Serving loop over a fitted nearest_neighbors object (MI355X extension,
    da_nn_server_*): submit() queues a request and returns its id without blocking,
    wait() returns (distances, indices, latency_ms)."""

    def __init__(self, nn, max_batch_queries=8192, max_device_bytes=1 << 30):
        self.nn = nn
        self.dtype = nn.dtype
        self.h = C.c_void_p()
        L.check(L.lib.da_nn_server_create(nn.handle.h, max_batch_queries, max_device_bytes,
                                          C.byref(self.h)))
        self._k = {}

    def submit(self, X, k):
        X = np.ascontiguousarray(X, dtype=self.dtype)
        rid = L.da_int()
        f = getattr(L.lib, f"da_nn_server_submit_{L.suffix(self.dtype)}")
        L.check(f(self.h, X.shape[0], X.shape[1], L.ptr(X), X.shape[1], k, C.byref(rid)))
        self._k[rid.value] = (X.shape[0], k)
        return rid.value

    def wait(self, rid):
        nq, k = self._k.pop(rid)
        ind = np.empty((nq, k), dtype=np.int32)
        dist = np.empty((nq, k), dtype=self.dtype)
        lat = C.c_double()
        f = getattr(L.lib, f"da_nn_server_wait_{L.suffix(self.dtype)}")
        L.check(f(self.h, rid, L.ptr(ind), L.ptr(dist), C.byref(lat)))
        return dist, ind, lat.value

    def stats(self):
        out = np.zeros(8)
        n = L.da_int(8)
        L.check(L.lib.da_nn_server_stats(self.h, L.ptr(out), C.byref(n)))
        keys = ("requests", "queries", "batches", "p50_ms", "p99_ms", "queries_per_batch",
                "elapsed_s", "fallback_queries")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if self.h:
            L.lib.da_nn_server_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class approximate_neighbors:
    """IVF-Flat approximate nearest neighbours (neighbors.py:578-780 of the reference):
    k-means centroids, inverted lists and probed-list search on the GPU."""

    def __init__(self, n_neighbors=5, algorithm="ivfflat", metric="sqeuclidean", n_list=1,
                 n_probe=1, kmeans_iter=10, train_fraction=1.0, seed=0, check_data=False):
        self.n_neighbors, self.algorithm, self.metric = n_neighbors, algorithm, metric
        self.n_list, self._n_probe, self.kmeans_iter = n_list, n_probe, kmeans_iter
        self.train_fraction, self.seed, self.check_data = train_fraction, seed, check_data
        self._h = None
        self.order = None
        self.dtype = None

    @property
    def n_probe(self):
        return self._n_probe

    @n_probe.setter
    def n_probe(self, value):
        self._n_probe = value
        if self._h is not None:
            self._h.set_int("n_probe", value)

    def _prepare(self, X):
        X, self.order, self.dtype = L.prepare(X)
        h = L.Handle(L.HANDLE_APPROX_NN, self.dtype)
        h.set_int("number of neighbors", self.n_neighbors)
        h.set_string("algorithm", self.algorithm)
        h.set_string("metric", self.metric.replace("_", " "))
        h.set_int("n_list", self.n_list)
        h.set_int("n_probe", self._n_probe)
        h.set_int("k-means_iter", self.kmeans_iter)
        h.set_int("seed", self.seed)
        h.set_real("train fraction", self.train_fraction)
        h.set_string("check data", "yes" if self.check_data else "no")
        h.set_string("storage order", "column-major" if self.order == L.COLUMN_MAJOR else "row-major")
        n, d = X.shape
        self._X = X  # the C API keeps the training pointer until training
        h.call("da_approx_nn_set_training_data", n, d, L.ptr(X), n if self.order == L.COLUMN_MAJOR else d)
        self._h = h
        return X

    def train(self, X_train):
        self._prepare(X_train)
        self._h.call("da_approx_nn_train")
        return self

    def train_and_add(self, X_train):
        self._prepare(X_train)
        self._h.call("da_approx_nn_train_and_add")
        return self

    def add(self, X_add):
        X, _, _ = L.prepare(X_add, dtype=self.dtype, order=self.order)
        n, d = X.shape
        self._h.call("da_approx_nn_add", n, d, L.ptr(X), n if self.order == L.COLUMN_MAJOR else d)
        return self

    def kneighbors(self, X_test, n_neighbors=0, return_distance=True):
        X, _, _ = L.prepare(X_test, dtype=self.dtype, order=self.order)
        nq, d = X.shape
        k = n_neighbors if n_neighbors > 0 else self.n_neighbors
        fo = "F" if self.order == L.COLUMN_MAJOR else "C"
        ind = np.empty((nq, k), dtype=np.int32, order=fo)
        dist = np.empty((nq, k), dtype=self.dtype, order=fo)
        self._h.call("da_approx_nn_kneighbors", nq, d, L.ptr(X), nq if fo == "F" else d,
                     L.ptr(ind), L.ptr(dist) if return_distance else None, k,
                     1 if return_distance else 0)
        return (dist, ind) if return_distance else ind

    @property
    def cluster_centroids(self):
        d = self._X.shape[1]
        r = np.zeros(self.n_list * d, dtype=self.dtype)
        dim = L.da_int(r.size)
        self._h.call("da_handle_get_result", L.ANN_CENTROIDS, C.byref(dim), L.ptr(r))
        if self.order == L.COLUMN_MAJOR:
            return r.reshape((self.n_list, d), order="F")
        return r.reshape((self.n_list, d))

    @property
    def list_sizes(self):
        r = np.zeros(self.n_list, dtype=np.int32)
        dim = L.da_int(self.n_list)
        L.check(L.lib.da_handle_get_result_int(self._h.h, L.ANN_LIST_SIZES, C.byref(dim), L.ptr(r)),
                self._h.h)
        return r

    @property
    def rinfo(self):
        r = np.zeros(4, dtype=self.dtype)
        dim = L.da_int(4)
        self._h.call("da_handle_get_result", L.RINFO, C.byref(dim), L.ptr(r))
        return r

    @property
    def n_index(self):
        """Number of vectors added to the index."""
        return int(self.rinfo[1])
