"""k-means and DBSCAN (python_interface/python_package/aoclda/clustering.py:36-441 of the
reference)."""
import ctypes as C

import numpy as np

from . import _lib as L


class kmeans:
    def __init__(self, n_clusters=1, initialization_method="k-means++", C=None, n_init=10,
                 max_iter=300, seed=-1, algorithm="lloyd", distance="euclidean",
                 normalize_data=False, tol=1.0e-4, empty_clusters="ignore",
                 afk_mcmc_samples=50, mixed_precision=False, low_precision_max_iter=200,
                 low_precision_tol=1.0e-2, check_data=False, precision="working"):
        # precision: operand precision of the MFMA cross term for fp32 data ("working" =
        # fp32, "bf16"); labels are re-checked exactly either way (MI355X extension option
        # "cross term precision").
        self.n_clusters_opt = n_clusters
        self.initialization_method = initialization_method
        self.C = C
        self.n_init, self.max_iter, self.seed = n_init, max_iter, seed
        self.algorithm, self.distance, self.normalize_data = algorithm, distance, normalize_data
        self.tol, self.empty_clusters = tol, empty_clusters
        self.afk_mcmc_samples = afk_mcmc_samples
        self.mixed_precision = mixed_precision
        self.low_precision_max_iter, self.low_precision_tol = low_precision_max_iter, low_precision_tol
        self.check_data = check_data
        self.precision = precision
        self.status = None
        self._h = None

    def fit(self, A):
        A, self.order, self.dtype = L.prepare(A)
        h = L.Handle(L.HANDLE_KMEANS, self.dtype)
        self._h = h
        if getattr(self, "profile", False):
            L.check(L.lib.da_amd_profile_enable(h.h, 1))
        init = self.initialization_method
        if self.C is not None:
            init = "supplied"
        h.set_string("initialization method", init)
        h.set_string("algorithm", self.algorithm)
        h.set_string("storage order", "column-major" if self.order == L.COLUMN_MAJOR else "row-major")
        h.set_string("empty clusters", self.empty_clusters)
        h.set_string("distance", self.distance)
        h.set_string("normalize data", "yes" if self.normalize_data else "no")
        h.set_string("mixed precision", "yes" if self.mixed_precision else "no")
        h.set_string("check data", "yes" if self.check_data else "no")
        h.set_string("cross term precision", self.precision)
        h.set_int("n_clusters", self.n_clusters_opt)
        h.set_int("n_init", self.n_init)
        h.set_int("max_iter", self.max_iter)
        h.set_int("seed", self.seed)
        h.set_int("afk-mc2 samples", self.afk_mcmc_samples)
        h.set_int("low precision max_iter", self.low_precision_max_iter)
        h.set_real("convergence tolerance", self.tol)
        h.set_real("low precision convergence tolerance", self.low_precision_tol)
        n, d = A.shape
        self._A = A  # the C API keeps the caller's pointer semantics
        ld = n if self.order == L.COLUMN_MAJOR else d
        h.call("da_kmeans_set_data", n, d, L.ptr(A), ld, allow=(L.CODE["incompatible_options"],))
        if self.C is not None:
            Cm, _, _ = L.prepare(self.C, dtype=self.dtype, order=self.order)
            self._C = Cm
            ldc = Cm.shape[0] if self.order == L.COLUMN_MAJOR else d
            h.call("da_kmeans_set_init_centres", L.ptr(Cm), ldc)
        self.status = h.call("da_kmeans_compute",
                             allow=(L.CODE["maxit"], L.CODE["incompatible_options"]))
        self.n_samples, self.n_features = n, d
        return self

    def _rinfo(self):
        r = np.zeros(6, dtype=self.dtype)
        dim = L.da_int(6)
        self._h.call("da_handle_get_result", L.RINFO, C.byref(dim), L.ptr(r))
        return r

    @property
    def cluster_centres(self):
        k, d = self.n_clusters, self.n_features
        r = np.zeros(k * d, dtype=self.dtype)
        dim = L.da_int(k * d)
        self._h.call("da_handle_get_result", L.KMEANS_CENTRES, C.byref(dim), L.ptr(r))
        if self.order == L.COLUMN_MAJOR:
            return r.reshape((k, d), order="F")
        return r.reshape((k, d))

    @property
    def labels(self):
        n = self.n_samples
        r = np.zeros(n, dtype=np.int32)
        dim = L.da_int(n)
        L.check(L.lib.da_handle_get_result_int(self._h.h, L.KMEANS_LABELS, C.byref(dim), L.ptr(r)),
                self._h.h)
        return r

    @property
    def inertia(self):
        return float(self._rinfo()[4])

    @property
    def n_iter(self):
        return int(self._rinfo()[3])

    @property
    def lp_n_iter(self):
        return int(self._rinfo()[5])

    @property
    def n_clusters(self):
        return int(self._rinfo()[2])

    def transform(self, X):
        X, _, _ = L.prepare(X, dtype=self.dtype, order=self.order)
        m, d = X.shape
        k = self.n_clusters
        out = np.empty((m, k), dtype=self.dtype, order="F" if self.order == L.COLUMN_MAJOR else "C")
        ldx = m if self.order == L.COLUMN_MAJOR else d
        ldo = m if self.order == L.COLUMN_MAJOR else k
        self._h.call("da_kmeans_transform", m, d, L.ptr(X), ldx, L.ptr(out), ldo)
        return out

    def kernel_times(self, names=("kmeans_assign", "kmeans_recheck", "kmeans_update")):
        """Per-kernel GPU time of the fit (set ``profile = True`` before fit): name ->
        (total ms, launches), from HIP events on the handle's stream."""
        out = {}
        for name in names:
            ms, cnt = C.c_double(), L.da_int()
            L.check(L.lib.da_amd_profile_get(self._h.h, name.encode(), C.byref(ms), C.byref(cnt)))
            if cnt.value:
                out[name] = (ms.value, cnt.value)
        return out

    @property
    def n_rechecked(self):
        """Samples re-assigned with exact double distances during the last fit (summed over
        iterations)."""
        v = L.da_int()
        L.check(L.lib.da_kmeans_last_compute_stats(self._h.h, C.byref(v)))
        return v.value

    def predict(self, Y):
        Y, _, _ = L.prepare(Y, dtype=self.dtype, order=self.order)
        m, d = Y.shape
        out = np.empty(m, dtype=np.int32)
        ld = m if self.order == L.COLUMN_MAJOR else d
        self._h.call("da_kmeans_predict", m, d, L.ptr(Y), ld, L.ptr(out))
        return out


def kmeans_batch(datasets, init_centres, max_iter=300, tol=1.0e-4, precision="working"):
    """Batched Lloyd k-means (MI355X extension da_kmeans_batch_device_s): instance i clusters
    the fp32 rows datasets[i] (n_i x d_i, d_i <= 64) from the initial centres init_centres[i]
    (k_i x d_i, k_i <= 64). precision: MFMA operand type of the assignment, "working" (fp32),
    "bf16" or "bf16x3" (hi + lo split; labels exact in every form). Returns (centres, labels, n_iter, inertia) as lists /
    arrays, one entry per instance."""
    X = [np.ascontiguousarray(a, dtype=np.float32) for a in datasets]
    C0 = [np.ascontiguousarray(c, dtype=np.float32) for c in init_centres]
    if len(X) != len(C0) or not X:
        raise ValueError("one initial centre array per dataset is required")
    n = np.array([a.shape[0] for a in X], dtype=np.int32)
    d = np.array([a.shape[1] for a in X], dtype=np.int32)
    k = np.array([c.shape[0] for c in C0], dtype=np.int32)
    xs = np.concatenate([a.ravel() for a in X])
    cs = np.concatenate([c.ravel() for c in C0])
    dx, dc, dl = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(L.lib.da_amd_malloc(xs.nbytes, C.byref(dx)))
    L.check(L.lib.da_amd_malloc(cs.nbytes, C.byref(dc)))
    L.check(L.lib.da_amd_malloc(int(n.sum()) * 4, C.byref(dl)))
    try:
        L.check(L.lib.da_amd_memcpy_htod(dx, L.ptr(xs), xs.nbytes))
        L.check(L.lib.da_amd_memcpy_htod(dc, L.ptr(cs), cs.nbytes))
        n_iter = np.zeros(len(X), dtype=np.int32)
        inertia = np.zeros(len(X), dtype=np.float32)
        ct = {"working": 0, "fp32": 0, "bf16": 1, "bf16x3": 2}[precision]
        L.check(L.lib.da_kmeans_batch_device_s(len(X), L.ptr(n), L.ptr(d), L.ptr(k), dx, dc, dl,
                                               max_iter, tol, ct, L.ptr(n_iter), L.ptr(inertia)))
        cout = np.empty_like(cs)
        lab = np.empty(int(n.sum()), dtype=np.int32)
        L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(cout), dc, cout.nbytes))
        L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(lab), dl, lab.nbytes))
    finally:
        for p in (dx, dc, dl):
            L.lib.da_amd_free(p)
        L.check(L.lib.da_kmeans_batch_release())
    centres, labels = [], []
    co = lo = 0
    for i in range(len(X)):
        centres.append(cout[co:co + k[i] * d[i]].reshape(k[i], d[i]))
        labels.append(lab[lo:lo + n[i]])
        co += k[i] * d[i]
        lo += n[i]
    return centres, labels, n_iter, inertia


class DBSCAN:
    """DBSCAN clustering (clustering.py:329-441 of the reference): the eps-neighbourhoods, core
    points, connected clusters and labels are computed on the GPU."""

    def __init__(self, min_samples=5, metric="euclidean", algorithm="auto", leaf_size=30,
                 eps=0.5, power=2.0, check_data=False):
        self.min_samples, self.metric, self.algorithm = min_samples, metric, algorithm
        self.leaf_size, self.eps, self.power, self.check_data = leaf_size, eps, power, check_data
        self._h = None

    def fit(self, A):
        A, self.order, self.dtype = L.prepare(A)
        h = L.Handle(L.HANDLE_DBSCAN, self.dtype)
        self._h = h
        h.set_int("min samples", self.min_samples)
        h.set_string("metric", self.metric)
        h.set_string("algorithm", self.algorithm.replace("_", " "))
        h.set_int("leaf size", self.leaf_size)
        h.set_real("eps", self.eps)
        h.set_real("power", self.power)
        h.set_string("check data", "yes" if self.check_data else "no")
        h.set_string("storage order", "column-major" if self.order == L.COLUMN_MAJOR else "row-major")
        n, d = A.shape
        h.call("da_dbscan_set_data", n, d, L.ptr(A), n if self.order == L.COLUMN_MAJOR else d)
        h.call("da_dbscan_compute")
        self._n = n
        return self

    def _int(self, query, size):
        r = np.zeros(max(size, 1), dtype=np.int32)
        dim = L.da_int(max(size, 1))
        L.check(L.lib.da_handle_get_result_int(self._h.h, query, C.byref(dim), L.ptr(r)),
                self._h.h)
        return r[:size]

    def _rinfo(self):
        r = np.zeros(9, dtype=self.dtype)
        dim = L.da_int(9)
        self._h.call("da_handle_get_result", L.RINFO, C.byref(dim), L.ptr(r))
        return r

    @property
    def labels(self):
        return self._int(L.DBSCAN_LABELS, self._n)

    @property
    def core_sample_indices(self):
        return self._int(L.DBSCAN_CORE_INDICES, self.n_core_samples)

    @property
    def n_clusters(self):
        return int(self._int(L.DBSCAN_N_CLUSTERS, 1)[0])

    @property
    def n_core_samples(self):
        return int(self._int(L.DBSCAN_N_CORE, 1)[0])

    @property
    def n_samples(self):
        return int(self._rinfo()[0])

    @property
    def n_features(self):
        return int(self._rinfo()[1])
