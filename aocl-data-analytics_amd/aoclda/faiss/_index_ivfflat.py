"""Stand-in for faiss.IndexIVFFlat on the GPU IVF-Flat index
(python_interface/python_package/aoclda/faiss/_index_ivfflat.py:39-165 of the reference).

faiss metric 1 (METRIC_L2) maps to squared euclidean distances, metric 0
(METRIC_INNER_PRODUCT) to "inner product". The coarse quantizer passed in only gives the
dimension: the centroids come from the index's own k-means, whose iteration count and
seed are cp.niter and cp.seed; cp.max_points_per_centroid caps the training sample at
that many points per list.
"""
import warnings

import numpy as np

from ..neighbors import approximate_neighbors
from ._kmeans import _ClusteringParameters
from ._shim import _Shim

_METRICS = {1: "sqeuclidean", 0: "inner product"}


class IndexIVFFlat(_Shim):
    _SUPPORTED_ATTRS = frozenset({"cp", "aocl", "nprobe"})

    def __init__(self, quantizer, d, nlist, metric=1):
        if quantizer.d != d:
            raise ValueError(f"the quantizer's dimension {quantizer.d} differs from d={d}")
        if type(quantizer).__name__ not in ("IndexFlatL2", "IndexFlatIP"):
            warnings.warn("IndexIVFFlat ignores a quantizer that is not a flat index.",
                          UserWarning, stacklevel=2)
        if metric not in _METRICS:
            raise ValueError(f"unknown metric {metric}: expected 0 (METRIC_INNER_PRODUCT) or "
                             "1 (METRIC_L2)")
        self._d, self._nlist, self._metric = int(d), int(nlist), metric
        self._ann = None
        self._nprobe = 1
        self.cp = _ClusteringParameters(niter=25, seed=0)
        self.aocl = True

    d = property(lambda self: self._d)
    nlist = property(lambda self: self._nlist)
    metric_type = property(lambda self: self._metric)
    is_trained = property(lambda self: self._ann is not None)
    ntotal = property(lambda self: 0 if self._ann is None else self._ann.n_index)
    quantizer = property(lambda self: None)

    @property
    def nprobe(self):
        return self._nprobe

    @nprobe.setter
    def nprobe(self, value):
        object.__setattr__(self, "_nprobe", value)
        if self._ann is not None:
            self._ann.n_probe = value

    def train(self, X):
        X = np.asarray(X)
        if X.ndim != 2 or X.shape[1] != self._d:
            raise ValueError(f"IndexIVFFlat.train: expected an array with {self._d} columns, "
                             f"got shape {X.shape}")
        frac = 1.0
        if self.cp.max_points_per_centroid is not None:
            frac = min(1.0, self.cp.max_points_per_centroid * self._nlist / X.shape[0])
        ann = approximate_neighbors(n_neighbors=1, metric=_METRICS[self._metric],
                                    n_list=self._nlist, n_probe=self._nprobe,
                                    kmeans_iter=self.cp.niter, seed=self.cp.seed,
                                    train_fraction=frac)
        ann.train(X)
        self._ann = ann

    def _need_trained(self, where):
        if self._ann is None:
            raise RuntimeError(f"IndexIVFFlat.{where}: the index is not trained; "
                               "call train() first")

    def add(self, X):
        self._need_trained("add")
        self._ann.add(X)

    def search(self, X, k):
        self._need_trained("search")
        dist, ind = self._ann.kneighbors(X, n_neighbors=k)
        return dist, ind.astype(np.int64)

    def reset(self):
        self._ann = None

    def __getstate__(self):
        # The reference pickles the index through its model persistence, which this build
        # does not provide.
        raise TypeError("IndexIVFFlat cannot be pickled: model persistence is not supported "
                        "by the AOCL-DA MI355X backend")
