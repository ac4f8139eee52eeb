"""Stand-in for faiss.IndexFlatL2: exact L2 search by the GPU brute-force kNN
(python_interface/python_package/aoclda/faiss/_index_flat_l2.py:32-165 of the reference).

Distances are squared euclidean, as faiss reports them. The indexed vectors are kept on
the host as well (faiss's reconstruct calls read them back) and the GPU index is rebuilt
over all of them on each add.
"""
import numpy as np

from ..neighbors import nearest_neighbors
from ._shim import _Shim

METRIC_L2 = 1


class IndexFlatL2(_Shim):
    _SUPPORTED_ATTRS = frozenset({"aocl"})

    def __init__(self, d):
        self._d = int(d)
        self._data = None
        self._nn = None
        self.aocl = True

    d = property(lambda self: self._d)
    ntotal = property(lambda self: 0 if self._data is None else self._data.shape[0])
    is_trained = property(lambda self: True)
    metric_type = property(lambda self: METRIC_L2)
    metric_arg = property(lambda self: 0)

    def _fit(self):
        self._nn = nearest_neighbors(metric="sqeuclidean")
        self._nn.fit(self._data)

    def add(self, X):
        X = np.asarray(X)
        if X.ndim != 2 or X.shape[1] != self._d:
            raise ValueError(f"IndexFlatL2.add: expected an array with {self._d} columns, "
                             f"got shape {X.shape}")
        self._data = X if self._data is None else np.concatenate([self._data, X], axis=0)
        self._fit()

    def train(self, X):
        """A flat index needs no training."""

    def search(self, X, k):
        if self._nn is None:
            raise RuntimeError("IndexFlatL2.search: the index is empty; call add() first")
        dist, ind = self._nn.kneighbors(X, n_neighbors=k)
        return dist, ind.astype(np.int64)

    def assign(self, X, k):
        if self._nn is None:
            raise RuntimeError("IndexFlatL2.assign: the index is empty; call add() first")
        return self.search(X, k)[1]

    def reset(self):
        self._data = None
        self._nn = None

    def _need_data(self, where):
        if self.ntotal == 0:
            raise RuntimeError(f"IndexFlatL2.{where}: the index is empty; call add() first")

    def reconstruct(self, i):
        self._need_data("reconstruct")
        if not 0 <= i < self.ntotal:
            raise IndexError(f"IndexFlatL2.reconstruct: index {i} is outside [0, {self.ntotal})")
        return self._data[i].copy()

    def reconstruct_n(self, i0, ni):
        self._need_data("reconstruct_n")
        return self._data[i0:i0 + ni].copy()

    def reconstruct_batch(self, ids):
        self._need_data("reconstruct_batch")
        return self._data[np.asarray(ids)].copy()

    def search_and_reconstruct(self, X, k):
        """(D, I, R): R[q, j] is the vector of neighbour I[q, j], zeros where I is -1."""
        self._need_data("search_and_reconstruct")
        dist, ind = self.search(X, k)
        rec = self._data[np.where(ind < 0, 0, ind)]
        rec[ind < 0] = 0
        return dist, ind, rec.astype(np.float32)

    # Pickling keeps the vectors, not the GPU handle; unpickling rebuilds the GPU index.
    def __getstate__(self):
        return {"d": self._d, "data": self._data}

    def __setstate__(self, state):
        object.__setattr__(self, "_d", state["d"])
        object.__setattr__(self, "_data", state["data"])
        object.__setattr__(self, "_nn", None)
        object.__setattr__(self, "aocl", True)
        if self._data is not None:
            self._fit()
