"""Stand-ins for faiss.ClusteringParameters and faiss.Kmeans on the GPU k-means
(python_interface/python_package/aoclda/faiss/_kmeans.py:38-192 of the reference).

faiss's niter, nredo, seed and spherical map to the k-means options max_iter, n_init,
seed and distance ("cosine" when spherical); the AOCL-DA options are keyword arguments.
"""
import warnings

import numpy as np

from ..clustering import kmeans
from ._shim import _Shim


class _ClusteringParameters(_Shim):
    """The one parameter struct that faiss shares between Kmeans and IndexIVF."""

    _SUPPORTED_ATTRS = frozenset(
        {"niter", "nredo", "seed", "spherical", "max_points_per_centroid"})

    def __init__(self, niter=25, nredo=1, seed=-1, spherical=False,
                 max_points_per_centroid=None):
        self.niter, self.nredo, self.seed = niter, nredo, seed
        self.spherical = spherical
        self.max_points_per_centroid = max_points_per_centroid


class Kmeans(_Shim):
    _SUPPORTED_ATTRS = frozenset({"aocl", "cp"})

    def __init__(self, d, k, niter=25, nredo=1, verbose=False, seed=-1, spherical=False, *,
                 algorithm="lloyd", initialization_method="k-means++", normalize_data=False,
                 tol=1.0e-4, empty_clusters="ignore", afk_mcmc_samples=50,
                 mixed_precision=False, low_precision_max_iter=10, low_precision_tol=1.0e-2,
                 **kwargs):
        for name in list(kwargs) + (["verbose"] if verbose else []):
            warnings.warn(f"Kmeans: '{name}' is not supported by the AOCL-DA MI355X backend.",
                          UserWarning, stacklevel=2)
        self._d, self._k = int(d), int(k)
        self._model = None
        self.aocl = True
        self.cp = _ClusteringParameters(niter=niter, nredo=nredo, seed=seed,
                                        spherical=spherical)
        self._opts = dict(algorithm=algorithm, initialization_method=initialization_method,
                          normalize_data=normalize_data, tol=tol,
                          empty_clusters=empty_clusters, afk_mcmc_samples=afk_mcmc_samples,
                          mixed_precision=mixed_precision,
                          low_precision_max_iter=low_precision_max_iter,
                          low_precision_tol=low_precision_tol)

    d = property(lambda self: self._d)
    k = property(lambda self: self._k)
    is_trained = property(lambda self: self._model is not None)

    @property
    def centroids(self):
        return None if self._model is None else self._model.cluster_centres

    def train(self, x):
        x = np.asarray(x)
        if x.ndim != 2 or x.shape[1] != self._d:
            raise ValueError(f"Kmeans.train: expected an array with {self._d} columns, "
                             f"got shape {x.shape}")
        model = kmeans(n_clusters=self._k, max_iter=self.cp.niter, n_init=self.cp.nredo,
                       seed=self.cp.seed, distance="cosine" if self.cp.spherical else "euclidean",
                       **self._opts)
        model.fit(x)
        self._model = model

    def assign(self, x):
        """(D, I): the transform value of each sample's nearest centre and that centre."""
        if self._model is None:
            raise RuntimeError("Kmeans.assign: not trained; call train() first")
        dist = self._model.transform(x)
        ind = dist.argmin(axis=1).astype(np.int64)
        return dist[np.arange(dist.shape[0]), ind], ind
