"""Python front end of the MI355X build of AOCL-DA's distance path.

Mirrors the reference package's entry points for this path
(python_interface/python_package/aoclda/{metrics,neighbors,clustering}.py):
``aoclda.metrics.pairwise_distances``, ``aoclda.neighbors.nearest_neighbors`` and
``aoclda.clustering.kmeans``. Every call goes through the C ABI of libaoclda_amd.so.
"""
from . import clustering, metrics, neighbors  # noqa: F401
from ._lib import DAError  # noqa: F401
