"""faiss front end on the GPU kNN, IVF-Flat and k-means
(python_interface/python_package/aoclda/faiss/ of the reference): stand-ins for
faiss.IndexFlatL2, faiss.IndexIVFFlat and faiss.Kmeans, and faiss_patch /
undo_faiss_patch to put them in faiss's namespace. ``python -m aoclda.faiss script.py``
runs a script with the patch applied.
"""
from ._index_flat_l2 import IndexFlatL2  # noqa: F401
from ._index_ivfflat import IndexIVFFlat  # noqa: F401
from ._kmeans import Kmeans  # noqa: F401
from .patch_faiss import faiss_patch, undo_faiss_patch

__all__ = ["faiss_patch", "undo_faiss_patch", "IndexFlatL2", "IndexIVFFlat", "Kmeans"]
