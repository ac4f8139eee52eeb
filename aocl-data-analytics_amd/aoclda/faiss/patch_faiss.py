"""Swap faiss's IndexFlatL2, IndexIVFFlat, Kmeans and index_factory for the GPU stand-ins
and back (python_interface/python_package/aoclda/faiss/patch_faiss.py:32-129 of the
reference). faiss itself is imported on the first call, so the stand-ins work without it.
"""
import re

_saved = {}


def _faiss():
    import faiss  # noqa: PLC0415  (only the patch needs faiss installed)
    if not _saved:
        _saved.update(IndexFlatL2=faiss.IndexFlatL2, IndexIVFFlat=faiss.IndexIVFFlat,
                      Kmeans=faiss.Kmeans, index_factory=faiss.index_factory)
    return faiss


def _index_factory(d, description, metric=None):
    """"IVF<nlist>,Flat" and "Flat" with METRIC_L2 build stand-ins; any other description
    goes to faiss's own factory."""
    from ._index_flat_l2 import IndexFlatL2
    from ._index_ivfflat import IndexIVFFlat
    faiss = _faiss()
    if metric is None:
        metric = faiss.METRIC_L2
    desc = description.strip()
    m = re.fullmatch(r"IVF(\d+),Flat", desc)
    if m:
        # The stand-in reads only the dimension from the quantizer, so a flat L2 quantizer
        # serves every metric.
        return IndexIVFFlat(IndexFlatL2(d), d, int(m.group(1)), metric)
    if desc == "Flat" and metric == faiss.METRIC_L2:
        return IndexFlatL2(d)
    return _saved["index_factory"](d, description, metric)


def _replacements():
    from ._index_flat_l2 import IndexFlatL2
    from ._index_ivfflat import IndexIVFFlat
    from ._kmeans import Kmeans
    return {"IndexIVFFlat": IndexIVFFlat, "IndexFlatL2": IndexFlatL2,
            "index_factory": _index_factory, "Kmeans": Kmeans}


def _names(args):
    if not args:
        return list(_replacements())
    if isinstance(args[0], str):
        return [args[0]]
    if isinstance(args[0], (list, tuple)):
        return list(args[0])
    raise TypeError("Unrecognized argument")


def _swap(args, print_patched, table, verb):
    done = []
    for name in _names(args):
        if name in table:
            setattr(_faiss(), name, table[name])
            done.append(name)
        else:
            print(f"The symbol {name} was not found.")
    if done and print_patched:
        print(f"AOCL Extension for faiss {verb} for the following symbols:")
        print(", ".join(done))


def faiss_patch(*args, print_patched=True):
    """Replace faiss symbols (all, one name, or a list of names) with the GPU stand-ins."""
    _faiss()
    _swap(args, print_patched, _replacements(), "enabled")


def undo_faiss_patch(*args, print_patched=True):
    """Put faiss's own symbols back."""
    _faiss()
    _swap(args, print_patched, _saved, "disabled")
