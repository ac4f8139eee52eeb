"""CPU tests of the faiss front end (aoclda.faiss), after the reference's
python_interface/tests/faissAPI/{patch_faiss,flat_l2_faiss,ivfflat_faiss,kmeans_faiss}_test.py:
symbol swapping, index_factory parsing, properties and argument checks that act before any
GPU call. faiss is not installed, so a stand-in module plays its part."""
import sys
import types

import numpy as np
import pytest

from aoclda.faiss import IndexFlatL2, IndexIVFFlat, Kmeans, faiss_patch, undo_faiss_patch
from aoclda.faiss import patch_faiss as pf


class _NativeFlat:
    def __init__(self, d):
        self.d = d


class _NativeIVF:
    pass


class _NativeKmeans:
    pass


def _native_factory(d, desc, metric=1):
    return ("native", d, desc, metric)


@pytest.fixture
def faiss(monkeypatch):
    mod = types.ModuleType("faiss")
    mod.IndexFlatL2, mod.IndexIVFFlat, mod.Kmeans = _NativeFlat, _NativeIVF, _NativeKmeans
    mod.index_factory = _native_factory
    mod.METRIC_L2, mod.METRIC_INNER_PRODUCT = 1, 0
    monkeypatch.setitem(sys.modules, "faiss", mod)
    monkeypatch.setattr(pf, "_saved", {})
    return mod


def test_patch_and_undo_all(faiss, capsys):
    faiss_patch()
    out = capsys.readouterr().out
    assert all(s in out for s in ("IndexIVFFlat", "IndexFlatL2", "index_factory", "Kmeans"))
    assert faiss.IndexFlatL2 is IndexFlatL2 and faiss.IndexIVFFlat is IndexIVFFlat
    assert faiss.Kmeans is Kmeans and faiss.index_factory is pf._index_factory
    undo_faiss_patch(print_patched=False)
    assert faiss.IndexFlatL2 is _NativeFlat and faiss.IndexIVFFlat is _NativeIVF
    assert faiss.Kmeans is _NativeKmeans and faiss.index_factory is _native_factory


def test_patch_one_and_list(faiss, capsys):
    faiss_patch("IndexFlatL2", print_patched=False)
    assert faiss.IndexFlatL2 is IndexFlatL2 and faiss.IndexIVFFlat is _NativeIVF
    faiss_patch(["Kmeans", "NoSuchSymbol"], print_patched=False)
    assert faiss.Kmeans is Kmeans
    assert "NoSuchSymbol was not found" in capsys.readouterr().out
    undo_faiss_patch(("IndexFlatL2", "Kmeans"), print_patched=False)
    assert faiss.IndexFlatL2 is _NativeFlat and faiss.Kmeans is _NativeKmeans
    with pytest.raises(TypeError):
        faiss_patch(3)


def test_patch_message(faiss, capsys):
    faiss_patch("Kmeans")
    out = capsys.readouterr().out
    assert "enabled for the following symbols" in out and "Kmeans" in out
    undo_faiss_patch("Kmeans")
    assert "disabled" in capsys.readouterr().out


def test_index_factory(faiss):
    faiss_patch(print_patched=False)
    ivf = faiss.index_factory(16, "IVF64,Flat")
    assert isinstance(ivf, IndexIVFFlat) and ivf.nlist == 64 and ivf.d == 16
    assert ivf.metric_type == faiss.METRIC_L2
    ip = faiss.index_factory(16, " IVF8,Flat ", faiss.METRIC_INNER_PRODUCT)
    assert isinstance(ip, IndexIVFFlat) and ip.metric_type == 0
    flat = faiss.index_factory(16, "Flat")
    assert isinstance(flat, IndexFlatL2) and flat.d == 16
    assert faiss.index_factory(16, "Flat", faiss.METRIC_INNER_PRODUCT) == \
        ("native", 16, "Flat", 0)
    assert faiss.index_factory(16, "HNSW32") == ("native", 16, "HNSW32", 1)
    undo_faiss_patch(print_patched=False)


def test_patch_needs_faiss(monkeypatch):
    monkeypatch.setitem(sys.modules, "faiss", None)
    monkeypatch.setattr(pf, "_saved", {})
    with pytest.raises(ImportError):
        faiss_patch(print_patched=False)


def test_flat_properties_and_checks():
    idx = IndexFlatL2(16)
    assert idx.d == 16 and idx.ntotal == 0 and idx.is_trained
    assert idx.metric_type == 1 and idx.metric_arg == 0 and idx.aocl is True
    idx.train(np.zeros((4, 16)))  # no-op
    with pytest.raises(ValueError):
        idx.add(np.zeros((4, 15), dtype=np.float32))
    with pytest.raises(ValueError):
        idx.add(np.zeros(16, dtype=np.float32))
    with pytest.raises(RuntimeError):
        idx.search(np.zeros((2, 16), dtype=np.float32), 3)
    with pytest.raises(RuntimeError):
        idx.assign(np.zeros((2, 16), dtype=np.float32), 3)
    for call in (lambda: idx.reconstruct(0), lambda: idx.reconstruct_n(0, 2),
                 lambda: idx.reconstruct_batch([0]),
                 lambda: idx.search_and_reconstruct(np.zeros((1, 16)), 1)):
        with pytest.raises(RuntimeError):
            call()


def test_unsupported_attributes():
    idx = IndexFlatL2(8)
    with pytest.raises(NotImplementedError):
        _ = idx.range_search
    with pytest.raises(AttributeError):
        _ = idx.__copy__
    with pytest.warns(UserWarning, match="verbose"):
        idx.verbose = True
    assert idx.verbose is True


def test_ivf_properties_and_checks():
    q = IndexFlatL2(16)
    idx = IndexIVFFlat(q, 16, 32)
    assert idx.d == 16 and idx.nlist == 32 and not idx.is_trained and idx.ntotal == 0
    assert idx.metric_type == 1 and idx.quantizer is None
    assert idx.cp.niter == 25 and idx.cp.seed == 0 and idx.cp.max_points_per_centroid is None
    assert idx.cp.spherical is False
    idx.nprobe = 4
    assert idx.nprobe == 4
    with pytest.raises(ValueError):
        IndexIVFFlat(IndexFlatL2(8), 16, 32)
    with pytest.raises(ValueError):
        IndexIVFFlat(q, 16, 32, metric=2)
    with pytest.warns(UserWarning, match="quantizer"):
        IndexIVFFlat(types.SimpleNamespace(d=16), 16, 32)
    with pytest.raises(ValueError):
        idx.train(np.zeros((10, 15), dtype=np.float32))
    for call in (lambda: idx.add(np.zeros((2, 16))), lambda: idx.search(np.zeros((2, 16)), 1)):
        with pytest.raises(RuntimeError):
            call()
    with pytest.raises(NotImplementedError):
        _ = idx.make_direct_map
    with pytest.warns(UserWarning, match="polysemous"):
        idx.cp.polysemous = 1


def test_kmeans_properties_and_checks():
    with pytest.warns(UserWarning, match="gpu"):
        km = Kmeans(8, 4, niter=7, nredo=2, seed=3, gpu=True)
    assert km.d == 8 and km.k == 4 and not km.is_trained and km.centroids is None
    assert km.cp.niter == 7 and km.cp.nredo == 2 and km.cp.seed == 3 and not km.cp.spherical
    with pytest.warns(UserWarning, match="verbose"):
        Kmeans(8, 4, verbose=True)
    with pytest.raises(ValueError):
        km.train(np.zeros((10, 7)))
    with pytest.raises(RuntimeError):
        km.assign(np.zeros((2, 8)))
    with pytest.raises(NotImplementedError):
        _ = km.obj
