"""Mps .npz dump/load, protocol "0.4" (reference mp.py:1085-1113,
mps.py:352-386): round trip, the keys the reference's loader reads, and
files laid out as the reference writes them.  Host only."""
import numpy as np
import pytest

from renormalizer_amd import Mps
from tests import helpers


def _state():
    model = helpers.holstein_model(3, 3, j=-1.0, g=1.0)
    mps, _ = helpers.holstein_initial(model, 6)
    return model, mps


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a.numpy_sites(), b.numpy_sites()):
        np.testing.assert_array_equal(x, y)
    assert [np.asarray(q).tolist() for q in a.qn] == [np.asarray(q).tolist() for q in b.qn]
    assert a.qnidx == b.qnidx and a.to_right == b.to_right
    np.testing.assert_array_equal(a.qntot, b.qntot)
    assert a.coeff == b.coeff


def test_round_trip(tmp_path):
    model, mps = _state()
    mps.coeff = 0.5 - 0.25j
    fname = tmp_path / "mps.npz"
    mps.dump(fname)
    _same(mps, Mps.load(model, fname))


def test_keys_read_by_reference_loader(tmp_path):
    """Every key the reference's load reads is present, ``qn`` as the
    object array it expects (reference mps.py:362-380)."""
    model, mps = _state()
    fname = tmp_path / "mps.npz"
    mps.dump(fname)
    with np.load(fname, allow_pickle=True) as f:   # a file written by this test
        keys = set(f.files)
        assert {"version", "nsites", "qn", "qnidx", "qntot", "to_right", "coeff"} <= keys
        assert str(f["version"]) == "0.4"
        n = int(f["nsites"])
        assert {f"mt_{i}" for i in range(n)} <= keys
        assert f["qn"].dtype == object and len(f["qn"]) == n + 1
        for i in range(n + 1):
            np.testing.assert_array_equal(np.asarray(f["qn"][i]), f[f"subqn_{i}"])


def _reference_layout(mps, version="0.4"):
    """The dictionary the reference's dump builds (mp.py:1093-1109)."""
    d = {"version": version, "nsites": len(mps)}
    for i, mt in enumerate(mps.numpy_sites()):
        d[f"mt_{i}"] = mt
    qn = [np.asarray(q) for q in mps.qn]
    arr = np.empty(len(qn), object)
    arr[:] = qn
    d["qn"] = arr
    if version == "0.4":   # only protocol 0.4 writes the subqn_<i> arrays
        for i, q in enumerate(qn):
            d[f"subqn_{i}"] = q
    d["qnidx"] = mps.qnidx
    d["qntot"] = np.asarray(mps.qntot)
    if version == "0.1":
        d["left"] = mps.to_right
    else:
        d["to_right"] = mps.to_right
        d["coeff"] = mps.coeff
    return d


def test_load_reference_layout(tmp_path):
    model, mps = _state()
    fname = tmp_path / "ref.npz"
    np.savez(fname, **_reference_layout(mps, "0.4"))
    _same(mps, Mps.load(model, fname))


@pytest.mark.parametrize("version", ["0.1", "0.2", "0.3", "9.9"])
def test_load_refuses(tmp_path, version):
    model, mps = _state()
    fname = tmp_path / "ref.npz"
    np.savez(fname, **_reference_layout(mps, version))
    with pytest.raises(ValueError):
        Mps.load(model, fname)
