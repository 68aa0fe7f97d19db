"""evolve_batch argument handling (host only)."""
import pytest

from renormalizer_amd.mps.batch import evolve_batch


def test_evolve_batch_arguments():
    assert evolve_batch([], None, 0.1) == []
    with pytest.raises(ValueError):
        evolve_batch([object(), object()], [object()], 0.1)
    with pytest.raises(ValueError):
        evolve_batch([object()], object(), 0.1, nstreams=0)
