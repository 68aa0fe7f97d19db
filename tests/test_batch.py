"""evolve_batch argument handling (host only)."""
import pytest

from renormalizer_amd.mps.batch import evolve_batch


def test_evolve_batch_arguments():
    assert evolve_batch([], None, 0.1) == []
    with pytest.raises(ValueError):
        evolve_batch([object(), object()], [object()], 0.1)
    with pytest.raises(ValueError):
        evolve_batch([object()], object(), 0.1, nstreams=0)


def test_qn_step_plan_matches_oracle_bookkeeping():
    """The tensor-free quantum-number pass of one time step (used to freeze
    the grouped sweep's block layout) ends with the bond quantum numbers the
    oracle's TDVP-PS step produces, and visits the sites in its order."""
    import numpy as np

    from oracle import tdvp as orc
    from renormalizer_amd.mps.batch import _qn_step_plan
    from tests import helpers
    model = helpers.holstein_model(4, 4, j=-1.0, g=1.0, elocal=[0.1, -0.3, 0.2, 0.0])
    mps, mpo = helpers.holstein_initial(model, 8)
    st = helpers.to_state(mps)
    halves, after = _qn_step_plan(mps)
    st1, _ = orc.tdvp_ps(st, helpers.mpo_arrays(mpo), 0.02)
    assert [np.array(q).tolist() for q in after.qn] == [np.array(q).tolist() for q in st1.qn]
    assert after.qnidx == st1.qnidx and after.to_right == st1.to_right
    n = len(mps)
    first = list(range(n)) if mps.to_right else list(range(n - 1, -1, -1))
    assert [s for s, _, _ in halves[0]] == first
    assert [s for s, _, _ in halves[1]] == first[::-1]
    assert halves[0][-1][1] is None and all(d is not None for _, d, _ in halves[0][:-1])
