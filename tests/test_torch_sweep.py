"""The plain-torch FP64 TDVP-PS step (the baseline the HIP path is timed
against) on the CPU, against the numpy oracle: same quantum numbers, the
same observables to 1e-8, the same Lanczos matvec counts except at
allclose checks decided at rounding level."""
import numpy as np
import pytest
import torch

from oracle import tdvp as orc
from renormalizer_amd.mps.torch_sweep import TorchSweepState, svd_qn_qr
from tests import helpers


def _mpo_t(mpo):
    return [torch.from_numpy(w).to(torch.complex128) for w in helpers.mpo_arrays(mpo)]


def _run(mps, mpo, dt, steps):
    st = helpers.to_state(mps)
    ts = TorchSweepState.from_mps(mps, "cpu")
    mpo_t = _mpo_t(mpo)
    for _ in range(steps):  # ts evolves in place: check each step before the next
        st, ls = orc.tdvp_ps(st, helpers.mpo_arrays(mpo), dt)
        ls2 = ts.evolve(mpo_t, dt)
        yield st, ts, ls, ls2


def _count_mismatch(ls, ls2):
    assert len(ls) == len(ls2)
    return sum(abs(a - b) > 0 for a, b in zip(ls, ls2)), max(abs(a - b) for a, b in zip(ls, ls2))


def test_torch_sweep_spin_boson():
    model = helpers.sbm_model(6, 4, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 8)
    for st, ts, ls, ls2 in _run(mps, mpo, 0.05, 4):
        for i in range(len(st)):
            np.testing.assert_allclose(ts.calc_1site_rdm(i).numpy(), orc.calc_1site_rdm(st, i), atol=1e-8)
        assert all(np.array_equal(a, b) for a, b in zip(st.qn, ts.qn))
        n, worst = _count_mismatch(ls, ls2)
        assert n <= max(1, len(ls) // 20) and worst <= 2


def test_torch_sweep_holstein_quantum_numbers():
    model = helpers.holstein_model(4, 4, g=0.7)
    mps, mpo = helpers.holstein_initial(model, 12)
    for st, ts, ls, ls2 in _run(mps, mpo, 0.1, 3):
        occ = helpers.e_occupations_oracle(st, model)
        occ2 = np.array([ts.calc_1site_rdm(i)[1, 1].real.item()
                         for i, b in enumerate(model.basis) if b.is_electron])
        np.testing.assert_allclose(occ2, occ, atol=1e-8)
        assert all(np.array_equal(a, b) for a, b in zip(st.qn, ts.qn))
        n, worst = _count_mismatch(ls, ls2)
        assert n <= max(1, len(ls) // 20) and worst <= 2


@pytest.mark.parametrize("system", ["L", "R"])
def test_torch_blocked_qr(system):
    """Reconstruction, isometry and triangularity of the blocked QR / RQ."""
    rng = np.random.default_rng(3)
    rows, cols = 24, 10
    c = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    if system == "R":
        c = c.T.copy()
        rows, cols = cols, rows
    qnl = rng.integers(0, 2, (rows, 1))
    qnr = 1 - rng.integers(0, 2, (cols, 1))
    blocks = orc.qn_blocks(qnl, qnr, np.array([1]))
    mask = np.zeros((rows, cols), dtype=bool)
    for _, _, ls, rs in blocks:
        mask[np.ix_(ls, rs)] = True
    c = np.where(mask, c, 0)
    u, ql, v, qr = svd_qn_qr(torch.from_numpy(c), qnl, qnr, np.array([1]), system)
    u, v = u.numpy(), v.numpy()
    np.testing.assert_allclose(u @ v.T, c, atol=1e-13)
    iso = u if system == "L" else v
    np.testing.assert_allclose(iso.conj().T @ iso, np.eye(iso.shape[1]), atol=1e-13)
    uo, qlo, vo, qro = orc.svd_qn(c, qnl, qnr, np.array([1]), QR=True, system=system, full_matrices=False)
    assert [tuple(q) for q in ql] == [tuple(q) for q in qlo]
    assert [tuple(q) for q in qr] == [tuple(q) for q in qro]
    # the triangular factor: upper triangular inside each block
    tri = v if system == "L" else u
    off = 0
    for _, _, ls, rs in blocks:
        k = min(len(ls), len(rs))
        sub = tri[(rs if system == "L" else ls)][:, off:off + k]
        if system == "L":
            assert np.allclose(np.tril(sub.T, -1), 0, atol=1e-13)
        else:
            assert np.allclose(np.tril(sub[len(ls) - k:], -1), 0, atol=1e-13)
        off += k
