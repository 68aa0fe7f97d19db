"""Device QN-blocked SVD (reno_block_svd, svd_qn's full-matrices branch)
against the host LAPACK path on the same blocks: singular values to 1e-12,
complete bases (unitary U and V), the reconstruction u diag(s) v^T of the
allowed blocks, the reference's column layout and quantum-number lists."""
import numpy as np
import pytest
import torch

from oracle import tdvp as orc
from renormalizer_amd.mps.gs import _blocked_svd, _blocked_svd_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols,qn", [(24, 10, True), (10, 24, True), (60, 60, False), (160, 40, True),
                                          (7, 3, False)])
def test_block_svd_matches_host(dev, rows, cols, qn):
    print(f"[block_svd {rows}x{cols} qn={qn}]", flush=True)
    rng = np.random.default_rng(rows * 7 + cols)
    c = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    if qn:
        qnl = rng.integers(0, 3, (rows, 1))
        qnr = 2 - rng.integers(0, 3, (cols, 1))
        qntot = np.array([2])
    else:
        qnl = np.zeros((rows, 1), dtype=int)
        qnr = np.zeros((cols, 1), dtype=int)
        qntot = np.array([0])
    mask = np.zeros((rows, cols), dtype=bool)
    for _, _, ls, rs in orc.qn_blocks(qnl, qnr, qntot):
        mask[np.ix_(ls, rs)] = True
    # nearly dependent columns, as the sweeps meet them, then the QN mask
    if cols > 4:
        c[:, :2] = c[:, 2:4] * 1e-17
    c = np.where(mask, c, 0)
    ct = torch.from_numpy(c).to(dev)
    U, SU, qnl_d, V, SV, qnr_d = _blocked_svd(ct, qnl, qnr, qntot)
    Uh, SUh, qnl_h, Vh, SVh, qnr_h = _blocked_svd_host(c, qnl, qnr, qntot)
    Uh, Vh = Uh.numpy(), Vh.numpy()
    assert qnl_d == qnl_h and qnr_d == qnr_h
    np.testing.assert_allclose(SU, SUh, atol=1e-12 * SUh.max())
    np.testing.assert_allclose(SV, SVh, atol=1e-12 * SVh.max())
    U, V = U.cpu().numpy(), V.cpu().numpy()
    assert U.shape == Uh.shape and V.shape == Vh.shape
    # complete bases inside each block
    np.testing.assert_allclose(U.conj().T @ U, np.eye(U.shape[1]), atol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(V.shape[1]), atol=1e-12)
    kt = sum(min(len(ls), len(rs)) for _, _, ls, rs in orc.qn_blocks(qnl, qnr, qntot))
    rec = (U[:, :kt] * SU[:kt]) @ V[:, :kt].T
    np.testing.assert_allclose(rec, c, atol=1e-12 * np.abs(c).max())
    # singular columns span the same spaces as the host's (distinct values)
    for j in range(kt):
        if SUh[j] > 1e-8 * SUh.max() and all(abs(SUh[j] - SUh[i]) > 1e-8 for i in range(kt) if i != j):
            ph = np.vdot(Uh[:, j], U[:, j])
            assert abs(abs(ph) - 1) < 1e-9
            np.testing.assert_allclose(U[:, j], ph * Uh[:, j], atol=1e-9)
            np.testing.assert_allclose(V[:, j], np.conj(ph) * Vh[:, j], atol=1e-9)


@pytest.mark.parametrize("n,qn", [(40, True), (64, False), (9, True)])
def test_block_eigh_matches_host(dev, n, qn):
    """Device per-sector eigendecomposition of a reduced density matrix
    (reno_block_eigh) against the host path: eigenvalues to 1e-12,
    orthonormal eigenvectors that diagonalise
    each sector, identical quantum-number lists."""
    print(f"[block_eigh {n} qn={qn}]", flush=True)
    from renormalizer_amd.mps.gs import _blocked_dm_eigh
    rng = np.random.default_rng(n)
    qnbig = rng.integers(0, 3, (n, 1)) if qn else np.zeros((n, 1), dtype=int)
    comp = 2 - rng.integers(0, 3, (n + 3, 1)) if qn else np.zeros((n + 3, 1), dtype=int)
    qntot = np.array([2]) if qn else np.array([0])
    a = rng.standard_normal((n, n + 3)) + 1j * rng.standard_normal((n, n + 3))
    mask = (qnbig + comp.T == qntot[0]) if qn else np.ones((n, n + 3), dtype=bool)
    a = np.where(mask, a, 0)
    ddm = a @ a.conj().T
    u, s, q = _blocked_dm_eigh(torch.from_numpy(ddm).to(dev), qnbig, comp, qntot)
    uh, sh, qh = _blocked_dm_eigh(ddm, qnbig, comp, qntot)
    assert q == qh
    # eigenvalues (the square roots of rounding-level eigenvalues of the
    # rank-deficient sectors differ at sqrt(eps))
    np.testing.assert_allclose(s ** 2, sh ** 2, atol=1e-12 * max(1.0, sh.max() ** 2))
    u = u.cpu().numpy()
    np.testing.assert_allclose(u.conj().T @ u, np.eye(u.shape[1]), atol=1e-12)
    d = u.conj().T @ ddm @ u
    np.testing.assert_allclose(np.diag(d).real, s ** 2, atol=1e-10 * max(1.0, sh.max() ** 2))
