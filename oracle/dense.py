"""Dense Hamiltonians and exact propagation for small models.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Built independently of the MPO construction: local operator matrices are
written from their closed forms (harmonic oscillator ladder operators,
Pauli matrices, two-level fermion) and combined with Kronecker products in
the site order of the reference models (spin-boson: spin, mode 0, mode 1,
...; Holstein scheme 2: e0, ph0, e1, ph1, ...).  Pinned by: the reference's
own closed-form cases (Delta-only spin oscillation, g = 0 exciton hopping)
in tests/test_oracle.py.
"""
import numpy as np
import scipy.linalg


def sho_ops(omega, n):
    """x, p^2, x^2 of a harmonic mode truncated to n levels with the
    reference's conventions (x = sqrt(1/2w)(b^+ + b), basis.py:227-265)."""
    b = np.diag(np.sqrt(np.arange(1, n)), k=1)
    bd = b.T
    x = np.sqrt(0.5 / omega) * (bd + b)
    # p = i sqrt(w/2) (b^+ - b); p^2 and x^2 from normal-ordered forms so the
    # highest level is treated like the reference (not x @ x)
    bb = np.diag(np.sqrt(np.arange(1, n - 1) * np.arange(2, n)), k=2) if n > 1 else np.zeros((1, 1))
    bdbd = bb.T
    nop = np.diag(np.arange(n)).astype(float)
    x2 = 0.5 / omega * (bdbd + nop + (nop + np.eye(n)) + bb)
    p2 = -omega / 2 * (bdbd - nop - (nop + np.eye(n)) + bb)
    return x, p2, x2


SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.diag([1.0, -1.0])
ADAG = np.array([[0.0, 0.0], [1.0, 0.0]])
NE = np.array([[0.0, 0.0], [0.0, 1.0]])


def _embed(ops_at, dims):
    """Kronecker product with ops_at = {site: matrix}, identity elsewhere."""
    out = np.ones((1, 1))
    for i, d in enumerate(dims):
        out = np.kron(out, ops_at.get(i, np.eye(d)))
    return out


def spin_boson_dense(eps, delta, omegas, displacements, nlevels):
    """H = eps sz + Delta sx + sum (p^2 + w^2 x^2)/2 - sz sum w^2 D x."""
    dims = [2] + [nlevels] * len(omegas)
    h = eps * _embed({0: SZ}, dims) + delta * _embed({0: SX}, dims)
    for k, (w, dis) in enumerate(zip(omegas, displacements)):
        x, p2, x2 = sho_ops(w, nlevels)
        h += _embed({k + 1: 0.5 * p2 + 0.5 * w ** 2 * x2}, dims)
        h += _embed({0: SZ, k + 1: -w ** 2 * dis * x}, dims)
    return h


def holstein_dense(elocal, j_matrix, omegas, displacements, nlevels):
    """One mode per molecule, scheme-2 site order [e0, ph0, e1, ph1, ...].
    H = sum_ij J_ij a_i^+ a_j (diag: elocalex + reorganisation energy)
        + sum (p^2 + w^2 x^2)/2 - sum w^2 D n_i x_i."""
    nmol = len(elocal)
    dims = []
    for _ in range(nmol):
        dims += [2, nlevels]
    h = np.zeros((int(np.prod(dims)),) * 2)
    for i in range(nmol):
        for j in range(nmol):
            if i == j:
                f = elocal[i] + 0.5 * displacements[i] ** 2 * omegas[i] ** 2
                h += f * _embed({2 * i: NE}, dims)
            elif j_matrix[i, j] != 0:
                # a_i^+ a_j between distinct two-level sites (no Jordan-Wigner
                # string: the reference's simple-electron basis is bosonic)
                h += j_matrix[i, j] * _embed({2 * i: ADAG, 2 * j: ADAG.T}, dims)
    for i in range(nmol):
        x, p2, x2 = sho_ops(omegas[i], nlevels)
        w = omegas[i]
        h += _embed({2 * i + 1: 0.5 * p2 + 0.5 * w ** 2 * x2}, dims)
        h += _embed({2 * i: NE, 2 * i + 1: -w ** 2 * displacements[i] * x}, dims)
    return h


def product_state(vectors):
    out = np.ones(1)
    for v in vectors:
        out = np.kron(out, v)
    return out


def propagate(h, psi0, times):
    """Exact states exp(-i H t) psi0 for each t (dense eigendecomposition)."""
    w, u = scipy.linalg.eigh(h)
    c = u.conj().T @ psi0
    return [u @ (np.exp(-1j * w * t) * c) for t in times]


def site_rdm(psi, dims, site):
    """One-site reduced density matrix in the convention of calc_1site_rdm
    (reference mps.py:1587-1596): rdm[i, j] = sum_rest conj(psi[i, rest]) psi[j, rest]."""
    t = psi.reshape(dims)
    t = np.moveaxis(t, site, 0).reshape(dims[site], -1)
    return t.conj() @ t.T


def expectation(psi, op):
    return np.vdot(psi, op @ psi)
