"""The Davidson-Liu solver of the DMRG local problem (mps/gs.py davidson,
reference lib/davidson/davidson.py:154-441) and its device vector kernels
(reno_dav_orth / reno_dav_dots / reno_dav_ritz / reno_dav_precond).

* The kernels against numpy on random data (GPU).
* The solver on dense Hermitian matrices with a known spectrum: the lowest
  one and three eigenvalues to 1e-10 and the eigenvectors to 1e-5 (the
  residual tolerance sqrt(1e-12)), run twice: on the GPU through the
  library, and on the CPU with the four kernels stood in by torch versions of
  their contracts (the solver's host logic without a GPU)."""
import numpy as np
import pytest
import torch

from renormalizer_amd import native
from renormalizer_amd.mps import gs


def _orth(V, x, thresh):
    for _ in range(2):
        if V.shape[0]:
            x -= V.T @ (V.conj() @ x)
    n2 = float(torch.vdot(x, x).real)
    if n2 >= thresh:
        x /= np.sqrt(n2)
    return n2


def _dots(V, y):
    return V.conj() @ y


def _ritz(V, AV, U, e):
    Ut = torch.from_numpy(np.ascontiguousarray(U)).to(V.dtype)
    X = Ut.T @ V
    AX = Ut.T @ AV
    R = AX - torch.from_numpy(np.asarray(e)).to(V.dtype)[:, None] * X
    return X, AX, R, (R.abs() ** 2).sum(dim=1).numpy()


def _precond(r, hdiag, mask, e):
    return torch.where(mask, r / (hdiag - e + 1e-4), torch.zeros_like(r))


@pytest.fixture(params=["stand-in", pytest.param("device", marks=pytest.mark.gpu)])
def target(request, monkeypatch):
    if request.param == "device":
        return torch.device("cuda", 0)
    for name, f in (("dav_orth", _orth), ("dav_dots", _dots), ("dav_ritz", _ritz), ("dav_precond", _precond)):
        monkeypatch.setattr(native, name, f)
    return torch.device("cpu")


def _problem(n, seed, masked):
    """Hermitian H = Q diag(w) Q^H with a gap above the lowest eigenvalues
    and a dominant diagonal (the case the diagonal preconditioner is made
    for); with ``masked`` the last fifth of the entries is outside the
    allowed set (H acts as zero there and the guesses vanish there)."""
    rng = np.random.default_rng(seed)
    m = n - n // 5 if masked else n
    a = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    h = np.diag(np.linspace(-3.0, 5.0, m)) + 0.05 * (a + a.conj().T)
    full = np.zeros((n, n), dtype=complex)
    full[:m, :m] = h
    mask = np.zeros(n, dtype=bool)
    mask[:m] = True
    return full, mask


@pytest.mark.parametrize("nroots", [1, 3])
@pytest.mark.parametrize("masked", [False, True])
def test_davidson_dense(target, nroots, masked):
    n = 300
    h, mask = _problem(n, 11 + nroots, masked)
    ht = torch.from_numpy(h).to(target)
    mt = torch.from_numpy(mask).to(target)

    def aop(x):
        return (ht @ x) * mt

    hdiag = torch.from_numpy(np.real(np.diag(h)).copy()).to(target)
    rng = np.random.default_rng(3)
    guesses = []
    for _ in range(nroots):
        g = np.zeros(n, dtype=complex)
        g[:5] = rng.standard_normal(5)
        guesses.append(torch.from_numpy(g).to(target) * mt)
    e, xs = gs.davidson(aop, guesses, hdiag, mt, nroots=nroots)
    sub = h[np.ix_(mask, mask)]
    w, v = np.linalg.eigh(sub)
    es = np.atleast_1d(e)
    np.testing.assert_allclose(es, w[:nroots], atol=1e-10)
    for i, x in enumerate(xs):
        x = x.cpu().numpy()
        assert np.abs(x[~mask]).max(initial=0.0) == 0.0
        ov = abs(np.vdot(v[:, i], x[mask]))
        assert abs(ov - 1) < 1e-8, ov


@pytest.mark.gpu
def test_dav_kernels_match_numpy():
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(7)

    def rnd(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    for n, k in [(1, 1), (257, 3), (5000, 12), (300001, 7)]:
        V = np.linalg.qr(rnd(n, k))[0].T.copy() if n >= k else rnd(k, n)
        x = rnd(n)
        Vd = torch.from_numpy(V).to(dev)
        xd = torch.from_numpy(x.copy()).to(dev)
        n2 = native.dav_orth(Vd, xd, 1e-14)
        y = x.copy()
        for _ in range(2):
            y -= V.T @ (V.conj() @ y)
        assert abs(n2 - np.vdot(y, y).real) <= 1e-12 * max(1.0, n2)
        expect = y / np.sqrt(np.vdot(y, y).real) if n2 >= 1e-14 else y  # a dependent vector is left as it is
        np.testing.assert_allclose(xd.cpu().numpy(), expect, atol=1e-12)
        yv = rnd(n)
        d = native.dav_dots(Vd, torch.from_numpy(yv).to(dev)).cpu().numpy()
        np.testing.assert_allclose(d, V.conj() @ yv, rtol=1e-12, atol=1e-12 * np.sqrt(n))
        AV = rnd(k, n)
        U = rnd(k, min(k, 3))
        e = rng.standard_normal(U.shape[1])
        X, AX, R, rn2 = native.dav_ritz(Vd, torch.from_numpy(AV).to(dev), U, e)
        Xr, AXr = U.T @ V, U.T @ AV
        Rr = AXr - e[:, None] * Xr
        np.testing.assert_allclose(X.cpu().numpy(), Xr, atol=1e-12)
        np.testing.assert_allclose(AX.cpu().numpy(), AXr, atol=1e-12)
        np.testing.assert_allclose(R.cpu().numpy(), Rr, atol=1e-12)
        np.testing.assert_allclose(rn2, (np.abs(Rr) ** 2).sum(axis=1), rtol=1e-12)
        hd = rng.standard_normal(n)
        mk = rng.random(n) < 0.7
        p = native.dav_precond(R[0].contiguous(), torch.from_numpy(hd).to(dev), torch.from_numpy(mk).to(dev), e[0])
        np.testing.assert_allclose(p.cpu().numpy(), np.where(mk, Rr[0] / (hd - e[0] + 1e-4), 0), rtol=1e-13)
    # an empty basis only normalises
    xd = torch.from_numpy(rnd(100)).to(dev)
    x0 = xd.cpu().numpy()
    n2 = native.dav_orth(torch.empty((0, 100), dtype=torch.complex128, device=dev), xd, 1e-14)
    np.testing.assert_allclose(n2, np.vdot(x0, x0).real, rtol=1e-13)
    np.testing.assert_allclose(xd.cpu().numpy(), x0 / np.sqrt(n2), atol=1e-14)
