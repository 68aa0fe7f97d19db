"""Parity of the HIP kernels (through the C ABI) with the numpy oracle.

Shapes cover the edge sites and the full-size sites of the spin-boson and
Holstein workloads (bond dims 16/64/256, local dims 2/8/10, MPO bond 3..5).
FP64 mode must agree with the complex128 oracle to 1e-12 relative (the
contraction order differs, so rounding differs); the mixed mode (complex64
environments, FP32 MFMA) to 2e-5 relative.
"""
import numpy as np
import pytest
import torch

from oracle import tdvp as orc
from renormalizer_amd import native

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12
RTOL_MIXED = 2e-5


def _rand(rng, shape, cplx=True):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _t(a, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("prec", [native.PREC_FP64, native.PREC_MIXED])
def test_mfma_lane_layout(dev, prec):
    """The output-register layout gemm.hip assumes for v_mfma_*_16x16x4:
    f64: lane l, reg r -> C[4r + l/16][l%16]; f32: C[4(l/16) + r][l%16]."""
    print(f"[mfma layout prec={prec}]", flush=True)
    rng = np.random.default_rng(0)
    a = rng.integers(-4, 5, (16, 4)).astype(np.float64)
    b = rng.integers(-4, 5, (4, 16)).astype(np.float64)
    dt = torch.float64 if prec == native.PREC_FP64 else torch.float32
    raw = native.probe_mfma16(_t(a, dev, dt), _t(b, dev, dt), prec).cpu().double().numpy().reshape(64, 4)
    c = a @ b
    lanes = np.arange(64)
    layouts = {
        "4r+q": np.stack([c[4 * r + lanes // 16, lanes % 16] for r in range(4)], axis=1),
        "4q+r": np.stack([c[4 * (lanes // 16) + r, lanes % 16] for r in range(4)], axis=1),
    }
    found = [k for k, v in layouts.items() if np.array_equal(v, raw)]
    want = "4r+q" if prec == native.PREC_FP64 else "4q+r"
    assert found == [want], f"layout found {found}, gemm.hip assumes {want}"


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (70, 33, 17), (128, 130, 64), (5, 200, 300)])
def test_zgemm(dev, M, N, K):
    print(f"[zgemm {M}x{N}x{K}]", flush=True)
    rng = np.random.default_rng(M * 1000 + N + K)
    A = _rand(rng, (M, K))
    B = _rand(rng, (K, N))
    C = native.zgemm(_t(A, dev), _t(B, dev)).cpu().numpy()
    assert _rel(C, A @ B) < RTOL64


SHAPES = [
    # (Ml, d, Mr, Dl, Dr)
    (1, 2, 2, 1, 3),       # left edge spin site
    (16, 8, 16, 3, 3),     # spin-boson M=16
    (8, 10, 1, 3, 1),      # right edge
    (64, 2, 64, 5, 4),     # Holstein electron site
    (64, 10, 64, 4, 5),    # Holstein phonon site
    (33, 7, 45, 2, 3),     # ragged
    (40, 6, 100, 3, 7),    # ragged split-K chunks (long K, few output tiles)
    (256, 10, 256, 3, 3),  # spin-boson M=256 full site
]


def _env_inputs(rng, Ml, d, Mr, Dl, Dr):
    L = _rand(rng, (Ml, Dl, Ml))
    R = _rand(rng, (Mr, Dr, Mr))
    W = rng.standard_normal((Dl, d, d, Dr))
    C = _rand(rng, (Ml, d, Mr))
    return L, W, R, C


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [native.PREC_FP64, native.PREC_MIXED])
def test_hop1(dev, shape, prec):
    print(f"[hop1 {shape} prec={prec}]", flush=True)
    rng = np.random.default_rng(sum(shape))
    L, W, R, C = _env_inputs(rng, *shape)
    cdt, rdt = native.prec_dtypes(prec)
    out = native.hop_apply(_t(L, dev, cdt), _t(W, dev, rdt), _t(R, dev, cdt), _t(C, dev), prec)
    ref = orc.hop1(L, W, R, C)
    assert _rel(out.cpu().numpy(), ref) < (RTOL64 if prec == native.PREC_FP64 else RTOL_MIXED)


@pytest.mark.parametrize("shape", [(1, 1, 3), (16, 16, 3), (64, 64, 5), (40, 23, 4), (256, 256, 3)])
@pytest.mark.parametrize("prec", [native.PREC_FP64, native.PREC_MIXED])
def test_hop0(dev, shape, prec):
    print(f"[hop0 {shape} prec={prec}]", flush=True)
    Ml, Mr, D = shape
    rng = np.random.default_rng(Ml + Mr + D)
    L = _rand(rng, (Ml, D, Ml))
    R = _rand(rng, (Mr, D, Mr))
    S = _rand(rng, (Ml, Mr))
    cdt, _ = native.prec_dtypes(prec)
    out = native.hop_apply(_t(L, dev, cdt), None, _t(R, dev, cdt), _t(S, dev), prec)
    ref = orc.hop0(L, R, S)
    assert _rel(out.cpu().numpy(), ref) < (RTOL64 if prec == native.PREC_FP64 else RTOL_MIXED)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("domain", ["L", "R"])
@pytest.mark.parametrize("prec", [native.PREC_FP64, native.PREC_MIXED])
def test_env_update(dev, shape, domain, prec):
    print(f"[env {domain} {shape} prec={prec}]", flush=True)
    Ml, d, Mr, Dl, Dr = shape
    rng = np.random.default_rng(7 * sum(shape))
    L, W, R, C = _env_inputs(rng, *shape)
    env = L if domain == "L" else R
    cdt, rdt = native.prec_dtypes(prec)
    out = native.env_update(_t(env, dev, cdt), _t(C, dev), _t(W, dev, rdt), domain, prec=prec)
    ref = orc.contract_one_site(env, C, W, domain)
    assert _rel(out.cpu().numpy(), ref) < (RTOL64 if prec == native.PREC_FP64 else RTOL_MIXED)


def test_env_update_explicit_bra(dev):
    print("[env explicit bra]", flush=True)
    rng = np.random.default_rng(3)
    L, W, R, C = _env_inputs(rng, 12, 3, 9, 2, 3)
    bra = _rand(rng, C.shape)
    for domain, env in (("L", L), ("R", R)):
        out = native.env_update(_t(env, dev), _t(C, dev), _t(W, dev), domain, ms_conj=_t(bra, dev))
        ref = orc.contract_one_site(env, C, W, domain, ms_conj=bra)
        assert _rel(out.cpu().numpy(), ref) < RTOL64


def _herm_env_inputs(rng, Ml, d, Mr, Dl, Dr):
    """Environments/MPO making H_eff Hermitian: L, R Hermitian in (bra, ket)
    for each MPO index and W real-symmetric in (d_out, d_in)."""
    L = _rand(rng, (Ml, Dl, Ml))
    L = L + L.transpose(2, 1, 0).conj()
    R = _rand(rng, (Mr, Dr, Mr))
    R = R + R.transpose(2, 1, 0).conj()
    W = rng.standard_normal((Dl, d, d, Dr))
    W = W + W.transpose(0, 2, 1, 3)
    return L / Ml, W, R / Mr


@pytest.mark.parametrize("shape", [(1, 2, 2, 1, 3), (4, 3, 5, 2, 2), (16, 8, 16, 3, 3), (64, 10, 64, 3, 3)])
def test_expm_krylov_matches_oracle(dev, shape):
    """Same matvec count and result as the restated reference Lanczos."""
    print(f"[krylov {shape}]", flush=True)
    Ml, d, Mr, Dl, Dr = shape
    rng = np.random.default_rng(11 + sum(shape))
    L, W, R = _herm_env_inputs(rng, *shape)
    v = _rand(rng, (Ml, d, Mr))
    dt = -0.05j / 2
    out, nmv = native.expm_krylov(_t(L, dev), _t(W, dev), _t(R, dev), _t(v, dev), dt)
    ref, nref = orc.expm_krylov(lambda y: orc.hop1(L, W, R, y.reshape(v.shape)).ravel(), dt, v.ravel())
    assert nmv == nref
    assert _rel(out.cpu().numpy().ravel(), ref) < 1e-11


@pytest.mark.parametrize("shape,dt", [
    ((1, 2, 2, 1, 3), -0.05j / 2),     # full-space exit (n = 4)
    ((16, 8, 16, 3, 3), -0.05j / 2),
    ((16, 8, 16, 3, 3), -3.0j),        # long Krylov space: several host chunks, Taylor substeps
    ((16, 8, 16, 3, 3), -0.4),         # imaginary time (real dt)
    ((64, 10, 64, 3, 3), -0.05j / 2),
])
def test_expm_krylov_async_matches_sync_and_oracle(dev, shape, dt):
    """The asynchronous entry point runs the same kernels as the synchronous
    one (bit-identical vector, same count); both match the oracle."""
    print(f"[krylov async {shape} dt={dt}]", flush=True)
    Ml, d, Mr, Dl, Dr = shape
    rng = np.random.default_rng(23 + sum(shape))
    L, W, R = _herm_env_inputs(rng, *shape)
    v = _rand(rng, (Ml, d, Mr))
    args = (_t(L, dev), _t(W, dev), _t(R, dev), _t(v, dev))
    out, nmv = native.expm_krylov(*args, dt)
    ref, nref = orc.expm_krylov(lambda y: orc.hop1(L, W, R, y.reshape(v.shape)).ravel(), dt, v.ravel())
    assert nmv == nref
    assert _rel(out.cpu().numpy().ravel(), ref) < 1e-11
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    out_a = native.expm_krylov_async(*args, dt, native.PREC_FP64, nmv + 5, info)
    assert info.cpu().tolist() == [nmv, 1]
    assert torch.equal(out_a, out)
    if nmv > 8:
        # budget too small: reported, not silently wrong
        info.zero_()
        native.expm_krylov_async(*args, dt, native.PREC_FP64, 6, info)
        assert info.cpu().tolist() == [6, 0]


def test_expm_krylov_breakdown(dev):
    """H_eff = identity: beta_0 = 0, breakdown after one matvec, exp(dt) v."""
    print("[krylov breakdown]", flush=True)
    Ml, d, Mr = 6, 3, 5
    L = np.eye(Ml).reshape(Ml, 1, Ml).astype(complex)
    R = np.eye(Mr).reshape(Mr, 1, Mr).astype(complex)
    W = np.eye(d).reshape(1, d, d, 1)
    rng = np.random.default_rng(2)
    v = _rand(rng, (Ml, d, Mr))
    dt = -0.3j
    args = (_t(L, dev), _t(W, dev), _t(R, dev), _t(v, dev))
    out, nmv = native.expm_krylov(*args, dt)
    ref, nref = orc.expm_krylov(lambda y: orc.hop1(L, W, R, y.reshape(v.shape)).ravel(), dt, v.ravel())
    assert nmv == nref == 1
    np.testing.assert_allclose(out.cpu().numpy(), np.exp(dt) * v, rtol=1e-13, atol=1e-14)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    out_a = native.expm_krylov_async(*args, dt, native.PREC_FP64, 12, info)
    assert info.cpu().tolist() == [1, 1]
    assert torch.equal(out_a, out)


def test_expm_krylov_bond(dev):
    print("[krylov bond]", flush=True)
    rng = np.random.default_rng(5)
    L, _, R = _herm_env_inputs(rng, 24, 1, 20, 3, 3)
    s = _rand(rng, (24, 20))
    dt = 0.05j / 2
    out, nmv = native.expm_krylov(_t(L, dev), None, _t(R, dev), _t(s, dev), dt)
    ref, nref = orc.expm_krylov(lambda y: orc.hop0(L, R, y.reshape(s.shape)).ravel(), dt, s.ravel())
    assert nmv == nref
    assert _rel(out.cpu().numpy().ravel(), ref) < 1e-11


@pytest.mark.parametrize("system", ["L", "R"])
@pytest.mark.parametrize("rows,cols,nblk", [(20, 8, 1), (8, 20, 1), (64, 32, 2), (30, 30, 2), (6, 3, 1), (3, 6, 1),
                                            (2560, 256, 1), (256, 2560, 1), (640, 64, 2), (1, 7, 1), (7, 1, 1),
                                            (3100, 40, 1), (40, 3100, 1), (363, 3630, 1), (3630, 363, 1),
                                            (300, 2800, 2), (2560, 37, 1), (1100, 45, 1), (45, 1900, 1)])
def test_block_qr(dev, system, rows, cols, nblk):
    print(f"[block_qr {system} {rows}x{cols} blocks={nblk}]", flush=True)
    rng = np.random.default_rng(rows * 31 + cols + nblk)
    C = _rand(rng, (rows, cols))
    if nblk == 1:
        rq = np.zeros((rows, 1), dtype=int)
        cq = np.zeros((cols, 1), dtype=int)
    else:
        rq = rng.integers(0, 2, (rows, 1))
        cq = 1 - rng.integers(0, 2, (cols, 1))
    qntot = np.array([0]) if nblk == 1 else np.array([1])
    blocks = orc.qn_blocks(rq, cq, qntot)
    mask = np.zeros((rows, cols), dtype=bool)
    for _, _, ls, rs in blocks:
        mask[np.ix_(ls, rs)] = True
    C = np.where(mask, C, 0)
    u, v = native.block_qr(system, _t(C, dev), [(ls, rs) for _, _, ls, rs in blocks])
    u = u.cpu().numpy()
    v = v.cpu().numpy()
    uo, _, vo, _ = orc.svd_qn(C, rq, cq, qntot, QR=True, system=system, full_matrices=False)
    assert u.shape == uo.shape and v.shape == vo.shape
    # reconstruction and isometry
    assert _rel(u @ v.T, C) < 1e-13
    iso = u if system == "L" else v
    np.testing.assert_allclose(iso.conj().T @ iso, np.eye(iso.shape[1]), atol=1e-13)
    # equal to scipy's factors up to a sign per column (Householder, real diagonal)
    for k in range(min(u.shape[1], 64)):
        j = np.argmax(np.abs(uo[:, k]))
        ph = u[j, k] / uo[j, k]
        assert abs(abs(ph) - 1) < 1e-10
        np.testing.assert_allclose(u[:, k], ph * uo[:, k], atol=1e-11)
        np.testing.assert_allclose(v[:, k], vo[:, k] / ph, atol=1e-11)


@pytest.mark.parametrize("system", ["L", "R"])
def test_block_qr_rank_deficient(dev, system):
    """Exactly zero and repeated columns: still an isometry and an exact
    reconstruction (Householder reflectors with tau = 0)."""
    print(f"[block_qr rank deficient {system}]", flush=True)
    rng = np.random.default_rng(9)
    C = _rand(rng, (300, 40))
    C[:, 5] = 0
    C[:, 7] = C[:, 3]
    C[:, 20:24] = 1e-17 * C[:, 30:34]
    if system == "R":
        C = C.T.copy()
    rows, cols = C.shape
    blocks = [(np.arange(rows), np.arange(cols))]
    u, v = native.block_qr(system, _t(C, dev), blocks)
    u = u.cpu().numpy()
    v = v.cpu().numpy()
    assert np.linalg.norm(u @ v.T - C) / np.linalg.norm(C) < 1e-14
    iso = u if system == "L" else v
    np.testing.assert_allclose(iso.conj().T @ iso, np.eye(iso.shape[1]), atol=1e-13)
    # v^T layout option gives the transpose of the same factor
    u2, vt = native.block_qr(system, _t(C, dev), blocks, v_transposed=True)
    np.testing.assert_array_equal(vt.cpu().numpy(), v.T)
