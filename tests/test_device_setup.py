"""The initial-state operations on device tensors (device_setup.py: direct
sum, MPO application, QR canonicalisation and SVD truncation sweeps, and
svd_qn_device_svd's rotation of the QR factors onto singular vectors),
checked against the host (numpy/LAPACK) versions of the same operations in
mp.py and against dense state vectors.

Each test runs twice: "stand-in" on CPU torch tensors with the native entry
points the code calls (reno_zgemm, reno_block_qr, reno_block_eigh, and
reno_env_update for the norms) replaced by numpy/LAPACK versions of the
same contracts, which checks the host logic without a GPU; "device" on the
GPU through the library itself."""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import tdvp as orc
from renormalizer_amd import CompressConfig, CompressCriteria, EvolveConfig, EvolveMethod, Mpo, Mps, native
from tests import helpers

CPU = torch.device("cpu")


def _zgemm(a, b):
    return a @ b


def _block_qr(system, coef, blocks, v_transposed=False):
    a = coef.numpy()
    rows, cols = a.shape
    kt = sum(min(len(r), len(c)) for r, c in blocks)
    u = np.zeros((rows, kt), dtype=complex)
    vt = np.zeros((kt, cols), dtype=complex)
    off = 0
    for r, c in blocks:
        sub = a[np.ix_(r, c)]
        k = min(sub.shape)
        if system == "L":
            x, y = scipy.linalg.qr(sub, mode="economic")
        else:
            x, y = scipy.linalg.rq(sub, mode="economic")
        u[r, off:off + k] = x
        vt[off:off + k][:, c] = y
        off += k
    assert v_transposed
    return torch.from_numpy(u), torch.from_numpy(vt)


def _block_eigh(dm, sectors):
    a = dm.numpy()
    n = a.shape[0]
    ut = sum(len(r) for r in sectors)
    u = np.zeros((n, ut), dtype=complex)
    w = np.zeros(ut)
    off = 0
    for r in sectors:
        sub = a[np.ix_(r, r)]
        ww, vv = scipy.linalg.eigh((sub + sub.conj().T) / 2)
        u[r, off:off + len(r)] = vv
        w[off:off + len(r)] = ww
        off += len(r)
    return torch.from_numpy(u), torch.from_numpy(w)


def _env_update(env, ms, w, domain, ms_conj=None):
    conj = None if ms_conj is None else ms_conj.numpy()
    return torch.from_numpy(orc.contract_one_site(env.numpy(), ms.numpy(), w.numpy().astype(complex), domain,
                                                  conj))


@pytest.fixture(params=["stand-in", pytest.param("device", marks=pytest.mark.gpu)])
def target(request, monkeypatch):
    """The torch device the state is moved to."""
    if request.param == "device":
        return torch.device("cuda", 0)
    monkeypatch.setattr(native, "zgemm", _zgemm)
    monkeypatch.setattr(native, "env_update", _env_update)
    monkeypatch.setattr(native, "block_qr", _block_qr)
    monkeypatch.setattr(native, "block_eigh", _block_eigh)
    return CPU


def dense(mps):
    s = mps.numpy_sites()
    v = s[0].reshape(-1, s[0].shape[-1])
    for t in s[1:]:
        v = (v @ t.reshape(t.shape[0], -1)).reshape(-1, t.shape[-1])
    return v.reshape(-1)


def moved(mps, target):
    new = mps.copy()
    new.to_device(target)
    return new


def same_qn(a, b):
    return (len(a.qn) == len(b.qn) and all(np.array_equal(np.array(x), np.array(y)) for x, y in zip(a.qn, b.qn))
            and a.qnidx == b.qnidx and a.to_right == b.to_right)


def _random_holstein(seed, m=12):
    np.random.seed(seed)
    model = helpers.holstein_model(4, 3)
    mps = Mps.random(model, 1, m)
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=6)
    return model, mps


def test_canonicalise_matches_host(target):
    _, mps = _random_holstein(3)
    h = mps.copy().canonicalise()
    d = moved(mps, target).canonicalise()
    assert same_qn(h, d)
    np.testing.assert_allclose(dense(d), dense(mps), atol=1e-13)
    for i in range(1, len(d)):  # right-canonical sites
        m = d.numpy_sites()[i].reshape(d[i].shape[0], -1)
        np.testing.assert_allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-12)


@pytest.mark.parametrize("through_left", [True, False])
def test_compress_matches_host(target, through_left):
    """Truncation 12 -> 6 on every bond: singular values, kept quantum
    numbers and the truncated state equal the LAPACK-SVD host sweep (the
    random state's spectra are non-degenerate)."""
    _, mps = _random_holstein(5)
    if through_left:  # right-canonical start, sweep to the right (system "L")
        mps.canonicalise()
    h, sh = mps.copy().compress(ret_s=True)
    d, sd = moved(mps, target).compress(ret_s=True)
    assert h.bond_dims == d.bond_dims
    assert same_qn(h, d)
    np.testing.assert_allclose(sd, sh, atol=1e-12)
    np.testing.assert_allclose(dense(d), dense(h), atol=1e-10)


def test_apply_and_add_dense(target):
    model, mps = _random_holstein(7)
    mpo = Mpo(model)
    ah = mpo @ mps
    ad = mpo @ moved(mps, target)
    assert same_qn(ah, ad)
    np.testing.assert_allclose(dense(ad), dense(ah), atol=1e-12)
    sh = mps + ah
    sd = moved(mps, target) + ad
    assert same_qn(sh, sd)
    np.testing.assert_allclose(dense(sd), dense(mps) + dense(ah), atol=1e-12)


def test_expand_matches_host(target):
    """expand_bond_dimension of a spin-boson ground state with a visible
    admixture (1e-3): same bond dimensions and quantum numbers, the same
    state to 1e-9."""
    model = helpers.sbm_model(6, 4)
    mpo = Mpo(model)

    def build(dev):
        s = Mps.ground_state(model, False)
        s.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=12)
        s.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        if dev:
            s.to_device(target)
        return s.expand_bond_dimension(mpo, coef=1e-3, include_ex=False)

    h, d = build(False), build(True)
    assert h.bond_dims == d.bond_dims
    assert same_qn(h, d)
    vh, vd = dense(h), dense(d)
    fid = abs(np.vdot(vh, vd)) / np.linalg.norm(vh) / np.linalg.norm(vd)
    assert fid > 1 - 1e-12, fid


@pytest.mark.gpu
def test_eigh_low_rank_gram():
    """Regression: an 81 x 81 Gram matrix of numerical rank about 5 from the
    device expansion (tests/golden/make_gram_low_rank.py), for which
    rocSOLVER's divide-and-conquer eigensolver returned NaN eigenvectors
    with info 0.  reno_block_eigh gives finite eigenpairs that reconstruct
    it, orthonormal vectors and LAPACK's eigenvalues."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "gram_low_rank_81.npy"))
    n = g.shape[0]
    dev = torch.device("cuda", 0)
    u, w = native.block_eigh(torch.from_numpy(g.astype(complex)).to(dev), [np.arange(n, dtype=np.int64)])
    u = u.cpu().numpy()
    w = w.cpu().numpy()
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))
    np.testing.assert_allclose(w, np.linalg.eigvalsh(g), atol=1e-15)
    np.testing.assert_allclose(u.conj().T @ u, np.eye(n), atol=1e-13)
    np.testing.assert_allclose((u * w) @ u.conj().T, g, atol=1e-15)


@pytest.mark.gpu
def test_expand_c4_device():
    """The benchmark start built on the device (spin + 64 modes x 10 levels,
    M = 256, admixture 1e-16): the host path's bond dimensions, a
    normalised state in canonical form, and <sigma_z> over 5 FP64 steps within
    the host path's own spread (tests/golden/make_expansion_c4.py: gesdd
    against gesvd for the truncation SVDs, 1.2e-9 after one step and 1.9e-7
    after five; the 1e-16 admixture directions are fixed by rounding, so
    no two builds agree more closely): at most 10 times that spread, and
    1e-9, per step."""
    import json
    import os
    import time
    from tests.golden.make_expansion_c4 import DT, M, STEPS, model_and_mpo, sigma_z
    with open(os.path.join(os.path.dirname(__file__), "golden", "expansion_c4.json")) as f:
        ref = json.load(f)
    model, mpo = model_and_mpo()
    mps = Mps.ground_state(model, False)
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=M)
    mps.to_device()
    torch.cuda.synchronize()
    t = time.time()
    mps = mps.expand_bond_dimension(mpo, coef=1e-16, include_ex=False)
    torch.cuda.synchronize()
    print(f"[device expansion of the M = {M} spin-boson start: {time.time() - t:.1f} s]", flush=True)
    assert mps.bond_dims == ref["bond_dims"]
    assert all(np.all(np.array(q) == 0) for q in mps.qn)
    assert abs(mps.mp_norm - 1) < 1e-12
    for i in (0, 1, 2, 32, 63, 64):
        if i == mps.qnidx:
            continue
        t = mps[i]
        m = t.reshape(-1, t.shape[-1]) if i < mps.qnidx else t.reshape(t.shape[0], -1).T
        g = m.conj().T @ m if i < mps.qnidx else m.T @ m.conj()
        eye = torch.eye(g.shape[0], dtype=g.dtype, device=g.device)
        assert (g - eye).abs().max().item() < 1e-12, i
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    zs = []
    for _ in range(STEPS):
        mps = mps.evolve(mpo, DT)
        zs.append(sigma_z(mps))
    zs = np.array(zs)
    a, b = np.array(ref["sigma_z_gesdd"]), np.array(ref["sigma_z_gesvd"])
    bound = np.maximum(10 * np.abs(a - b), 1e-9)
    print("device - host(gesdd):", np.abs(zs - a).tolist(), "bound", bound.tolist(), flush=True)
    assert np.all(np.abs(zs - a) <= bound)


def qn_conserved(mps):
    """Every nonzero entry A_i[l, d, r] of every site joins bond labels with
    qn_l + sigma_d = qn_r (left-block numbers on every bond)."""
    m = mps.copy()
    m.move_qnidx(len(m) - 1)
    sites = m.numpy_sites()
    for i, a in enumerate(sites):
        ql = np.array(m.qn[i]).reshape(a.shape[0], -1)
        # the last bond is stored as zeros; its left-block number is qntot
        qr = (np.array(m.qn[i + 1]) if i + 1 < len(m) else np.array([m.qntot])).reshape(a.shape[2], -1)
        sig = np.array(m._get_sigmaqn(i)).reshape(a.shape[1], -1)
        big = np.abs(a) > 1e-14 * np.abs(a).max()
        for l, dd, r in zip(*np.nonzero(big)):
            if not np.array_equal(ql[l] + sig[dd], qr[r]):
                return False
    return True


def test_expand_holstein_matches_host(target):
    """A quantum-number-conserving expansion (one electron on a 4-site
    Holstein chain, several sectors per bond, include_ex=True): same bond
    dimensions and the same state as the host path, with bond quantum
    numbers that the site tensors conserve.  The labels themselves are not
    compared: where the fixed bond dimension exceeds a bond's rank, the
    states kept beyond it have zero weight and their sectors (like the
    order of degenerate singular values) are set by rounding."""
    model = helpers.holstein_model(4, 3)
    mpo = Mpo(model)

    def build(dev):
        gs = Mps.ground_state(model, False)
        s = Mpo.onsite(model, r"a^\dagger", dof_set={1}) @ gs
        s.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=8)
        s.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        if dev:
            s.to_device(target)
        return s.expand_bond_dimension(mpo, coef=1e-3)

    h, d = build(False), build(True)
    assert h.bond_dims == d.bond_dims
    assert qn_conserved(h) and qn_conserved(d)
    vh, vd = dense(h), dense(d)
    fid = abs(np.vdot(vh, vd)) / np.linalg.norm(vh) / np.linalg.norm(vd)
    assert fid > 1 - 1e-12, fid
