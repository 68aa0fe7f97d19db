"""DMRG (optimize_mps) on the device: pins and parity.

What pins the local eigenproblem and the sweep (the reference holds no DMRG
golden vectors, only known energies):
* the H_eff diagonal / dense kernels against the brute-force H_eff;
* the reference's known answers on its 3-molecule Holstein test model
  (renormalizer/tests/parameter.py:8-30): ground state 0.08401412 + gs_zpe
  (test_gs.py:21, rel 1e-5) with the one- and two-site algorithms, and the
  four lowest states [0.08401412, 0.08449771, 0.08449801, 0.08449945] +
  gs_zpe with the state-averaged algorithm (test_gs.py:48-61);
* exact diagonalisation of small Holstein / spin-boson chains (oracle.dense).
CPU tests: basis selection and the random initial MPS.
"""
import numpy as np
import pytest

from oracle import dense
from renormalizer_amd import (Mol, Mps, Mpo, HolsteinModel, Phonon, Quantity, construct_mps_mpo,
                              optimize_mps)
from renormalizer_amd.mps.svd_qn import add_outer, get_qn_mask, select_basis
from renormalizer_amd.utils import constant
from tests import helpers


reference_holstein_model = helpers.reference_holstein_model


PROCEDURE = [[10, 0.4], [20, 0.2], [30, 0.1], [40, 0], [40, 0]]


# ---------------------------------------------------------------- CPU


def test_select_basis_percent():
    """percent of the kept states shared among the sectors, largest weights
    first inside each, then the largest weights overall (lib.py:253-322)."""
    s = np.array([0.9, 0.1, 0.5, 0.4, 0.3, 0.05])
    qn = [(0,), (1,), (0,), (1,), (2,), (2,)]
    sidx, qns, w = select_basis(None, s, qn, None, 4, percent=0)
    assert sidx == [0, 2, 3, 4]
    # percent 1: int(4 * 1 / 3) = 1 from every sector, then one more by weight
    sidx, qns, w = select_basis(None, s, qn, None, 4, percent=1.0)
    assert sorted(sidx[:3]) == [0, 3, 4] and sidx[3] == 2
    assert [tuple(q) for q in qns] == [qn[i] for i in sidx]


def test_random_mps_quantum_numbers():
    """Random initial MPS: left-canonical, normalised, every site tensor zero
    outside its quantum-number blocks, total quantum number as requested."""
    np.random.seed(7)
    model = helpers.holstein_model(3, 3)
    mps = Mps.random(model, 1, 6, percent=1)
    assert mps.is_left_canonical and mps.check_left_canonical()
    assert abs(np.linalg.norm(mps._mp[-1]) - 1) < 1e-12
    for i, site in enumerate(mps._mp):
        # qn[k] holds left-block numbers for k <= qnidx, right-block numbers beyond
        right = mps.qntot - mps.qn[i + 1] if i + 1 <= mps.qnidx else mps.qn[i + 1]
        qnmat = add_outer(add_outer(mps.qn[i], np.array(model.basis[i].sigmaqn)), right)
        mask = get_qn_mask(qnmat, mps.qntot)
        assert np.all(site[~mask] == 0)
    assert max(mps.bond_dims) <= 6


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(3, 4, 5, 2, 3), (8, 6, 7, 3, 4)])
def test_heff_kernels(dev, shape):
    import torch

    from renormalizer_amd import native
    from tests.test_oracle import _dense_heff
    print(f"[heff {shape}]", flush=True)
    Ml, d, Mr, Dl, Dr = shape
    rng = np.random.default_rng(sum(shape))
    L = rng.standard_normal((Ml, Dl, Ml)) + 1j * rng.standard_normal((Ml, Dl, Ml))
    R = rng.standard_normal((Mr, Dr, Mr)) + 1j * rng.standard_normal((Mr, Dr, Mr))
    W = rng.standard_normal((Dl, d, d, Dr))
    h = _dense_heff(L, W, R)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    diag = native.heff_diag(t(L), t(W), t(R), (Ml, d, Mr)).cpu().numpy().ravel()
    np.testing.assert_allclose(diag, np.diag(h), rtol=1e-12, atol=1e-12)
    idx = np.sort(rng.choice(Ml * d * Mr, size=min(40, Ml * d * Mr), replace=False)).astype(np.int64)
    sub = native.heff_dense(t(L), t(W), t(R), (Ml, d, Mr), t(idx)).cpu().numpy()
    np.testing.assert_allclose(sub, h[np.ix_(idx, idx)], rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["1site", "2site"])
def test_reference_ground_state(dev, method):
    """Reference test_gs.py:24-36: ground-state energy of the 3-molecule
    Holstein model within rel 1e-5 of 0.08401412 + gs_zpe."""
    print(f"[dmrg gs {method}]", flush=True)
    np.random.seed(2)
    model = reference_holstein_model()
    gs_e = 0.08401412 + model.gs_zpe
    mps, mpo = construct_mps_mpo(model, PROCEDURE[0][0], 1)
    mps.optimize_config.procedure = PROCEDURE
    mps.optimize_config.method = method
    energies, mps_opt = optimize_mps(mps.copy(), mpo)
    assert energies[-1] == pytest.approx(gs_e, rel=1e-5)
    assert mps_opt.expectation(mpo) == pytest.approx(gs_e, rel=1e-5)
    assert abs(mps_opt.mp_norm - 1) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["1site", "2site"])
def test_reference_multistate(dev, method):
    """Reference test_gs.py:39-61: four lowest states by the state-averaged
    algorithm."""
    print(f"[dmrg multistate {method}]", flush=True)
    np.random.seed(3)
    model = reference_holstein_model()
    mps, mpo = construct_mps_mpo(model, PROCEDURE[0][0], 1)
    mps.optimize_config.procedure = PROCEDURE
    mps.optimize_config.nroots = 4
    mps.optimize_config.method = method
    mps.optimize_config.e_atol = 1e-6
    mps.optimize_config.e_rtol = 1e-6
    energy, mpss = optimize_mps(mps, mpo)
    energy_std = np.array([0.08401412, 0.08449771, 0.08449801, 0.08449945]) + model.gs_zpe
    assert np.allclose(energy[-1], energy_std)
    assert np.allclose([m.expectation(mpo) for m in mpss], energy_std)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["1site", "2site"])
@pytest.mark.parametrize("nroots", [1, 4])
def test_reference_omega_target(dev, method, nroots):
    """Reference test_gs.py:63-85 (test_ex): omega = 0.084 targets the states
    nearest 0.084 through (H - omega)^2; <H> of the result equals the
    reference's known energies (np.allclose, as the reference asserts)."""
    print(f"[dmrg omega {method} nroots={nroots}]", flush=True)
    np.random.seed(5)
    model = reference_holstein_model()
    mps, mpo = construct_mps_mpo(model, PROCEDURE[0][0], 1)
    mps.optimize_config.procedure = PROCEDURE
    mps.optimize_config.nroots = nroots
    mps.optimize_config.method = method
    mps.optimize_config.e_atol = 1e-6
    mps.optimize_config.e_rtol = 1e-6
    energy, res = optimize_mps(mps, mpo, omega=0.084)
    energy_std = np.array([0.08401412, 0.08449771, 0.08449801, 0.08449945]) + model.gs_zpe
    if nroots == 1:
        assert np.allclose(res.expectation(mpo), energy_std[0])
    else:
        assert np.allclose([m.expectation(mpo) for m in res], energy_std)


def test_folded_mpo_dense():
    """The merged two-layer MPO of the omega target equals (H - omega)^2 as a
    dense matrix (Kronecker-product Hamiltonian of a 2-molecule Holstein
    chain, oracle.dense)."""
    from renormalizer_amd.mps.gs import folded_mpo
    model = helpers.holstein_model(2, 3, g=0.8)
    mpo = Mpo(model)
    omega = 0.37
    f = folded_mpo(mpo, omega)
    dis = 0.8 * np.sqrt(2.0)
    jm = np.array([[0.0, -1.0], [-1.0, 0.0]])
    h = dense.holstein_dense([0.0] * 2, jm, [1.0] * 2, [dis] * 2, 3)
    np.testing.assert_allclose(mpo.todense(), h, atol=1e-12)
    shifted = h - omega * np.eye(h.shape[0])
    np.testing.assert_allclose(f.todense(), shifted @ shifted, atol=1e-12)
    assert max(f.bond_dims) == (max(mpo.bond_dims) + 1) ** 2


def _ed_holstein(model_args):
    nmol, nlev, j, g = model_args
    omega = 1.0
    dis = g * np.sqrt(2.0 / omega)
    jm = np.zeros((nmol, nmol))
    for i in range(nmol - 1):
        jm[i, i + 1] = jm[i + 1, i] = j
    h = dense.holstein_dense([0.0] * nmol, jm, [omega] * nmol, [dis] * nmol, nlev)
    # one-exciton sector: electron occupations sum to 1
    dims = [2, nlev] * nmol
    occ = np.zeros(h.shape[0])
    for k, idx in enumerate(np.ndindex(*dims)):
        occ[k] = sum(idx[2 * i] for i in range(nmol))
    sel = np.where(occ == 1)[0]
    return np.linalg.eigvalsh(h[np.ix_(sel, sel)])


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["1site", "2site"])
def test_ed_holstein_chain(dev, method):
    """4-molecule Holstein chain (4 levels): DMRG ground state and the
    three lowest states against exact diagonalisation of the one-exciton
    sector."""
    print(f"[dmrg ed holstein {method}]", flush=True)
    np.random.seed(5)
    model = helpers.holstein_model(4, 4, j=-1.0, g=0.7)
    ed = _ed_holstein((4, 4, -1.0, 0.7))
    mps, mpo = construct_mps_mpo(model, 10, 1)
    mps.optimize_config.procedure = [[10, 0.4], [20, 0.2], [40, 0.1], [64, 0], [64, 0], [64, 0]]
    mps.optimize_config.method = method
    energies, mps_opt = optimize_mps(mps.copy(), mpo)
    assert energies[-1] == pytest.approx(ed[0], abs=1e-7)
    mps.optimize_config.nroots = 3
    mps.optimize_config.e_atol = 1e-8
    energies, _ = optimize_mps(mps, mpo)
    np.testing.assert_allclose(energies[-1], ed[:3], atol=1e-6)


@pytest.mark.gpu
def test_ed_spin_boson_direct_and_davidson(dev):
    """Spin + 3 modes (no quantum numbers): every site small enough for the
    direct eigensolver with 1site, Davidson with a forced algo; both give
    the exact ground-state energy."""
    print("[dmrg ed sbm]", flush=True)
    np.random.seed(11)
    model = helpers.sbm_model(3, 5, alpha=0.4, omega_c=4.0, eps=0.3)
    mpo = Mpo(model)
    ed = np.linalg.eigvalsh(mpo.todense())[0]
    for algo in ("davidson", "direct"):
        mps = Mps.random(model, 0, 12, percent=1)
        mps.optimize_config.procedure = [[12, 0.3], [30, 0.1], [60, 0], [60, 0], [60, 0]]
        mps.optimize_config.method = "2site"
        mps.optimize_config.algo = algo
        energies, m = optimize_mps(mps, mpo)
        assert energies[-1] == pytest.approx(ed, abs=1e-8), algo
