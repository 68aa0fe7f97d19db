"""Parity of the device TDVP-PS time step (Mps.evolve, EvolveMethod.tdvp_ps)
with the numpy oracle on identical initial tensors.

FP64 mode: observables agree with the complex128 oracle to 1e-8 (absolute,
for O(1) populations / <sigma_z>), the Lanczos matvec counts of every local
step are identical, and the quantum-number bookkeeping (qn, qnidx) is
bit-identical.  Mixed mode (FP32 matvecs in the tail of each Lanczos
iteration): observables within 1e-8 of the FP64 oracle; the all-FP32 mode
within 1e-5 on these short runs.  Reference integration targets are re-checked through the
device path against exact diagonalisation.
"""
import numpy as np
import pytest

from oracle import dense
from oracle import tdvp as orc
from renormalizer_amd import EvolveConfig, EvolveMethod
from tests import helpers

pytestmark = pytest.mark.gpu


def _assert_steps_match(sd, so):
    """Lanczos matvec counts per local step.  The stopping test is
    allclose(res, new_res) on every element (krylov.py:76-81); when an
    element sits at the threshold, rounding differences between the device
    and numpy can flip one decision by one check interval (2 matvecs) without
    changing the result beyond the Krylov tolerance.  Require identical
    counts on >= 97% of the local steps and differences of exactly 2."""
    flat_d = [x for s in sd for x in s]
    flat_o = [x for s in so for x in s]
    assert len(flat_d) == len(flat_o)
    diff = [(a, b) for a, b in zip(flat_d, flat_o) if a != b]
    assert len(diff) <= 0.03 * len(flat_d), diff
    assert all(abs(a - b) == 2 for a, b in diff), diff


def _sz_dev(mps):
    r = mps.calc_1site_rdm(0)[0]
    return (r[0, 0] - r[1, 1]).real


def _sz_orc(st):
    r = orc.calc_1site_rdm(st, 0)
    return (r[0, 0] - r[1, 1]).real


def _run_pair(mps, mpo, dt, nsteps, observe_dev, observe_orc, precision="fp64"):
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    dev_mps = mps.copy()
    dev_mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, precision=precision)
    dev_mps.to_device()
    obs_d, obs_o, steps_d, steps_o = [observe_dev(dev_mps)], [observe_orc(st)], [], []
    for _ in range(nsteps):
        dev_mps = dev_mps.evolve(mpo, dt)
        st, ls = orc.tdvp_ps(st, w, dt)
        obs_d.append(observe_dev(dev_mps))
        obs_o.append(observe_orc(st))
        steps_d.append(list(dev_mps.evolve_config.local_steps))
        steps_o.append(ls)
    return np.array(obs_d), np.array(obs_o), steps_d, steps_o, dev_mps, st


def test_spin_boson_fp64_parity():
    print("[tdvp sbm fp64 6x4 M=8]", flush=True)
    model = helpers.sbm_model(6, 4, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 8)
    od, oo, sd, so, dmps, st = _run_pair(mps, mpo, 0.05, 10, _sz_dev, _sz_orc)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    _assert_steps_match(sd, so)
    assert [q.tolist() for q in dmps.qn] == [np.array(q).tolist() for q in st.qn]
    assert dmps.qnidx == st.qnidx and dmps.to_right == st.to_right
    assert abs(dmps.mp_norm - 1) < 1e-12


def test_holstein_qn_fp64_parity():
    print("[tdvp holstein qn fp64 4x4 M=12]", flush=True)
    model = helpers.holstein_model(4, 4, j=-1.0, g=1.0)
    mps, mpo = helpers.holstein_initial(model, 12)
    obs_d = lambda m: m.e_occupations
    obs_o = lambda s: helpers.e_occupations_oracle(s, model)
    od, oo, sd, so, dmps, st = _run_pair(mps, mpo, 0.02, 10, obs_d, obs_o)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    np.testing.assert_allclose(od.sum(axis=1), 1.0, atol=1e-10)
    _assert_steps_match(sd, so)
    assert [np.array(q).tolist() for q in dmps.qn] == [np.array(q).tolist() for q in st.qn]
    # exact zeros outside the allowed blocks (one electron)
    def left_numbers(b):  # L-block numbers of bond b
        q = np.array(dmps.qn[b])[:, 0]
        return q if b <= dmps.qnidx else dmps.qntot[0] - q

    for i, site in enumerate(dmps.numpy_sites()):
        ql, qr = left_numbers(i), left_numbers(i + 1)
        sig = model.basis[i].sigmaqn[:, 0]
        allowed = (ql[:, None, None] + sig[None, :, None]) == qr[None, None, :]
        assert np.all(site[~allowed] == 0)
        assert np.any(site[allowed] != 0)


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-8), ("fp32", 1e-5)])
def test_spin_boson_mixed(precision, tol):
    print(f"[tdvp sbm {precision} 6x4 M=8]", flush=True)
    model = helpers.sbm_model(6, 4, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 8)
    od, oo, sd, so, *_ = _run_pair(mps, mpo, 0.05, 10, _sz_dev, _sz_orc, precision=precision)
    np.testing.assert_allclose(od, oo, atol=tol, rtol=0)
    if precision == "mixed":
        _assert_steps_match(sd, so)


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-8), ("fp32", 1e-5)])
def test_holstein_mixed(precision, tol):
    print(f"[tdvp holstein {precision} 4x4 M=12]", flush=True)
    model = helpers.holstein_model(4, 4, j=-1.0, g=1.0)
    mps, mpo = helpers.holstein_initial(model, 12)
    obs_d = lambda m: m.e_occupations
    obs_o = lambda s: helpers.e_occupations_oracle(s, model)
    od, oo, sd, so, *_ = _run_pair(mps, mpo, 0.02, 10, obs_d, obs_o, precision=precision)
    np.testing.assert_allclose(od, oo, atol=tol, rtol=0)
    if precision == "mixed":
        _assert_steps_match(sd, so)


def _device_pair(mps, mpo, dt, nsteps, observe, modes=("fp64", "mixed"), tau=1e-2):
    """The same start evolved on the device in several precision modes:
    observables, matvec counts and FP32 matvecs per step."""
    out = {}
    for mode in modes:
        m = mps.copy()
        m.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, precision=mode, mixed_tau=tau)
        m.to_device()
        obs, steps, low = [observe(m)], [], []
        for _ in range(nsteps):
            m = m.evolve(mpo, dt)
            obs.append(observe(m))
            steps.append(sum(m.evolve_config.local_steps))
            low.append(getattr(m.evolve_config, "fp32_matvecs", 0))
        out[mode] = (np.array(obs), np.array(steps), np.array(low), m)
    return out


def test_mixed_tail_c2_trajectory():
    """The mixed mode on the C2 workload (Holstein, 16 molecules, d = 10,
    M = 64, dt = 0.02), 40 steps: populations within 1e-8 of the FP64 device
    run, total matvec counts within 2 % per step, and a large share of the
    matvecs on FP32 matrix cores once the switch points are known (from
    the third step on)."""
    print("[tdvp mixed C2 holstein 32 sites M=64, 40 steps]", flush=True)
    model = helpers.holstein_model(16, 10, j=-1.0, g=1.0)
    mps, mpo = helpers.holstein_initial(model, 64)
    r = _device_pair(mps, mpo, 0.02, 40, lambda m: m.e_occupations)
    (o64, s64, _, _), (omx, smx, low, _) = r["fp64"], r["mixed"]
    err = np.abs(omx - o64).max()
    print(f"[mixed C2] max |dP| {err:.2e}, matvecs {s64.sum()} vs {smx.sum()}, "
          f"fp32 share {low[2:].sum() / smx[2:].sum():.3f}", flush=True)
    assert err < 1e-8
    np.testing.assert_allclose(smx, s64, rtol=0.02)
    assert low[2:].sum() > 0.3 * smx[2:].sum()


def test_reference_holstein_target_device():
    """reference test_evolve.py:123-131 through the device path vs exact
    diagonalisation: mean deviation of the populations < 1e-4."""
    print("[tdvp reference 3-site holstein]", flush=True)
    from renormalizer_amd import HolsteinModel, Mol, Phonon, Quantity
    ph = Phonon.simple_phonon(Quantity(1), Quantity(1), 2)
    model = HolsteinModel([Mol(Quantity(0), [ph])] * 3, Quantity(1), 3)
    mps, mpo = helpers.holstein_initial(model, 32, center=0)
    h = mpo.todense()
    st = helpers.to_state(mps)
    res = np.ones((1, 1))
    for s in st.sites:
        res = np.tensordot(res, s, axes=1).reshape(-1, s.shape[-1])
    dims = [b.nbas for b in model.basis]
    dt, nstep = 0.4, round(5 / 0.4)
    exact = dense.propagate(h, res[:, 0], [dt * k for k in range(nstep + 1)])
    m = mps.copy()
    m.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    m.to_device()
    got = [m.e_occupations]
    for _ in range(nstep):
        m = m.evolve(mpo, dt)
        got.append(m.e_occupations)
    ref = [[dense.site_rdm(exact[k], dims, 2 * i)[1, 1].real for i in range(3)] for k in range(nstep + 1)]
    assert np.abs(np.array(got) - np.array(ref)).mean() < 1e-4


def test_observables_device_vs_oracle():
    print("[observables]", flush=True)
    model = helpers.holstein_model(3, 3, j=-0.5, g=0.8)
    mps, mpo = helpers.holstein_initial(model, 10)
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    dev = mps.copy().to_device()
    np.testing.assert_allclose(dev.expectation(mpo), orc.expectation(st, w), rtol=1e-12, atol=1e-13)
    for i in range(len(st)):
        np.testing.assert_allclose(dev.calc_1site_rdm(i)[i], orc.calc_1site_rdm(st, i), atol=1e-13)
    np.testing.assert_allclose(dev.e_occupations, helpers.e_occupations_oracle(st, model), atol=1e-13)
    np.testing.assert_allclose(dev.mp_norm, orc.mp_norm(st), rtol=1e-13)


def test_spin_boson_c1_shape_parity():
    """The C1 workload shape (1 spin + 16 modes x 8 levels, M = 16, dt 0.05),
    first 10 steps, FP64 device vs oracle."""
    print("[tdvp C1 17 sites M=16, 10 steps]", flush=True)
    model = helpers.sbm_model(16, 8, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 16)
    od, oo, sd, so, *_ = _run_pair(mps, mpo, 0.05, 10, _sz_dev, _sz_orc)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    _assert_steps_match(sd, so)


def test_async_budget_redo_matches_sync():
    """Asynchronous Krylov calls with budgets that are too small: the step is
    redone with host waits and the state equals the synchronous step."""
    print("[tdvp async redo sbm 6x4 M=8]", flush=True)
    from renormalizer_amd import Mps
    model = helpers.sbm_model(6, 4, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 8)
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    mps.to_device()
    first = mps.evolve(mpo, 0.5)  # synchronous (no budgets yet), big dt: long Krylov spaces
    assert max(first.evolve_config.local_steps) > 12
    sync = first.copy()
    sync.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, krylov_async=False)
    sync = sync.evolve(mpo, 0.5)
    # shrink the recorded budgets so every asynchronous call runs out
    cache = first._env_cache
    first._env_cache = cache[:3] + ((cache[3][0], {k: 1 for k in cache[3][1]}, False, None),) + cache[4:]
    before = Mps.krylov_redo_count
    redone = first.evolve(mpo, 0.5)
    assert Mps.krylov_redo_count == before + 1
    assert redone.evolve_config.local_steps == sync.evolve_config.local_steps
    for a, b in zip(redone.numpy_sites(), sync.numpy_sites()):
        np.testing.assert_allclose(a, b, atol=1e-10)
    # with the budgets of the redone step the next step runs asynchronously
    # (or is redone again if a count grew past its margin): same result
    nxt = redone.evolve(mpo, 0.5)
    ref = sync.evolve(mpo, 0.5)
    assert nxt.evolve_config.local_steps == ref.evolve_config.local_steps
    np.testing.assert_allclose(_sz_dev(nxt), _sz_dev(ref), atol=1e-10)


def test_holstein_c2_shape_parity():
    """The C2 workload shape: Holstein chain of 16 molecules in scheme 2
    (32 sites, electron d = 2, phonon d = 10, omega = 1, g = 1, J = -1),
    Franck-Condon electron on molecule 8, M = 64, dt = 0.02, one-electron
    quantum numbers.  First 3 steps, FP64 device vs oracle: populations to
    1e-8, matvec counts, quantum numbers bit-identical."""
    print("[tdvp C2 holstein 32 sites M=64, 3 steps]", flush=True)
    model = helpers.holstein_model(16, 10, j=-1.0, g=1.0)
    mps, mpo = helpers.holstein_initial(model, 64)
    assert max(mps.bond_dims) == 64
    obs_d = lambda m: m.e_occupations
    obs_o = lambda s: helpers.e_occupations_oracle(s, model)
    od, oo, sd, so, dmps, st = _run_pair(mps, mpo, 0.02, 3, obs_d, obs_o)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    np.testing.assert_allclose(od.sum(axis=1), 1.0, atol=1e-10)
    _assert_steps_match(sd, so)
    assert [np.array(q).tolist() for q in dmps.qn] == [np.array(q).tolist() for q in st.qn]


def test_spin_boson_c5_shape_parity():
    """The C5 workload shape: spin + 32 modes x 8 levels, alpha = 0.4,
    omega_c = 5, M = 64, dt = 0.05.  First 2 steps: FP64 device vs oracle to
    1e-8 with the matvec counts; the mixed mode (complex64 environments, FP32
    matrix cores in the tail of each Lanczos iteration) within 1e-8 of the
    FP64 oracle on <sigma_z>."""
    print("[tdvp C5 sbm 33 sites M=64, 2 steps]", flush=True)
    model = helpers.sbm_model(32, 8, alpha=0.4, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 64)
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    runs = {}
    for precision in ("fp64", "mixed"):
        m = mps.copy()
        m.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, precision=precision)
        m.to_device()
        sz, steps = [_sz_dev(m)], []
        for _ in range(2):
            m = m.evolve(mpo, 0.05)
            sz.append(_sz_dev(m))
            steps.append(list(m.evolve_config.local_steps))
        runs[precision] = (np.array(sz), steps)
    oo, so = [_sz_orc(st)], []
    for _ in range(2):
        st, ls = orc.tdvp_ps(st, w, 0.05)
        oo.append(_sz_orc(st))
        so.append(ls)
    np.testing.assert_allclose(runs["fp64"][0], oo, atol=1e-8, rtol=0)
    _assert_steps_match(runs["fp64"][1], so)
    np.testing.assert_allclose(runs["mixed"][0], oo, atol=1e-8, rtol=0)


def _ps2_pair(mps, mpo, dt, nsteps, m, observe_dev, observe_orc):
    from renormalizer_amd import CompressConfig, CompressCriteria
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    dev = mps.copy()
    dev.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=m)
    dev.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps2)
    dev.to_device()
    od, oo, sd, so = [observe_dev(dev)], [observe_orc(st)], [], []
    for _ in range(nsteps):
        dev = dev.evolve(mpo, dt)
        st, ls = orc.tdvp_ps2(st, w, dt, max_dim=m)
        od.append(observe_dev(dev))
        oo.append(observe_orc(st))
        sd.append(list(dev.evolve_config.local_steps))
        so.append(ls)
    return np.array(od), np.array(oo), sd, so, dev, st


def test_tdvp_ps2_spin_boson_parity():
    """Two-site TDVP-PS (bond dimension grown from 6 and truncated at
    M = 8) on the device vs the oracle: <sigma_z> to 1e-8, the same bond
    dimensions and Lanczos matvec counts.

    The start is the spin-up vacuum expanded along the Hamiltonian with a
    1e-3 admixture: its singular values are well separated, and the oracle
    moves by under 3e-15 when 1e-15 noise is added to it.  From the bare
    product state the two-site steps keep zero-weight SVD completions whose
    directions are fixed by rounding, and the truncated dynamics then moves
    by up to 1e-7 under the same noise, in the oracle alone."""
    print("[tdvp ps2 sbm 6x4 M=8]", flush=True)
    model = helpers.sbm_model(6, 4, alpha=0.5, omega_c=5.0)
    mps, mpo = helpers.sbm_initial(model, 6, coef=1e-3)
    od, oo, sd, so, dmps, st = _ps2_pair(mps, mpo, 0.05, 5, 8, _sz_dev, _sz_orc)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    assert [m.shape[-1] for m in dmps.numpy_sites()] == [s.shape[-1] for s in st.sites]
    assert max(m.shape[-1] for m in dmps.numpy_sites()) == 8
    _assert_steps_match(sd, so)


def test_tdvp_ps2_holstein_qn_parity():
    """Two-site TDVP-PS with one-electron quantum numbers (Holstein, 4
    molecules, scheme 2, Franck-Condon state expanded to M = 6), bond
    dimension limit 10: populations to 1e-8, quantum numbers identical to
    the oracle's."""
    print("[tdvp ps2 holstein qn 4x4 M=10]", flush=True)
    model = helpers.holstein_model(4, 4, j=-1.0, g=1.0)
    mps, mpo = helpers.holstein_initial(model, 6)
    obs_d = lambda m: m.e_occupations
    obs_o = lambda s: helpers.e_occupations_oracle(s, model)
    od, oo, sd, so, dmps, st = _ps2_pair(mps, mpo, 0.05, 5, 10, obs_d, obs_o)
    np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
    np.testing.assert_allclose(od.sum(axis=1), 1.0, atol=1e-10)
    assert [np.array(q).tolist() for q in dmps.qn] == [np.array(q).tolist() for q in st.qn]
    _assert_steps_match(sd, so)


def test_tdvp_ps2_product_state_exact():
    """Two-site TDVP-PS on the device from the bare product state, bond
    dimension unlimited: the bonds grow from 1 and the state follows
    exp(-iHt) psi (dense propagator, spin-boson 3 modes x 3 levels) to the
    Krylov tolerance, independent of the oracle."""
    print("[tdvp ps2 product state exact]", flush=True)
    from renormalizer_amd import CompressConfig, CompressCriteria, Mpo, Mps
    model = helpers.sbm_model(3, 3, alpha=0.5, omega_c=4.0)
    mpo = Mpo(model)
    mps = Mps.ground_state(model, False)
    st = helpers.to_state(mps)
    psi0 = np.ones((1, 1))
    for s in st.sites:
        psi0 = np.tensordot(psi0, s, axes=1).reshape(-1, s.shape[-1])
    dt, nstep = 0.1, 6
    exact = dense.propagate(mpo.todense(), psi0[:, 0], [dt * k for k in range(nstep + 1)])
    dev = mps.copy()
    dev.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=1000)
    dev.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps2)
    dev.to_device()
    for k in range(1, nstep + 1):
        dev = dev.evolve(mpo, dt)
        psi = np.ones((1, 1))
        for s in dev.numpy_sites():
            psi = np.tensordot(psi, s, axes=1).reshape(-1, s.shape[-1])
        np.testing.assert_allclose(psi[:, 0], exact[k], atol=1e-7)


def test_holstein_disorder_c3_shape_parity():
    """The C3 workload shape, three of its instances: Holstein chains of 8,
    14 and 20 molecules (scheme 2, phonon d = 8), site energies drawn from
    N(0, 0.5) with a fixed seed, J = -1, omega = 1, g = 0.7, Franck-Condon
    electron in the middle, bond dimensions 16, 32 and 48, one-electron
    quantum numbers, dt = 0.02.  Two steps each, FP64 device vs oracle:
    populations to 1e-8, matvec counts and quantum numbers."""
    rng = np.random.default_rng(2024)
    for nmol, m in ((8, 16), (14, 32), (20, 48)):
        print(f"[tdvp C3 instance {nmol} molecules M={m}]", flush=True)
        elocal = rng.normal(0.0, 0.5, nmol)
        model = helpers.holstein_model(nmol, 8, j=-1.0, g=0.7, elocal=list(elocal))
        mps, mpo = helpers.holstein_initial(model, m)
        obs_d = lambda x: x.e_occupations
        obs_o = lambda s, model=model: helpers.e_occupations_oracle(s, model)
        od, oo, sd, so, dmps, st = _run_pair(mps, mpo, 0.02, 2, obs_d, obs_o)
        np.testing.assert_allclose(od, oo, atol=1e-8, rtol=0)
        np.testing.assert_allclose(od.sum(axis=1), 1.0, atol=1e-10)
        _assert_steps_match(sd, so)
        assert [np.array(q).tolist() for q in dmps.qn] == [np.array(q).tolist() for q in st.qn]
