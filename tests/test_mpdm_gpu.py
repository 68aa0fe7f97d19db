"""Purified density operators on the device (reference mps/mpdm.py,
mps/thermalprop.py): the reference's own MpDm tests and imaginary-time
parity with the oracle's 4-leg restatement."""
import numpy as np
import pytest

from oracle import tdvp as orc
from renormalizer_amd import CompressConfig, CompressCriteria, EvolveConfig, EvolveMethod, Mpo, Mps, Quantity
from renormalizer_amd.mps.mpdm import MpDm
from renormalizer_amd.mps.thermalprop import ThermalProp
from tests import helpers
from tests.test_mpdm import _asymmetric_model, _expanded_mpdm, _gibbs_occupations
from tests.test_tdvp_gpu import _assert_steps_match

pytestmark = pytest.mark.gpu


def test_from_mps_occupations():
    """reference test_mpdm.py:11-17: the purification of a random MPS has
    its electron occupations, before and after canonicalisation."""
    print("[mpdm from_mps]", flush=True)
    np.random.seed(2)
    model = helpers.reference_holstein_model()
    gs = Mps.random(model, 1, 20)
    dm = MpDm.from_mps(gs)
    np.testing.assert_allclose(gs.e_occupations, dm.e_occupations, rtol=1e-10, atol=1e-12)
    gs = gs.canonicalise()
    dm = dm.canonicalise()
    np.testing.assert_allclose(gs.e_occupations, dm.e_occupations, rtol=1e-10, atol=1e-12)


def test_thermal_prop_reference_values():
    """reference test_mpdm.py:20-57 (adaptive TDVP-PS row): the 3-molecule
    Holstein model at 298 K, electron occupations and internal energy
    within rtol 5e-3 of the exact values quoted there."""
    print("[mpdm thermal prop 298 K]", flush=True)
    model = helpers.reference_holstein_model()
    init_mps = MpDm.max_entangled_ex(model)
    beta = Quantity(298, "K").to_beta()
    evolve_time = beta / 2j
    evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, adaptive=True, guess_dt=0.1 / 1j)
    tp = ThermalProp(init_mps, evolve_config=evolve_config)
    tp.evolve(evolve_dt=evolve_time, nsteps=1)
    etot_std = 0.0853388 + model.gs_zpe
    occ_std = [0.20896541050347484, 0.35240029674394463, 0.4386342927525734]
    np.testing.assert_allclose(tp.e_occupations_array[-1], occ_std, rtol=5e-3)
    np.testing.assert_allclose(tp.energies[-1], etot_std, rtol=5e-3)
    assert abs(tp.e_occupations_array[-1].sum() - 1) < 1e-8
    assert tp.vn_entropy_array.shape[0] == 2


def test_imaginary_time_parity_with_oracle():
    """Imaginary-time TDVP-PS of a density operator on the device (merged
    system/ancilla index) vs the oracle's 4-leg contractions: occupations
    to 1e-8 and the same Lanczos matvec counts, and both on the dense Gibbs
    state of the one-exciton space."""
    print("[mpdm imaginary time parity]", flush=True)
    model = _asymmetric_model()
    mpo = Mpo(model)
    dm = _expanded_mpdm(model, mpo)
    st = helpers.mpdm_state(dm)
    w = helpers.mpo_arrays(mpo)
    dev = dm.copy()
    dev.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    dev.to_device()
    dtau = 0.1
    sd, so = [], []
    for k in range(1, 7):
        dev = dev.evolve(mpo, -1j * dtau)
        st, ls = orc.tdvp_ps(st, w, -1j * dtau)
        sd.append(list(dev.evolve_config.local_steps))
        so.append(ls)
        got = dev.e_occupations
        np.testing.assert_allclose(got, helpers.e_occupations_mpdm_oracle(st, model), atol=1e-8)
        np.testing.assert_allclose(got, _gibbs_occupations(model, mpo, 2 * dtau * k), atol=1e-7)
    _assert_steps_match(sd, so)
    assert isinstance(dev, MpDm)


def test_expectation_is_trace_with_ancilla():
    """<H> of a truncated density operator on the device equals the 4-leg
    oracle contraction (ancilla traced between bra and ket)."""
    print("[mpdm expectation]", flush=True)
    model = _asymmetric_model()
    mpo = Mpo(model)
    dm = MpDm.max_entangled_ex(model)
    dm.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=6)
    dm = dm.expand_bond_dimension(mpo, coef=0.3).canonicalise()
    st = helpers.mpdm_state(dm)
    want = orc.expectation(st, helpers.mpo_arrays(mpo))
    dm.to_device()
    np.testing.assert_allclose(dm.expectation(mpo), want, rtol=1e-12)
