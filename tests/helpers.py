"""Shared builders for the parity tests: models, initial states, and the
conversion of a host Mps into the oracle's State."""
import numpy as np

from oracle import tdvp as orc
from renormalizer_amd import (CompressConfig, CompressCriteria, EvolveConfig, EvolveMethod, HolsteinModel,
                              Mol, Mpo, Mps, Phonon, Quantity, SpinBosonModel)


def sbm_model(nmode, nlev, alpha=0.5, omega_c=5.0, eps=0.0, delta=1.0, log_grid=True):
    """Ohmic spin-boson model discretised with the reference's Wang1 grid
    (sbm/lib.py:116-125, post_process 89-103)."""
    from renormalizer_amd.sbm import SpectralDensityFunction
    sdf = SpectralDensityFunction(alpha, omega_c)
    if log_grid:
        w, c2 = sdf.Wang1(nmode)
    else:
        w, c2 = sdf.trapz(nmode, 0.0, omega_c * 3)
    omegas, dis = sdf.post_process(w, c2)
    ph = [Phonon.simple_phonon(o, d, nlev) for o, d in zip(omegas, dis)]
    return SpinBosonModel(Quantity(eps), Quantity(delta), ph)


def holstein_model(nmol, nlev, j=-1.0, g=1.0, omega=1.0, elocal=None):
    dis = g * np.sqrt(2.0 / omega)
    mols = []
    for i in range(nmol):
        e = 0.0 if elocal is None else elocal[i]
        mols.append(Mol(Quantity(e), [Phonon.simple_phonon(Quantity(omega), Quantity(dis), nlev)]))
    return HolsteinModel(mols, Quantity(j), scheme=2)


def sbm_initial(model, m, coef=1e-16):
    """SpinBosonDynamics.init_mps (reference sbm/sbm.py:55-59): spin up x
    vacuum, expanded along the Hamiltonian to bond dimension m."""
    mpo = Mpo(model)
    mps = Mps.ground_state(model, False)
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=m)
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    mps = mps.expand_bond_dimension(mpo, coef=coef, include_ex=False)
    return mps, mpo


def holstein_initial(model, m, center=None, energy_offset=True):
    """ChargeDiffusionDynamics.init_mps with a Franck-Condon electron
    (reference transport/dynamics.py:138-199)."""
    center = model.mol_num // 2 if center is None else center
    tentative = Mpo(model)
    gs = Mps.ground_state(model, False)
    mps = Mpo.onsite(model, r"a^\dagger", dof_set={center}) @ gs
    if energy_offset:
        e = orc.expectation(to_state(mps), [np.asarray(w) for w in tentative])
        mpo = Mpo(model, offset=Quantity(e))
    else:
        mpo = tentative
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=m)
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    mps = mps.expand_bond_dimension(mpo)
    mps.canonicalise()
    return mps, mpo


def to_state(mps) -> orc.State:
    sig = [b.sigmaqn for b in mps.model.basis]
    return orc.State(mps.numpy_sites(), [np.array(q) for q in mps.qn], mps.qnidx, mps.qntot,
                     mps.to_right, sig)


def mpo_arrays(mpo):
    return [np.asarray(w) for w in mpo]


def e_occupations_oracle(state, model):
    out = []
    for i, b in enumerate(model.basis):
        if b.is_electron:
            out.append(orc.calc_1site_rdm(state, i)[1, 1].real)
    return np.array(out)


def reference_holstein_model():
    """renormalizer/tests/parameter.py:8-30 (3 molecules, two modes each,
    4 levels, ground / excited frequencies equal)."""
    from renormalizer_amd.utils import constant
    elocalex = Quantity(2.67, "eV")
    j_matrix = np.array([[0.0, -0.1, -0.2], [-0.1, 0.0, -0.3], [-0.2, -0.3, 0.0]]) / constant.au2ev
    omega_q = [Quantity(106.51, "cm^{-1}"), Quantity(1555.55, "cm^{-1}")]
    omega = [[omega_q[0], omega_q[0]], [omega_q[1], omega_q[1]]]
    dis_q = [Quantity(30.1370, "a.u."), Quantity(8.7729, "a.u.")]
    displacement = [[Quantity(0), dis_q[0]], [Quantity(0), dis_q[1]]]
    ph_list = [Phonon(*args) for args in zip(omega, displacement, [4, 4])]
    return HolsteinModel([Mol(elocalex, ph_list, 15.45)] * 3, j_matrix)


def mpdm_state(mpdm) -> orc.State:
    """Oracle state with the reference's 4-leg density-operator sites."""
    sig = [orc.add_outer(np.asarray(b.sigmaqn), np.zeros_like(np.asarray(b.sigmaqn)))
           for b in mpdm.model.basis]
    return orc.State(mpdm.mpdm_sites(), [np.array(q) for q in mpdm.qn], mpdm.qnidx, mpdm.qntot,
                     mpdm.to_right, sig)


def e_occupations_mpdm_oracle(state, model):
    """<a_i^+ a_i> of a 4-leg purified state, normalised by its norm."""
    from renormalizer_amd import Mpo, Op
    nrm2 = orc.mp_norm(state) ** 2
    return np.array([orc.expectation(state, mpo_arrays(Mpo(model, Op(r"a^\dagger a", dof)))) / nrm2
                     for dof in model.e_dofs])
