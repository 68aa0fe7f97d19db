"""Host parts of the adjoint gradients (renormalizer_amd.mps.gradient): the
Hamiltonian builder behind the autograd interface is H_0 + sum_k c_k
sigma_z x_k, pinned on dense matrices (the model's own MPO at its couplings
c_k = -omega_k^2 dis_k, reference model/model.py:436, and linearity in
every c_k)."""
import numpy as np

from oracle import dense
from renormalizer_amd import Mpo
from renormalizer_amd.mps.gradient import spin_boson_mpo_of
from tests import helpers


def test_spin_boson_mpo_of_is_linear_in_couplings():
    model = helpers.sbm_model(3, 3)
    build = spin_boson_mpo_of(model)
    c = np.array([-ph.omega[1] ** 2 * ph.dis[1] for ph in model.ph_list])
    h0 = Mpo(model).todense()
    np.testing.assert_allclose(build(c).todense(), h0, atol=1e-12)
    nlev = 3
    sz = np.diag([1.0, -1.0])
    for k in range(3):
        cp = c.copy()
        cp[k] += 0.3
        dh = build(cp).todense() - h0
        # the change is 0.3 sigma_z (x) x_k, x_k of the mode's oscillator
        mats = [sz] + [np.eye(nlev)] * 3
        mats[k + 1] = dense.sho_ops(model.ph_list[k].omega[0], nlev)[0]
        op = mats[0]
        for m in mats[1:]:
            op = np.kron(op, m)
        np.testing.assert_allclose(dh, 0.3 * op, atol=1e-12)
