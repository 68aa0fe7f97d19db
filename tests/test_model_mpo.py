"""Model / operator algebra / MPO construction (host side, CPU).

Pins: reference doctest answers of Op/OpSum (model/op.py:56-73, 433-447,
466-471), closed-form basis matrices, dense MPO == independently built
Kronecker Hamiltonian (the pattern of reference test_mpo.py:25-42), MPO
bond dimensions (spin-boson D=3, Holstein 4/5), offset shift
(test_mpo.py:90-102), Hermiticity.
"""
import numpy as np
import pytest

from oracle import dense
from renormalizer_amd import (BasisHalfSpin, BasisSHO, BasisSimpleElectron, HolsteinModel, Model, Mol,
                              Mpo, Op, OpSum, Phonon, Quantity, SpinBosonModel)


def test_op_algebra_reference_examples():
    x = Op("X", 0, 0.5)
    y = Op("Y", 1, 0.2)
    assert str(3 * x) == "Op('X', [0], 1.5)"
    assert str(x + y) == "[Op('X', [0], 0.5), Op('Y', [1], 0.2)]"
    assert str(x - y) == "[Op('X', [0], 0.5), Op('Y', [1], -0.2)]"
    assert str(x * y) == "Op('X Y', [0, 1], 0.1)"
    assert str((y + x) * (x + y)) == ("[Op('Y X', [1, 0], 0.1), Op('Y Y', [1, 1], 0.04000000000000001), "
                                      "Op('X X', [0, 0], 0.25), Op('X Y', [0, 1], 0.1)]")
    op = Op(r"a^\dagger a", ["site0", "site1"], 2.0, qn=[1, -1])
    assert str(op) == "Op('a^\\\\dagger a', ['site0', 'site1'], 2.0, [[1], [-1]])"
    assert str(Op("X I Y I", [0, 1, 2, 3], 0.5).squeeze_identity()) == "Op('X Y', [0, 2], 0.5)"
    ops = Op("X Y", [3, 2], 0.5) * Op("Y X", [2, 3], 3.0) * Op("Z Z", [2, 2], 1.0)
    elem, f = ops.split_elementary({2: 0, 3: 1})
    assert str(elem) == "[Op('Y Y Z Z', [2, 2, 2, 2], 1.0), Op('X X', [3, 3], 1.0)]" and f == 1.5
    s = Op("X", 0, 1.0) + Op("Y", 1, 2.0)
    assert str((s + s).simplify()) == "[Op('X', [0], 2.0), Op('Y', [1], 4.0)]"
    assert str(s / 0.5) == "[Op('X', [0], 2.0), Op('Y', [1], 4.0)]"
    assert str((s - s).simplify()) == "[]"
    assert str(OpSum([Op("X I Y I", [0, 1, 2, 3], 0.5), Op("X Y", [0, 2], 0.5),
                      Op("Z", [1], 1e-4)]).simplify(atol=1e-3)) == "[Op('X Y', [0, 2], 1.0)]"
    # default quantum numbers of the electron ladder operators
    assert Op(r"a^\dagger", 0).qn.tolist() == [1] and Op("a", 0).qn.tolist() == [-1]


def test_basis_matrices_closed_forms():
    b = BasisSimpleElectron(0)
    np.testing.assert_array_equal(b.op_mat(r"a^\dagger"), [[0, 0], [1, 0]])
    s = BasisHalfSpin(0)
    np.testing.assert_array_equal(s.op_mat("X"), [[0, 1], [1, 0]])
    np.testing.assert_array_equal(-1 * s.op_mat("iY") @ s.op_mat("iY"), np.eye(2))
    np.testing.assert_array_equal(s.op_mat("sigma_z"), np.diag([1, -1]))
    h = BasisSHO(0, 0.7, 6)
    x = h.op_mat("x")
    # [x, p] = i away from the truncation edge
    p = h.op_mat("p")
    comm = x @ p - p @ x
    np.testing.assert_allclose(np.diag(comm)[:-1], 1j, atol=1e-12)
    # p^2/2 + w^2 x^2/2 = w (n + 1/2) exactly (normal-ordered forms)
    hh = 0.5 * h.op_mat("p^2") + 0.5 * 0.7 ** 2 * h.op_mat("x^2")
    np.testing.assert_allclose(hh, np.diag(0.7 * (np.arange(6) + 0.5)), atol=1e-12)
    np.testing.assert_allclose(h.op_mat(Op("x", 0, 2.0)), 2 * x)


def _sbm(nmode, nlev, eps=0.3, delta=1.0):
    omegas = np.linspace(0.4, 2.0, nmode)
    dis = np.linspace(0.3, 1.1, nmode)
    ph = [Phonon.simple_phonon(Quantity(w), Quantity(d), nlev) for w, d in zip(omegas, dis)]
    return SpinBosonModel(Quantity(eps), Quantity(delta), ph), omegas, dis


def _holstein(nmol, nlev, j=-1.0, g=1.0):
    omega = 1.0
    dis = g * np.sqrt(2 / omega)
    mol = Mol(Quantity(0), [Phonon.simple_phonon(Quantity(omega), Quantity(dis), nlev)])
    return HolsteinModel([mol] * nmol, Quantity(j), scheme=2), omega, dis


def test_spin_boson_mpo_dense():
    model, omegas, dis = _sbm(3, 4)
    mpo = Mpo(model)
    assert mpo.bond_dims == [1, 3, 3, 3, 1]
    h = dense.spin_boson_dense(0.3, 1.0, omegas, dis, 4)
    np.testing.assert_allclose(mpo.todense(), h, atol=1e-12)
    assert mpo.is_hermitian()


def test_holstein_mpo_dense():
    model, omega, dis = _holstein(3, 3)
    mpo = Mpo(model)
    assert max(mpo.bond_dims) in (4, 5)
    jm = model.j_matrix
    h = dense.holstein_dense([0.0] * 3, jm, [omega] * 3, [dis] * 3, 3)
    np.testing.assert_allclose(mpo.todense(), h, atol=1e-12)


def test_mpo_offset_and_identity():
    model, omegas, dis = _sbm(2, 3)
    h = Mpo(model).todense()
    ho = Mpo(model, offset=Quantity(0.7)).todense()
    np.testing.assert_allclose(ho, h - 0.7 * np.eye(len(h)), atol=1e-12)
    np.testing.assert_allclose(Mpo.identity(model).todense(), np.eye(len(h)))


def test_mpo_single_term_and_qn():
    model, _, _ = _holstein(4, 2)
    mpo = Mpo.onsite(model, r"a^\dagger", dof_set={2})
    assert mpo.bond_dims == [1] * 9
    assert mpo.qntot.tolist() == [1]
    dims = [2, 2] * 4
    np.testing.assert_allclose(mpo.todense(), dense._embed({4: dense.ADAG}, dims))


def test_large_model_bond_dims():
    """C4-size spin-boson (64 modes x 10 levels): D = 3 on every interior bond."""
    model, _, _ = _sbm(64, 10)
    mpo = Mpo(model)
    assert mpo.bond_dims == [1] + [3] * 64 + [1]
    assert all(w.shape[1] == w.shape[2] for w in mpo)


def test_model_errors():
    with pytest.raises(ValueError):
        Model([BasisHalfSpin(0), BasisHalfSpin(0)], [])
    with pytest.raises(ValueError):
        Model([BasisHalfSpin(0)], [Op("X", 1)])
    m = Model([BasisHalfSpin(0)], [Op("X", 0, 0.0), Op("Z", 0, 1.0)])
    assert len(m.ham_terms) == 1
