"""Local basis sets and operator matrices (reference model/basis.py).

Only the bases on the TDVP-PS path are provided: harmonic oscillator
(BasisSHO, basis.py:110-339), spin-1/2 (BasisHalfSpin, 932-996) and the
two-level electron (BasisSimpleElectron, 882-929).  Matrices are the
standard second-quantisation closed forms with the reference's conventions
(x = sqrt(1/2w)(b^+ + b), p^2 = -w/2 (b^+b^+ - b^+b - bb^+ + bb), factor
included in the returned matrix).
"""
from typing import List, Union

import numpy as np

from renormalizer_amd.model.op import Op


class BasisSet:
    is_electron = False
    is_phonon = False
    is_spin = False
    multi_dof = False

    def __init__(self, dof, nbas: int, sigmaqn: List):
        self.dof = dof
        assert type(nbas) is int
        self.nbas = nbas
        qns = []
        for qn in sigmaqn:
            if isinstance(qn, (int, np.integer)):
                qn = [int(qn)]
            qns.append(np.array(qn))
        self.sigmaqn: np.ndarray = np.array(qns)

    def __str__(self):
        ret = f"dof: {self.dof}, nbas: {self.nbas}"
        if not np.all(self.sigmaqn == 0):
            ret += f", qn: {self.sigmaqn.tolist()}"
        return f"{self.__class__.__name__}({ret})"

    __repr__ = __str__

    def op_mat(self, op):
        raise NotImplementedError

    @property
    def dofs(self):
        return tuple(self.dof) if self.multi_dof else (self.dof,)

    def copy(self, new_dof):
        raise NotImplementedError


def _as_op(op) -> Op:
    return op if isinstance(op, Op) else Op(op, None)


class BasisSHO(BasisSet):
    is_phonon = True

    def __init__(self, dof, omega, nbas, x0=0.0, dvr=False, general_xp_power=False):
        if dvr:
            raise NotImplementedError("DVR representation is outside the TDVP-PS hot path")
        self.omega = omega
        self.x0 = x0
        super().__init__(dof, nbas, [0] * nbas)
        self.general_xp_power = general_xp_power
        self.dvr = False

    def __str__(self):
        return f"BasisSHO(dof: {self.dof}, x0: {self.x0}, omega: {self.omega}, nbas: {self.nbas})"

    def _ladder(self, symbol):
        n = self.nbas
        if symbol == "b":
            return np.diag(np.sqrt(np.arange(1, n)), k=1)
        if symbol == r"b^\dagger":
            return np.diag(np.sqrt(np.arange(1, n)), k=-1)
        if symbol == "b b":
            if n == 1:
                return np.zeros((1, 1))
            return np.diag(np.sqrt(np.arange(1, n - 1) * np.arange(2, n)), k=2)
        if symbol == r"b^\dagger b^\dagger":
            if n == 1:
                return np.zeros((1, 1))
            return np.diag(np.sqrt(np.arange(1, n - 1) * np.arange(2, n)), k=-2)
        if symbol == r"b^\dagger b":
            return np.diag(np.arange(n)).astype(float)
        if symbol == r"b b^\dagger":
            return np.diag(np.arange(n) + 1).astype(float)
        raise KeyError(symbol)

    def _mat(self, symbol):
        n = self.nbas
        w = self.omega
        lad = self._ladder
        if symbol in ("b", "b b", r"b^\dagger", r"b^\dagger b^\dagger", r"b^\dagger b", r"b b^\dagger"):
            return lad(symbol)
        if symbol == r"b^\dagger+b":
            return lad(r"b^\dagger") + lad("b")
        if symbol == r"b^\dagger-b":
            return lad(r"b^\dagger") - lad("b")
        if symbol == "x":
            return np.sqrt(0.5 / w) * (lad(r"b^\dagger") + lad("b")) + np.eye(n) * self.x0
        if symbol in ("x^2", "x x"):
            return (np.eye(n) * self.x0 ** 2
                    + 2 * self.x0 * np.sqrt(0.5 / w) * (lad(r"b^\dagger") + lad("b"))
                    + 0.5 / w * (lad(r"b^\dagger b^\dagger") + lad(r"b^\dagger b")
                                 + lad(r"b b^\dagger") + lad("b b")))
        if symbol == "p":
            return 1j * np.sqrt(w / 2) * (lad(r"b^\dagger") - lad("b"))
        if symbol in ("p^2", "p p"):
            return -w / 2 * (lad(r"b^\dagger b^\dagger") - lad(r"b^\dagger b")
                             - lad(r"b b^\dagger") + lad("b b"))
        if symbol == "x p":
            return -1.0j / 2 * (lad("b b") - lad(r"b^\dagger b^\dagger")
                                + lad(r"b b^\dagger") - lad(r"b^\dagger b"))
        if symbol == "p x":
            return -1.0j / 2 * (lad("b b") - lad(r"b^\dagger b^\dagger")
                                - lad(r"b b^\dagger") + lad(r"b^\dagger b"))
        if symbol == "dx":
            return (self._mat("p") / -1.0j).real
        if symbol in ("dx^2", "dx dx"):
            return -self._mat("p^2")
        if symbol == "I":
            return np.eye(n)
        if symbol == "n":
            return np.diag(np.arange(n)).astype(float)
        raise ValueError(f"op_symbol:{symbol} is not supported. ")

    def op_mat(self, op: Union[Op, str]):
        op = _as_op(op)
        symbol = op.symbol.replace("partialx", "dx").replace(r"b^\dagger + b", r"b^\dagger+b")
        return self._mat(symbol) * op.factor

    def copy(self, new_dof):
        return self.__class__(new_dof, omega=self.omega, nbas=self.nbas, x0=self.x0)


class BasisSimpleElectron(BasisSet):
    is_electron = True

    def __init__(self, dof, sigmaqn=None):
        super().__init__(dof, 2, [0, 1] if sigmaqn is None else sigmaqn)

    def op_mat(self, op):
        op = _as_op(op)
        mat = np.zeros((2, 2))
        if op.symbol == r"a^\dagger":
            mat[1, 0] = 1.0
        elif op.symbol == "a":
            mat[0, 1] = 1.0
        elif op.symbol == r"a^\dagger a":
            mat[1, 1] = 1.0
        elif op.symbol == "I":
            mat = np.eye(2)
        else:
            raise ValueError(f"op_symbol:{op.symbol} is not supported")
        return mat * op.factor

    def copy(self, new_dof):
        return self.__class__(new_dof)


class BasisHalfSpin(BasisSet):
    is_spin = True

    def __init__(self, dof, sigmaqn: List = None):
        super().__init__(dof, 2, [0, 0] if sigmaqn is None else sigmaqn)

    @staticmethod
    def _single(symbol):
        if symbol == "I":
            return np.eye(2)
        if symbol in ("sigma_x", "X", "x"):
            return np.array([[0.0, 1.0], [1.0, 0.0]])
        if symbol in ("sigma_y", "Y", "y"):
            return np.array([[0.0, -1.0j], [1.0j, 0.0]])
        if symbol in ("isigma_y", "iY", "iy"):
            return np.array([[0.0, 1.0], [-1.0, 0.0]])
        if symbol in ("sigma_z", "Z", "z"):
            return np.diag([1.0, -1.0])
        if symbol in ("sigma_-", "-"):
            return np.array([[0.0, 0.0], [1.0, 0.0]])
        if symbol in ("sigma_+", "+"):
            return np.array([[0.0, 1.0], [0.0, 0.0]])
        raise ValueError(f"op_symbol:{symbol} is not supported")

    def op_mat(self, op: Union[Op, str]):
        op = _as_op(op)
        syms = op.split_symbol
        if len(syms) == 1:
            mat = self._single(syms[0])
        else:
            mat = np.eye(2)
            for s in syms:
                mat = mat @ self._single(s)
        return mat * op.factor

    def copy(self, new_dof):
        return self.__class__(new_dof, self.sigmaqn)
