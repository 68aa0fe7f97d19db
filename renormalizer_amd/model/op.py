"""Symbolic operators (reference model/op.py:11-528).

``Op`` is an immutable product of elementary symbols with DoF names, a
factor and per-symbol quantum numbers; ``OpSum`` is a list of ``Op``.  The
behaviour the MPO construction relies on (splitting into per-site
elementary operators, default quantum numbers of ``a^\\dagger``/``a``,
equality/hash through ``to_tuple``) follows the reference.
"""
from collections import defaultdict
from itertools import chain
from typing import List, Tuple, Union

import numpy as np

from renormalizer_amd.utils import Quantity


def _default_qn(symbol: str) -> int:
    if symbol == r"a^\dagger":
        return 1
    if symbol == "a":
        return -1
    return 0


class Op:
    @classmethod
    def product(cls, op_list: List["Op"]) -> "Op":
        symbol = " ".join(op.symbol for op in op_list)
        dofs = list(chain.from_iterable(op.dofs for op in op_list))
        factor = np.prod([op.factor for op in op_list])
        qn = list(chain.from_iterable(op.qn_list for op in op_list))
        return cls(symbol, dofs, factor, qn)

    @classmethod
    def identity(cls, dof, qn_size=1, factor=1.0):
        if isinstance(dof, list):
            return cls(" ".join(["I"] * len(dof)), dof, factor=factor,
                       qn=[np.zeros(qn_size, dtype=int)] * len(dof))
        return cls("I", dof, factor=factor, qn=[np.zeros(qn_size, dtype=int)])

    def __init__(self, symbol: str, dof, factor: Union[float, complex, Quantity] = 1.0,
                 qn: Union[List, int] = None):
        if not isinstance(symbol, str):
            raise TypeError(f"symbol should be a str. Got {symbol} as {type(symbol)}")
        self.symbol = symbol
        # "b^\dagger + b" is one simple symbol; everything else splits on spaces
        self.split_symbol: List[str] = symbol.replace(r"b^\dagger + b", r"b^\dagger+b").split(" ")
        nsym = len(self.split_symbol)
        if nsym == 1:
            if isinstance(dof, list):
                assert len(dof) == 1
                dofs = dof
            else:
                dofs = [dof]
            if isinstance(qn, list):
                if len(qn) != 1:
                    raise ValueError(f"Incompatible sizes of quantum number {qn} and symbol {self.split_symbol}")
                qn_list = qn
            elif qn is None:
                qn_list = None
            else:
                qn_list = [qn]
        else:
            if isinstance(dof, list):
                if nsym != len(dof):
                    raise ValueError("symbol and DoF name not match")
                dofs = dof
            else:
                dofs = [dof] * nsym
            if isinstance(qn, list):
                if nsym != len(qn):
                    raise ValueError("symbol and qn length not match")
                qn_list = qn
            elif qn is None:
                qn_list = None
            else:
                raise ValueError("qn should be a list.")
        if qn_list is None:
            qn_list = [_default_qn(s) for s in self.split_symbol]
        for d in dofs:
            if d.__hash__ is None:
                raise ValueError(f"dof name should be hashable. Got {d}.")
        self.dofs: List = dofs
        if isinstance(factor, Quantity):
            factor = factor.as_au()
        self._factor = factor + 0.0
        self.qn_list: List[np.ndarray] = [np.array(q).reshape(-1) for q in qn_list]

    def split_elementary(self, dof_to_siteidx) -> Tuple[List["Op"], Union[float, complex]]:
        """Group the symbols by MPS site (order kept within a site, sites in
        increasing order), factors set to 1 (reference op.py:185-228)."""
        if len(self.dofs) == 1:
            return [Op(self.symbol, self.dofs, qn=self.qn_list)], self.factor
        grouped = defaultdict(list)
        for sym, name, qn in zip(self.split_symbol, self.dofs, self.qn_list):
            site = dof_to_siteidx.get(name)
            if site is None:
                raise ValueError(f"Unknown DoF name {name} in {self}.")
            grouped[site].append(Op(sym, name, qn=qn))
        return [Op.product(grouped[s]) for s in sorted(grouped)], self.factor

    @property
    def factor(self):
        return self._factor

    @property
    def qn(self) -> np.ndarray:
        return sum(self.qn_list)

    @property
    def qn_size(self) -> int:
        return len(self.qn)

    @property
    def is_identity(self) -> bool:
        return set(self.split_symbol) == {"I"}

    def squeeze_identity(self):
        if self.is_identity:
            return self.__class__.identity(self.dofs[0], factor=self.factor, qn_size=self.qn_size)
        syms, dofs, qns = [], [], []
        for s, d, q in zip(self.split_symbol, self.dofs, self.qn_list):
            if s == "I":
                continue
            syms.append(s)
            dofs.append(d)
            qns.append(q)
        return Op(" ".join(syms), dofs, self.factor, qns)

    def same_term(self, other) -> bool:
        return self.symbol == other.symbol and self.dofs == other.dofs

    def to_tuple(self):
        return self.symbol, tuple(self.dofs), self.factor, tuple(tuple(t) for t in self.qn_list)

    def __hash__(self):
        return hash(self.to_tuple())

    def __eq__(self, other):
        return self.to_tuple() == other.to_tuple()

    def __str__(self):
        ret = ", ".join([repr(self.symbol), str(self.dofs), str(self.factor)])
        if not np.all(np.array(self.qn_list) == 0):
            ret += f", {[q.tolist() for q in self.qn_list]}"
        return f"Op({ret})"

    __repr__ = __str__

    def __add__(self, other):
        if isinstance(other, (int, float)) and other == 0:
            return OpSum([self])
        if isinstance(other, np.ndarray) and np.array_equal(other, np.array(0)):
            return OpSum([self])
        if isinstance(other, Op):
            return OpSum([self, other])
        if isinstance(other, list):
            return OpSum([self] + other)
        raise TypeError(f"Unknown operand type {type(other)}")

    def __radd__(self, other):
        if (isinstance(other, (int, float)) and other == 0) or \
                (isinstance(other, np.ndarray) and np.array_equal(other, np.array(0))):
            return OpSum([self])
        raise TypeError(f"Unknown operand type {type(other)}")

    def __neg__(self):
        return Op(self.symbol, self.dofs, -self.factor, self.qn_list)

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        if isinstance(other, np.generic):
            other = other.item()
        if isinstance(other, Op):
            return Op.product([self, other])
        if isinstance(other, (int, float, complex)):
            return Op(self.symbol, self.dofs, self.factor * other, self.qn_list)
        if isinstance(other, list):
            for item in other:
                if not isinstance(item, Op):
                    raise TypeError(f"Operand must be a list of `Op`. Got {type(item)}")
            return OpSum([self * item for item in other])
        raise TypeError(f"Unsupported type: {type(other)}")

    def __rmul__(self, other):
        if isinstance(other, (int, float, complex, np.generic)):
            return self * other
        if isinstance(other, list):
            return OpSum(other) * self
        raise TypeError(f"Unknwon type {type(other)}")


class OpSum(list):
    @classmethod
    def product(cls, op_list):
        if len(op_list) == 0:
            return cls()
        prod = op_list[0]
        for op in op_list[1:]:
            prod = prod * op
        return prod

    def copy(self):
        return OpSum(super().copy())

    def simplify(self, atol=0) -> "OpSum":
        """Combine identical terms, then drop |factor| <= atol (reference
        op.py:451-490, same term order)."""
        pending = OpSum(reversed([op.squeeze_identity() for op in self]))
        merged = []
        while pending:
            op = pending.pop()
            same = [i for i, o in enumerate(pending) if op.same_term(o)]
            op = Op(op.symbol, op.dofs, op.factor + sum(pending[i].factor for i in same), op.qn_list)
            for k, i in enumerate(same):
                pending.pop(i - k)
            merged.append(op)
        return OpSum([op for op in merged if np.abs(op.factor) > atol])

    def __add__(self, other):
        if not isinstance(other, (Op, list)):
            raise TypeError("OpSum can only add with `Op` or list of `Op`")
        if isinstance(other, Op):
            other = [other]
        return OpSum(super().__add__(other))

    def __iadd__(self, other):
        if isinstance(other, Op):
            self.append(other)
            return self
        return super().__iadd__(other)

    def __neg__(self):
        return OpSum([-op for op in self])

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        if isinstance(other, list):
            res = []
            for op1 in self:
                res.extend(op1 * other)
            return OpSum(res)
        if isinstance(other, (int, float, complex, np.generic, Op)):
            return OpSum([op * other for op in self])
        return OpSum(super().__mul__(other))

    def __rmul__(self, other):
        if isinstance(other, (int, float, complex, np.generic)):
            return self * other
        return OpSum(super().__rmul__(other))

    def __truediv__(self, other):
        assert isinstance(other, (int, float, complex, np.generic))
        return self * (1 / other)

    __array_ufunc__ = None
