"""Matrix product density operators in the purification picture (reference
renormalizer/mps/mpdm.py:15-161).

A site of the reference's MpDm is a 4-leg tensor (M_l, d, d, M_r): the first
physical index carries the system, the second an ancilla copy, and the
density operator is Tr_ancilla |psi><psi|.  Here the two indices are merged
into one of dimension d * d (system index major), so the state is an ordinary
MPS over a larger local space and every device kernel of the MPS path
applies unchanged.  An operator acts on the system index only: its MPO site
W (D_l, d, d, D_r) becomes W (x) 1_d on the merged index (``ancilla_mpo``),
which is the reference's contraction of the operator with the first index
(hop_expr.py:87-91, 111-115; lib.py:206-211, 238-243).  Cost: the two matrix
products that touch the state (S1, S3 of the effective Hamiltonian) see the
merged dimension d * d exactly as the reference's 4-leg contraction does;
only the small operator stage (S2) carries the extra factor d.
"""
from typing import List

import numpy as np

from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps, expand_bond_dimension_general
from renormalizer_amd.mps.svd_qn import add_outer
from renormalizer_amd.model import Op, OpSum


def ancilla_mpo(mpo: Mpo) -> Mpo:
    """The operator on the merged (system, ancilla) index: site tensors
    W[a, (p r), (q s), b] = W[a, p, q, b] delta_rs.  Cached on ``mpo``."""
    cached = getattr(mpo, "_ancilla_ext", None)
    if cached is not None:
        return cached
    ext = mpo.metacopy()
    sites = []
    for w in mpo._mp:
        w = np.asarray(w)
        d = w.shape[1]
        t = np.einsum("apqb,rs->aprqsb", w, np.eye(d, dtype=w.dtype))
        sites.append(np.ascontiguousarray(t.reshape(w.shape[0], d * d, d * d, w.shape[3])))
    ext._mp = sites
    ext._ancilla_of = mpo
    mpo._ancilla_ext = ext
    return ext


class MpDm(Mps):
    """Purified density operator (reference mpdm.py:15): the sites are kept
    with the system and ancilla indices merged, (M_l, d * d, M_r)."""

    @classmethod
    def random(cls, mpo, qntot, m_max, percent=0):
        raise ValueError("MpDm don't have to produce random state")

    @classmethod
    def ground_state(cls, model, max_entangled, normalize=True, condition=None):
        raise ValueError("Use max_entangled_ex or max_entangled_gs for matrix product density matrix")

    @classmethod
    def from_mps(cls, mps: Mps) -> "MpDm":
        """rho = |mps><mps| lifted to a purification with the ancilla equal
        to the system index, mo[:, i, i, :] = ms[:, i, :] (reference
        mpdm.py:27-48)."""
        new = cls()
        new.model = mps.model
        new._mp = []
        for ms in mps.numpy_sites():
            d = ms.shape[1]
            mo = np.zeros((ms.shape[0], d, d, ms.shape[2]), dtype=ms.dtype)
            for i in range(d):
                mo[:, i, i, :] = ms[:, i, :]
            new._mp.append(mo.reshape(ms.shape[0], d * d, ms.shape[2]))
        new.dtype = mps.dtype
        new.coeff = mps.coeff
        new.optimize_config = mps.optimize_config
        new.evolve_config = mps.evolve_config
        new.qn = [np.array(q).copy() for q in mps.qn]
        new.qntot = np.array(mps.qntot).copy()
        new.qnidx = mps.qnidx
        new.to_right = mps.to_right
        new.compress_config = mps.compress_config.copy()
        return new

    @classmethod
    def max_entangled_ex(cls, model, normalize=True) -> "MpDm":
        """Infinite-temperature state of the one-exciton space: sum_i a_i^+
        applied to the locally maximally entangled ground state (reference
        mpdm.py:54-65)."""
        mps = Mps.ground_state(model, max_entangled=True)
        ex_mps = Mpo.onsite(model, r"a^\dagger") @ mps
        if normalize:
            ex_mps.normalize("mps_and_coeff")
        return cls.from_mps(ex_mps)

    @classmethod
    def max_entangled_gs(cls, model) -> "MpDm":
        return cls.from_mps(Mps.ground_state(model, max_entangled=True))

    def _get_sigmaqn(self, idx):
        """System quantum numbers on the merged index; the ancilla carries
        none (reference mpdm.py:67-70)."""
        up = np.asarray(self.model.basis[idx].sigmaqn)
        return add_outer(up, np.zeros_like(up)).reshape(-1, up.shape[-1])

    @property
    def is_mps(self):
        return False

    @property
    def is_mpo(self):
        return False

    @property
    def is_mpdm(self):
        return True

    # ---- 4-leg view --------------------------------------------------------
    def mpdm_sites(self) -> List[np.ndarray]:
        """Site tensors in the reference's layout (M_l, d, d, M_r)."""
        out = []
        for m, d in zip(self.numpy_sites(), self.model.pbond_list):
            out.append(m.reshape(m.shape[0], d, d, m.shape[2]))
        return out

    def todense(self) -> np.ndarray:
        """The purification as an operator matrix (system x ancilla), as the
        reference's Mpo.todense of the 4-leg sites."""
        res = np.ones((1, 1, 1, 1))
        for mt in self.mpdm_sites():
            d1 = res.shape[1] * mt.shape[1]
            d2 = res.shape[2] * mt.shape[2]
            res = np.tensordot(res, mt, axes=1).transpose((0, 1, 3, 2, 4, 5)).reshape(1, d1, d2, mt.shape[-1])
        return res[0, :, :, 0] * self.coeff

    # ---- operators act on the system index --------------------------------
    def evolve(self, mpo, evolve_dt, normalize=True) -> "MpDm":
        return super().evolve(ancilla_mpo(mpo), evolve_dt, normalize)

    def expectation(self, mpo, self_conj: "Mps" = None):
        if isinstance(mpo, (Op, OpSum)):
            mpo = Mpo(self.model, mpo)
        return super().expectation(ancilla_mpo(mpo), self_conj)

    def expectations(self, mpos, self_conj: "Mps" = None, opt: bool = True) -> np.ndarray:
        mpos = [Mpo(self.model, m) if isinstance(m, (Op, OpSum)) else m for m in mpos]
        return super().expectations([ancilla_mpo(m) for m in mpos], self_conj, opt)

    def calc_1site_rdm(self, idx=None):
        raise NotImplementedError("one-site reduced density matrices are defined for MPS")

    def expand_bond_dimension(self, hint_mpo=None, coef=1e-10, include_ex=True):
        """reference mps.py:1934-1960 with the MpDm branch: the expansion
        states come from the one-exciton infinite-temperature state."""
        if hint_mpo is not None and include_ex:
            assert int(np.asarray(self.qntot).ravel()[0]) == 1
            ex_state = MpDm.max_entangled_ex(self.model)
            ex_state.compress_config = self.compress_config
            ex_state.move_qnidx(self.qnidx)
            ex_state.to_right = self.to_right
        else:
            ex_state = None
        return expand_bond_dimension_general(self, hint_mpo, coef, ex_state)

    # ---- on-disk format: the reference's 4-leg site arrays ----------------
    def dump(self, fname):
        data = {"version": "0.4", "nsites": len(self)}
        for i, mt in enumerate(self.mpdm_sites()):
            data[f"mt_{i}"] = mt
        qn = [np.array(q) for q in self.qn]
        arr = np.empty(len(qn), object)
        arr[:] = qn
        data["qn"] = arr
        for i, q in enumerate(qn):
            data[f"subqn_{i}"] = q
        data["qnidx"] = self.qnidx
        data["qntot"] = np.array(self.qntot)
        data["to_right"] = self.to_right
        data["coeff"] = self.coeff
        np.savez(fname, **data)

    @classmethod
    def load(cls, model, fname: str) -> "MpDm":
        mps = Mps.load.__func__(cls, model, fname)
        mps._mp = [m.reshape(m.shape[0], m.shape[1] * m.shape[2], m.shape[3]) if m.ndim == 4 else m
                   for m in mps._mp]
        return mps

    def __str__(self):
        return f"MpDm(nsite={len(self)}, bond dims={self.bond_dims})"
