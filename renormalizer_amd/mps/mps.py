"""Matrix product state with the one-site TDVP-PS time step on the device
(reference mps/mps.py).

Entry points kept from the reference: ``Mps.ground_state``,
``Mps.hartree_product_state``, ``expand_bond_dimension``, ``evolve`` (with
``EvolveMethod.tdvp_ps``, optionally ``adaptive``), ``normalize``,
``expectation``/``expectations``, ``e_occupations``/``ph_occupations``,
``calc_1site_rdm``, ``mp_norm``/``norm``/``dot``/``distance``, ``dump``/``load``.

The time step (``_evolve_tdvp_ps``, reference mps.py:1267-1404) runs entirely
through ``libreno_hip.so``: Lanczos exponentials of the one- and zero-site
effective Hamiltonians, blocked QR/RQ gauge steps, environment updates and
bond absorption; only quantum-number bookkeeping and the Lanczos scalars
touch the host.  Environments from the previous time step are reused: the
reference rebuilds all of them every step (mps.py:1283-1285), but the left
environments it rebuilds are bit-for-bit the ones the previous
left-to-right half-sweep produced (the final normalisation only rescales the
last site), and the rebuilt right environments are overwritten before use.
"""
import logging
import threading
from functools import wraps
from typing import Dict, List, Union

import numpy as np
import torch
import scipy.linalg
from scipy import stats

from renormalizer_amd import native
from renormalizer_amd.model import Model, Op, OpSum
from renormalizer_amd.model import basis as ba
from renormalizer_amd.mps.lib import Environ
from renormalizer_amd.mps.mp import MatrixProduct, _is_dev
from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.svd_qn import add_outer, get_qn_mask, select_basis, svd_qn_device, svd_qn_plan
from renormalizer_amd.utils import (CompressConfig, EvolveConfig, EvolveMethod, OptimizeConfig)

logger = logging.getLogger(__name__)

# environment / operator precision of each mode: "mixed" keeps FP64
# environments and switches only the tail of each Lanczos iteration to FP32
# (native.expm_krylov_mixed_async); "fp32" runs every contraction on FP32
# matrix cores with complex64 environments
_PREC = {"fp64": native.PREC_FP64, "mixed": native.PREC_FP64, "fp32": native.PREC_MIXED}


def _min_abs(t1, t2):
    return t1 if abs(t1) < abs(t2) else t2


def adaptive_tdvp(fun):
    """Step-size control by comparing one dt step with two dt/2 steps
    (reference mps.py:46-115)."""

    @wraps(fun)
    def adaptive_fun(self: "Mps", mpo, evolve_target_t):
        if not self.evolve_config.adaptive:
            return fun(self, mpo, evolve_target_t)
        config: EvolveConfig = self.evolve_config.copy()
        config.check_valid_dt(evolve_target_t)
        cur_mps = self
        del self
        p_restart, p_min, p_max = 0.5, 0.1, 2.0
        evolved_t = 0
        while True:
            dt = _min_abs(config.guess_dt, evolve_target_t - evolved_t)
            mps_half1 = fun(cur_mps, mpo, dt / 2)
            mps_half2 = fun(mps_half1, mpo, dt / 2)
            mps = fun(cur_mps, mpo, dt)
            dis = mps.distance(mps_half2)
            del mps_half1, mps
            p = (0.75 * config.adaptive_rtol / (dis / mps_half2.mp_norm + 1e-30)) ** (1.0 / 3)
            p = min(max(p, p_min), p_max)
            if p < p_restart:
                config.guess_dt = dt * p
                continue
            evolved_t += dt
            if np.allclose(evolved_t, evolve_target_t):
                mps_half2.evolve_config.guess_dt = config.guess_dt
                return mps_half2
            config.guess_dt *= p
            cur_mps = mps_half2

    return adaptive_fun



def _describe(local_steps):
    """scipy.stats.describe of the Krylov matvec counts (reference
    mps.py:1400-1402); equal counts make its skewness undefined, which scipy
    reports as a precision-loss warning."""
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        return stats.describe(local_steps)

def _fp32_matvecs(counts, switches):
    """Matvecs of a step that ran on the FP32 operator: per Krylov call, the
    iterations from its switch point on."""
    if not switches:
        return 0
    return int(sum(max(0, c - max(1, switches.get(k, native.SWITCH_NEVER))) for k, c in counts))


class Mps(MatrixProduct):
    def __init__(self):
        super().__init__()
        self.coeff = 1
        self.optimize_config: OptimizeConfig = OptimizeConfig()
        self.evolve_config: EvolveConfig = EvolveConfig()
        self._env_cache = None

    # ---- constructors (reference mps.py:187-350) -------------------------
    @classmethod
    def _empty(cls, model, n):
        mps = cls()
        mps.model = model
        mps._mp = [None] * n
        return mps

    @classmethod
    def hartree_product_state(cls, model, condition: Dict = None, qn_idx: int = None):
        condition = {} if condition is None else dict(condition)
        mps = cls._empty(model, model.nsite)
        qn_size = model.qn_size
        mps.qn = [np.zeros((1, qn_size), dtype=int)]
        index = [model.dof_to_siteidx[k] for k in condition]
        assert len(index) == len(set(index))
        condition = {model.dof_to_siteidx[k]: v for k, v in condition.items()}
        for isite, local_basis in enumerate(model.basis):
            ms = np.zeros((1, local_basis.nbas, 1))
            local_state = condition.pop(isite, 0)
            if isinstance(local_state, int):
                ms[0, local_state, 0] = 1.0
                qn = local_basis.sigmaqn[local_state]
            else:
                ms[0, :, 0] = local_state
                all_qn = np.array(local_basis.sigmaqn)[np.nonzero(local_state)]
                if not np.allclose(all_qn.std(axis=0), 0):
                    raise ValueError("Quantum numbers are mixed in the condition.")
                qn = all_qn[0]
            mps._mp[isite] = ms
            mps.qn.append(mps.qn[-1] + np.array(qn).reshape(1, qn_size))
        if condition:
            raise ValueError(f"Condition not completely used: {condition}")
        mps.qntot = mps.qn[-1][0]
        mps.qnidx = model.nsite
        mps.move_qnidx(model.nsite - 1 if qn_idx is None else qn_idx)
        mps.to_right = False
        return mps

    @classmethod
    def ground_state(cls, model: Model, max_entangled: bool, normalize: bool = True, condition: Dict = None):
        mps = cls._empty(model, model.nsite)
        mps.qn = [np.zeros((1, model.qn_size), dtype=int)] * (model.nsite + 1)
        mps.qnidx = model.nsite - 1
        mps.to_right = False
        mps.qntot = np.zeros(model.qn_size, dtype=int)
        if condition is not None:
            raise NotImplementedError("multi-electron bases are outside the TDVP-PS hot path")
        for isite, local_basis in enumerate(model.basis):
            pdim = local_basis.nbas
            ms = np.zeros((1, pdim, 1))
            if local_basis.is_phonon:
                if max_entangled:
                    ms[0, :, 0] = 1.0 / np.sqrt(pdim) if normalize else 1.0
                else:
                    ms[0, 0, 0] = 1.0
            elif isinstance(local_basis, ba.BasisSimpleElectron):
                ms[0, 0, 0] = 1.0
            elif isinstance(local_basis, ba.BasisHalfSpin):
                if max_entangled:
                    ms[0, :, 0] = 1.0 / np.sqrt(2.0) if normalize else 1.0
                else:
                    ms[0, 0, 0] = 1.0
            else:
                raise NotImplementedError(type(local_basis))
            mps._mp[isite] = ms
        return mps

    @classmethod
    def random(cls, model: Model, qntot, m_max, percent=1.0) -> "Mps":
        """Random left-canonical MPS in the sector ``qntot`` (reference
        mps.py:120-185): on every bond, random orthonormal vectors inside each
        allowed quantum-number sector (eigenvectors of random symmetric
        blocks) with ``select_basis`` choosing the kept ones; the last site is
        random and normalised.  Host tensors, numpy's global generator."""
        mps = cls()
        mps.model = model
        if isinstance(qntot, (int, np.integer)):
            qntot = np.array([qntot])
        qntot = np.asarray(qntot)
        qn_size = len(qntot)
        assert qn_size == model.qn_size
        mps.qn = [np.zeros((1, qn_size), dtype=int)]
        mps._mp = []
        dim_list = [1]
        for imps in range(model.nsite - 1):
            qnbig = add_outer(mps.qn[imps], np.array(model.basis[imps].sigmaqn)).reshape(-1, qn_size)
            u_set, s_set, qnset = [], [], []
            for iblock in set([tuple(t) for t in qnbig]):
                if np.all(qntot < np.array(iblock)):
                    continue
                indices = [i for i, x in enumerate(qnbig) if tuple(x) == iblock]
                a = np.random.random([len(indices), len(indices)]) - 0.5
                a = a + a.T
                s, u = scipy.linalg.eigh(a=a)
                rec = np.zeros((len(qnbig), len(indices)))
                rec[indices, :] = u
                u_set.append(rec)
                s_set.append(s)
                qnset += [iblock] * len(indices)
            u_set = np.concatenate(u_set, axis=1)
            s_set = np.concatenate(s_set)
            m_max2 = m_max[imps + 1] if isinstance(m_max, (list, tuple, np.ndarray)) else m_max
            sidx, mpsqn, _ = select_basis(u_set, s_set, qnset, None, m_max2, percent=percent)
            mt = u_set[:, sidx]
            dim_list.append(len(sidx))
            mps._mp.append(mt.reshape((dim_list[imps], -1, dim_list[imps + 1])))
            mps.qn.append(mpsqn.reshape(len(sidx), qn_size))
        mps.qn.append(np.zeros((1, qn_size), dtype=int))
        dim_list.append(1)
        last_mt = np.random.random([dim_list[-2], model.basis[-1].nbas, dim_list[-1]]) - 0.5
        qnmask = get_qn_mask(add_outer(mps.qn[-2], np.array(model.basis[-1].sigmaqn)), qntot)
        last_mt[~qnmask] = 0
        last_mt /= np.linalg.norm(last_mt.ravel())
        mps._mp.append(last_mt)
        mps.qnidx = len(mps._mp) - 1
        mps.to_right = False
        mps.qntot = qntot
        mps.dtype = np.float64
        return mps

    # ---- bookkeeping -------------------------------------------------------
    def _get_sigmaqn(self, idx):
        return self.model.basis[idx].sigmaqn

    @property
    def is_mps(self):
        return True

    @property
    def is_mpo(self):
        return False

    @property
    def is_mpdm(self):
        return False

    @property
    def nexciton(self):
        return self.qntot

    def _invalidate(self):
        self._env_cache = None

    def metacopy(self) -> "Mps":
        new = super().metacopy()
        new.coeff = self.coeff
        new.optimize_config = self.optimize_config.copy()
        new.evolve_config = self.evolve_config.copy()
        new._env_cache = None
        return new

    def to_complex(self, inplace=False) -> "Mps":
        new = super().to_complex(inplace=inplace)
        new.coeff = complex(new.coeff)
        return new

    # ---- norms -----------------------------------------------------------
    def _device(self):
        return self._mp[0].device if self.on_device else torch.device("cuda", torch.cuda.current_device())

    def _device_sites(self):
        """Complex128 device tensors of the sites (uploading host tensors)."""
        dev = self._device()
        out = []
        for m in self._mp:
            if _is_dev(m):
                out.append(m if m.dtype == torch.complex128 else m.to(torch.complex128))
            else:
                out.append(torch.from_numpy(np.ascontiguousarray(m, dtype=np.complex128)).to(dev))
        return out

    def dot(self, other: "MatrixProduct") -> complex:
        """sum of products of the site tensors without conjugation: use
        ``a.conj().dot(b)`` for <a|b> (reference mp.py:933-956)."""
        assert len(self) == len(other)
        if not self.on_device and not other.on_device:
            return self.dot_host(other)
        bra = self._device_sites()
        ket = other._device_sites()
        e = torch.ones((1, 1, 1), dtype=torch.complex128, device=ket[0].device)
        for b, k in zip(bra, ket):
            e = native.env_update(e, k, _identity_w(k.shape[1], k.device), "L", ms_conj=b)
        return complex(e.reshape(-1)[0].item())

    @property
    def mp_norm(self) -> float:
        res = self.conj().dot(self).real
        if res < 0:
            assert abs(res) < 1e-8
            res = 0
        return float(np.sqrt(res))

    @property
    def norm(self):
        return np.linalg.norm(self.coeff) * self.mp_norm

    def distance(self, other) -> float:
        l1 = self.conj().dot(self)
        l2 = other.conj().dot(other)
        l12 = self.conj().dot(other)
        d2 = (l1 + l2 - l12 - l12.conjugate()).real
        if d2 < 0:
            assert d2 / l1.real < 1e-8
            return 0.0
        return float(np.sqrt(d2))

    def scale(self, val, inplace=False):
        new = self if inplace else self.copy()
        if np.iscomplex(val):
            new.to_complex(inplace=True)
        else:
            val = val.real
        m = new._mp[self.qnidx]
        new._mp[self.qnidx] = m * val
        new._invalidate()
        return new

    def normalize(self, kind):
        """reference mps.py:2025-2059."""
        nrm = self.mp_norm
        if kind == "mps_only":
            new_coeff = self.coeff
        elif kind == "mps_and_coeff":
            new_coeff = self.coeff / np.linalg.norm(self.coeff)
        elif kind == "mps_norm_to_coeff":
            new_coeff = self.coeff * nrm
        else:
            raise ValueError(f"kind={kind} is not valid.")
        m = self._mp[self.qnidx]
        self._mp[self.qnidx] = m * (1.0 / nrm)
        self.coeff = new_coeff
        return self

    # ---- initial-state expansion (host; reference mps.py:1934-2023) ------
    def expand_bond_dimension(self, hint_mpo=None, coef=1e-10, include_ex=True):
        return expand_bond_dimension(self, hint_mpo, coef, include_ex)

    # ---- time evolution ---------------------------------------------------
    def evolve(self, mpo, evolve_dt, normalize=True) -> "Mps":
        method = {
            EvolveMethod.tdvp_ps: self._evolve_tdvp_ps,
            EvolveMethod.tdvp_ps2: self._evolve_tdvp_ps2,
        }.get(self.evolve_config.method)
        if method is None:
            raise NotImplementedError(
                f"{self.evolve_config.method} is not on the device path; use EvolveMethod.tdvp_ps")
        new_mps = method(mpo, evolve_dt)
        if normalize:
            new_mps.normalize("mps_and_coeff" if np.iscomplex(evolve_dt) else "mps_only")
        return new_mps

    @adaptive_tdvp
    def _evolve_tdvp_ps(self, mpo: Mpo, evolve_dt) -> "Mps":
        """One-site TDVP projector splitting, two half-sweeps (reference
        mps.py:1267-1404), on the device."""
        if self.evolve_config.ivp_solver != "krylov":
            raise NotImplementedError("only the Krylov propagator is on the device path")
        prec = _PREC[self.evolve_config.precision]
        n = len(self)
        cache = self._env_cache
        # a step starts at one end of the chain and returns there: right to
        # left from the last site (left environments kept) or left to right
        # from the first (right environments kept)
        at_end = self.qnidx == (0 if self.to_right else n - 1)
        reuse = cache is not None and cache[0] is mpo and cache[1] == prec and at_end \
            and cache[4] == self.to_right
        # the sweep rewrites the cached environments in place: take them away
        # from self so that a second evolution of self rebuilds its own
        self._env_cache = None
        budgets = None
        wide = True
        if reuse and self.evolve_config.krylov_async and cache[3] is not None \
                and cache[3][0] == complex(evolve_dt):
            budgets = cache[3][1]
            # counts of a step run with host waits (the first one, or a redo)
            # move more than those of consecutive asynchronous steps
            wide = cache[3][2]
        switches = cache[3][3] if budgets is not None else None
        adaptive = self.evolve_config.precision == "mixed"
        mps, environ, counts, switch = self._tdvp_ps_sweep(
            mpo, evolve_dt, prec, cache[2] if reuse else None, budgets, wide,
            switches if adaptive else None, self.evolve_config.mixed_tau if adaptive else None)
        synchronous = budgets is None
        if counts is None:
            # an iteration budget was too small: redo the step with host waits
            mps, environ, counts, switch = self._tdvp_ps_sweep(mpo, evolve_dt, prec, None, None)
            synchronous = True
        local_steps = [c for _, c in counts]
        mps.evolve_config.stat = _describe(local_steps)
        mps.evolve_config.local_steps = local_steps
        mps.evolve_config.fp32_matvecs = _fp32_matvecs(counts, switches if adaptive and not synchronous
                                                       else None)
        mps._env_cache = (mpo, prec, environ, (complex(evolve_dt), dict(counts), synchronous, switch),
                          mps.to_right)
        return mps

    @adaptive_tdvp
    def _evolve_tdvp_ps2(self, mpo: Mpo, evolve_dt) -> "Mps":
        """Two-site TDVP projector splitting with SVD truncation, two
        half-sweeps (reference mps.py:1406-1517), on the device: per bond the
        two-site Krylov step (-i dt/2) with the merged two-site MPO, the
        truncation by the blocked SVD and basis selection of the compress
        config (mp.py:651-780, host LAPACK as in the reference), the
        environment update, and the one-site step (+i dt/2) of the site left
        behind with the device QR that moves the centre on (mp.py:890-908).
        The bond dimension can grow up to the compress config's limit."""
        if self.evolve_config.ivp_solver != "krylov":
            raise NotImplementedError("only the Krylov propagator is on the device path")
        from renormalizer_amd.mps.gs import _merge_w, _update_mps
        prec = _PREC[self.evolve_config.precision]
        mps = self.copy() if np.iscomplex(evolve_dt) else self.to_complex()
        if not mps.on_device:
            mps.to_device()
        dev = mps._mp[0].device
        wt = mpo.device_tensors(prec, dev)
        rdt = torch.float64 if prec == native.PREC_FP64 else torch.float32
        n = len(mps)
        environ = Environ(mps._mp, wt, None, prec=prec)
        w2_cache = {}
        local_steps = []
        for _ in range(2):
            for imps in mps.iter_idx_list(full=False):
                if mps.to_right:
                    lidx, c0, c1, ridx = imps - 1, imps, imps + 1, imps + 2
                    c2, last = c1, n - 2
                else:
                    lidx, c0, c1, ridx = imps - 2, imps - 1, imps, imps + 1
                    c2, last = c0, 1
                l_array = environ.read("L", lidx)
                r_array = environ.read("R", ridx)
                ms2 = torch.tensordot(mps._mp[c0], mps._mp[c1], dims=1)
                shape2 = list(ms2.shape)
                if (c0, c1) not in w2_cache:
                    w2 = _merge_w(np.asarray(mpo[c0], dtype=np.float64), np.asarray(mpo[c1], dtype=np.float64))
                    w2_cache[(c0, c1)] = torch.from_numpy(w2).to(device=dev, dtype=rdt).contiguous()
                fused = ms2.reshape(shape2[0], shape2[1] * shape2[2], shape2[3]).contiguous()
                mps_t, j = native.expm_krylov(l_array, w2_cache[(c0, c1)], r_array, fused,
                                              -1j * evolve_dt / 2, prec)
                local_steps.append(j)
                qnbigl, qnbigr, _ = mps._get_big_qn([c0, c1], with_qnmat=False)
                _update_mps(mps, mps_t.reshape(shape2), [c0, c1], qnbigl, qnbigr, 0)
                if imps == last:
                    continue
                if mps.to_right:
                    l_array = environ.GetLR("L", lidx + 1, mps._mp, wt, itensor=l_array, method="System")
                else:
                    r_array = environ.GetLR("R", ridx - 1, mps._mp, wt, itensor=r_array, method="System")
                site = mps._mp[c2].contiguous()
                shape1 = list(site.shape)
                mps_t, j = native.expm_krylov(l_array, wt[c2], r_array, site, 1j * evolve_dt / 2, prec)
                local_steps.append(j)
                # move the canonical centre on (_push_cano) with the device QR
                qnbigl, qnbigr, _ = mps._get_big_qn([c2], with_qnmat=False)
                system = "L" if mps.to_right else "R"
                u, qnlset, vt, qnrset = svd_qn_device(mps_t, qnbigl, qnbigr, mps.qntot, system)
                if mps.to_right:
                    kt = u.shape[1]
                    mps._mp[c2] = u.reshape(shape1[:-1] + [kt])
                    nxt = mps._mp[c2 + 1]
                    mps._mp[c2 + 1] = native.zgemm(vt, nxt.reshape(nxt.shape[0], -1)).reshape(
                        [kt] + list(nxt.shape[1:]))
                    mps.qn[c2 + 1] = np.array(qnlset, dtype=int).reshape(kt, -1)
                    mps.qnidx = c2 + 1
                else:
                    kt = vt.shape[0]
                    mps._mp[c2] = vt.reshape([kt] + shape1[1:])
                    prev = mps._mp[c2 - 1]
                    mps._mp[c2 - 1] = native.zgemm(prev.reshape(-1, prev.shape[-1]), u).reshape(
                        list(prev.shape[:-1]) + [kt])
                    mps.qn[c2] = np.array(qnrset, dtype=int).reshape(kt, -1)
                    mps.qnidx = c2 - 1
            mps._switch_direction()
        mps._invalidate()
        mps.evolve_config.stat = _describe(local_steps)
        mps.evolve_config.local_steps = local_steps
        return mps

    # number of time steps redone with host waits because an asynchronous
    # Krylov budget proved too small (diagnostics)
    krylov_redo_count = 0

    def _tdvp_ps_sweep(self, mpo, evolve_dt, prec, environ, budgets, wide=True, switches=None,
                       tau=None):
        """The sweep itself.  With ``budgets`` (matvec counts of the previous
        step keyed by position) every Krylov call is queued asynchronously and
        the counts are read once at the end; returns counts None when a
        budget ran out.  With ``tau`` (the mixed mode, asynchronous steps
        only) each Krylov call switches to FP32 matvecs at the iteration
        ``switches`` holds for its position (all FP64 where it holds none)
        and measures the switch point for the next step; returns (mps,
        environ, counts, measured switch points or None)."""
        mps = self.to_complex()
        if not mps.on_device:
            mps.to_device()
        dev = mps._mp[0].device
        wt = mpo.device_tensors(prec, dev)
        n = len(mps)
        if environ is None:
            # only the left environments are read before being rebuilt
            environ = Environ(mps._mp, wt, "L" if not mps.to_right else "R", prec=prec)
        use_async = budgets is not None
        adaptive = use_async and tau is not None
        # tile autotuning of the matvec GEMMs (once per operator shape)
        tune = use_async and getattr(self.evolve_config, "autotune", False)
        if adaptive:
            wt32 = mpo.device_tensors(native.PREC_MIXED, dev)
        if use_async:
            info = torch.zeros((4 * n + 4, 3), dtype=torch.int32, device=dev)
        keys = []
        counts = []

        def krylov(key, l_array, w, r_array, v, dt):
            if not use_async:
                out, j = native.expm_krylov(l_array, w, r_array, v, dt, prec)
                counts.append((key, j))
                return out
            # the previous step's count at the same place plus a margin: between
            # consecutive asynchronous steps the counts move by 2 at the 99th
            # percentile and by 6 at most (spin-boson, M = 256); after a step
            # with host waits they move more
            prev = budgets.get(key, 0)
            budget = prev + (prev // 4 + 8 if wide else prev // 16 + 6)
            budget = max(budget, 12)
            if tune:
                native.autotune(l_array, w, r_array, v, prec)
            if adaptive:
                js = switches.get(key, native.SWITCH_NEVER) if switches else native.SWITCH_NEVER
                w32 = None if w is None else wt32[key[1]]
                l32, r32 = l_array.to(torch.complex64), r_array.to(torch.complex64)
                if tune:
                    native.autotune(l32, w32, r32, v, native.PREC_MIXED)
                out = native.expm_krylov_mixed_async(
                    l_array, w, r_array, l32, w32, r32, v, dt, budget, js, tau, info[len(keys)])
            else:
                out = native.expm_krylov_async(l_array, w, r_array, v, dt, prec, budget, info[len(keys)])
            keys.append(key)
            return out

        for half in range(2):
            for imps in mps.iter_idx_list(full=True):
                system = "L" if mps.to_right else "R"
                l_array = environ.read("L", imps - 1)
                r_array = environ.read("R", imps + 1)
                site = mps._mp[imps]
                shape = list(site.shape)
                # the quantum-number bookkeeping of the gauge step does not
                # depend on the evolved tensor: do it before queueing the
                # Krylov call, so the QR is queued right after it
                qnbigl, qnbigr, _ = mps._get_big_qn([imps], with_qnmat=False)
                plan = svd_qn_plan(qnbigl, qnbigr, mps.qntot)
                mps_t = krylov((half, imps, 1), l_array, wt[imps], r_array, site, -1j * evolve_dt / 2)
                u, qnlset, vt, qnrset = svd_qn_device(mps_t, qnbigl, qnbigr, mps.qntot, system, plan)
                if not mps.to_right and imps != 0:
                    kt = vt.shape[0]
                    mps._mp[imps] = vt.reshape([kt] + shape[1:])
                    mps.qn[imps] = np.array(qnrset, dtype=int).reshape(kt, -1)
                    mps.qnidx = imps - 1
                    r_array = environ.GetLR("R", imps, mps._mp, wt, itensor=r_array, method="System")
                    u_t = krylov((half, imps, 0), l_array, None, r_array, u, 1j * evolve_dt / 2)
                    prev = mps._mp[imps - 1]
                    mps._mp[imps - 1] = native.zgemm(prev.reshape(-1, prev.shape[-1]), u_t).reshape(
                        list(prev.shape[:-1]) + [kt])
                elif mps.to_right and imps != n - 1:
                    kt = u.shape[1]
                    mps._mp[imps] = u.reshape(shape[:-1] + [kt])
                    mps.qn[imps + 1] = np.array(qnlset, dtype=int).reshape(kt, -1)
                    mps.qnidx = imps + 1
                    l_array = environ.GetLR("L", imps, mps._mp, wt, itensor=l_array, method="System")
                    vt_t = krylov((half, imps, 0), l_array, None, r_array, vt, 1j * evolve_dt / 2)
                    nxt = mps._mp[imps + 1]
                    mps._mp[imps + 1] = native.zgemm(vt_t, nxt.reshape(nxt.shape[0], -1)).reshape(
                        [kt] + list(nxt.shape[1:]))
                else:
                    mps._mp[imps] = mps_t
            mps._switch_direction()
        if use_async:
            got = info[:len(keys)].cpu().numpy()
            # a call that ran out of budget left its output unwritten, so
            # everything after it is void: redo (checked before the zero-vector
            # flag, which such garbage can raise downstream)
            if (got[:, 1] == 0).any():
                Mps.krylov_redo_count += 1
                logger.info("Krylov budget exhausted at %s; redoing the step with host waits",
                            [k for k, g in zip(keys, got) if g[1] == 0][:4])
                return mps, environ, None, None
            if (got[:, 1] < 0).any():
                raise AssertionError("zero vector handed to expm_krylov (krylov.py:44)")
            counts = [(k, int(c)) for k, c in zip(keys, got[:, 0])]
            if adaptive:
                return mps, environ, counts, {k: int(x) for k, x in zip(keys, got[:, 2])}
        return mps, environ, counts, None

    # ---- observables --------------------------------------------------------
    def expectation(self, mpo: Union[Mpo, Op, OpSum], self_conj: "Mps" = None):
        """<psi|O|psi> (reference mps.py:471-525), FP64 on the device."""
        if isinstance(mpo, (Op, OpSum)):
            mpo = Mpo(self.model, mpo)
        ket = self._device_sites()
        bra = [b.conj().resolve_conj() for b in ket] if self_conj is None else self_conj._device_sites()
        wt = mpo.device_tensors(native.PREC_FP64, ket[0].device)
        e = torch.ones((1, 1, 1), dtype=torch.complex128, device=ket[0].device)
        for k, b, w in zip(ket, bra, wt):
            e = native.env_update(e, k, w, "L", ms_conj=b)
        val = complex(e.reshape(-1)[0].item())
        if np.isclose(val.imag, 0):
            return float(val.real)
        return val

    def expectations(self, mpos, self_conj: "Mps" = None, opt: bool = True) -> np.ndarray:
        """Several expectation values (reference mps.py:527-575).  Operators
        that act as the identity outside a window of sites reuse shared
        identity environments on both sides of the window."""
        mpos = [Mpo(self.model, m) if isinstance(m, (Op, OpSum)) else m for m in mpos]
        if not opt or self_conj is not None:
            res = np.array([self.expectation(m, self_conj) for m in mpos])
        else:
            ket = self._device_sites()
            dev = ket[0].device
            windows = [_identity_window(m) for m in mpos]
            n = len(ket)
            lid, rid = _identity_environments(ket, dev)
            vals = []
            for m, (lo, hi) in zip(mpos, windows):
                wt = m.device_tensors(native.PREC_FP64, dev)
                e = lid[lo - 1]
                for i in range(lo, hi + 1):
                    e = native.env_update(e, ket[i], wt[i], "L")
                r = rid[hi + 1] if hi + 1 < n else torch.ones((1, 1, 1), dtype=torch.complex128, device=dev)
                vals.append(complex(native.env_close(e, r).item()))
            res = np.array(vals)
        if np.allclose(np.imag(res), 0):
            return np.real(res)
        return res

    @property
    def e_occupations(self):
        key = "e_occupations"
        if key not in self.model.mpos:
            self.model.mpos[key] = [Mpo(self.model, Op(r"a^\dagger a", dof)) for dof in self.model.e_dofs]
        return self.expectations(self.model.mpos[key])

    @property
    def ph_occupations(self):
        key = "ph_occupations"
        if key not in self.model.mpos:
            self.model.mpos[key] = [Mpo(self.model, Op("n", dof)) for dof in self.model.v_dofs]
        return self.expectations(self.model.mpos[key])

    def calc_1site_rdm(self, idx=None) -> Dict[int, np.ndarray]:
        """One-site reduced density matrices rho[i, j] = <i|Tr_rest|psi><psi||j>
        (reference mps.py:1547-1598), FP64 on the device."""
        ket = self._device_sites()
        dev = ket[0].device
        n = len(ket)
        if idx is None:
            idx = list(range(n))
        elif isinstance(idx, int):
            idx = [idx]
        else:
            idx = list(idx)
        lid, rid = _identity_environments(ket, dev, upto=max(idx))
        rdm = {}
        for i in idx:
            d = ket[i].shape[1]
            x = native.env_update(lid[i - 1], ket[i], _pair_w(d, dev), "L")
            r = rid[i + 1] if i + 1 < n else torch.ones((1, 1, 1), dtype=torch.complex128, device=dev)
            rdm[i] = native.env_close(x, r).reshape(d, d).cpu().numpy()
        return rdm

    def calc_bond_singular_values(self) -> np.ndarray:
        """Singular values on every bond (reference mps.py:1759-1773): a host
        copy made right-canonical, then one left-to-right SVD sweep without
        truncation; rows padded with zeros."""
        mps = self.copy().to_host()
        mps.ensure_right_canonical()
        _, s_array = mps.compress(temp_m_trunc=np.iinfo(np.int32).max, ret_s=True)
        return s_array

    def calc_entropy(self, entropy_type: str):
        """reference mps.py:1689-1757 for entropy_type "bond" (the only one
        on this path): the von Neumann entropy of every bond."""
        if entropy_type != "bond":
            raise NotImplementedError(f"entropy type {entropy_type!r}: only 'bond' is provided")
        return self.calc_bond_entropy()

    def calc_bond_entropy(self, s_array: np.ndarray = None) -> np.ndarray:
        """Von Neumann entropy -sum p ln p, p = s^2 / sum s^2, on every bond
        (reference mps.py:1775-1793, utils.py:41-48)."""
        if s_array is None:
            s_array = self.calc_bond_singular_values()
        out = []
        for sigma in s_array:
            p = np.asarray(sigma, dtype=float) ** 2
            p = p / p.sum()
            p = p[p > 0]
            out.append(float(-(p * np.log(p)).sum()))
        return np.array(out)

    # ---- on-disk format (reference mp.py:1085-1113, mps.py:352-386) -------
    def dump(self, fname):
        """Protocol "0.4" (reference mp.py:1085-1113): ``mt_<i>`` site
        arrays, ``subqn_<i>`` bond quantum numbers, ``qn`` (the same as an
        object array, which the reference loader reads), ``qnidx``,
        ``qntot``, ``to_right``, ``coeff``, ``version``, ``nsites``."""
        data = {"version": "0.4", "nsites": len(self)}
        for i, mt in enumerate(self.numpy_sites()):
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
    def load(cls, model: Model, fname: str):
        """Reads files of protocol "0.4" (reference mp.py:1085-1113,
        mps.py:352-386) without unpickling: the bond quantum numbers come from
        the ``subqn_<i>`` arrays that protocol writes beside its pickled
        ``qn``.  Older protocols ("0.1", "0.2", "0.3") keep the quantum
        numbers only in the pickled object array and are refused with a
        ValueError, as is an unknown version."""
        with np.load(fname, allow_pickle=False) as f:
            version = str(f["version"])
            if version not in ("0.1", "0.2", "0.3", "0.4"):
                raise ValueError(f"Unknown dump version: {version}")
            n = int(f["nsites"])
            missing = [i for i in range(n + 1) if f"subqn_{i}" not in f.files]
            if missing:
                raise ValueError(
                    f"dump protocol {version} file without subqn_<i> arrays: its quantum numbers "
                    f"are stored only pickled, which is not read")
            mps = cls()
            mps.model = model
            mps._mp = [f[f"mt_{i}"] for i in range(n)]
            mps.dtype = np.complex128 if any(np.iscomplexobj(m) for m in mps._mp) else np.float64
            mps.qn = [f[f"subqn_{i}"].astype(int).reshape(-1, model.qn_size) for i in range(n + 1)]
            mps.qnidx = int(f["qnidx"])
            mps.qntot = f["qntot"].astype(int)
            if version == "0.1":
                mps.to_right = bool(f["left"])
                mps.coeff = 1
            elif version in ("0.3", "0.4"):
                mps.to_right = bool(f["to_right"])
                mps.coeff = f["coeff"].item(0)
            else:
                raise ValueError("dump protocol 0.2 stores its coefficient pickled; not read")
        return mps

    def __str__(self):
        return f"Mps(nsite={len(self)}, bond dims={self.bond_dims})"


_ID_CACHE = {}
_ID_LOCK = threading.Lock()


def _identity_w(d, dev):
    key = ("id", d, str(dev))
    with _ID_LOCK:
        if key not in _ID_CACHE:
            _ID_CACHE[key] = torch.eye(d, dtype=torch.float64, device=dev).reshape(1, d, d, 1).contiguous()
        return _ID_CACHE[key]


def _pair_w(d, dev):
    """W[0, i', j', i*d + j] = delta(i', i) delta(j', j): the environment
    update with it keeps the (bra, ket) physical pair open."""
    key = ("pair", d, str(dev))
    with _ID_LOCK:
        if key not in _ID_CACHE:
            w = np.zeros((1, d, d, d * d))
            for i in range(d):
                for j in range(d):
                    w[0, i, j, i * d + j] = 1.0
            _ID_CACHE[key] = torch.from_numpy(w).to(dev)
        return _ID_CACHE[key]


def _identity_environments(ket, dev, upto=None):
    """Identity-operator environments: lid[i] folds sites 0..i from the
    left, rid[i] folds sites i..N-1 from the right (sentinels at -1 / N)."""
    n = len(ket)
    one = torch.ones((1, 1, 1), dtype=torch.complex128, device=dev)
    lid = {-1: one}
    last = n - 1 if upto is None else upto
    for i in range(0, last):
        lid[i] = native.env_update(lid[i - 1], ket[i], _identity_w(ket[i].shape[1], dev), "L")
    rid = {n: one}
    for i in range(n - 1, 0, -1):
        rid[i] = native.env_update(rid[i + 1], ket[i], _identity_w(ket[i].shape[1], dev), "R")
    return lid, rid


def _identity_window(mpo: Mpo):
    """[lo, hi] such that every site outside is an identity with unit MPO
    bonds; the whole chain if there is no such window."""
    n = len(mpo)
    is_id = []
    for m in mpo:
        m = np.asarray(m)
        is_id.append(m.shape[0] == 1 and m.shape[3] == 1 and np.array_equal(m[0, :, :, 0], np.eye(m.shape[1])))
    lo = 0
    while lo < n - 1 and is_id[lo]:
        lo += 1
    hi = n - 1
    while hi > lo and is_id[hi]:
        hi -= 1
    return lo, hi


# ---------------------------------------------------------------------------
# initial-state expansion (host tensors; reference mps.py:1934-2023 and
# lib.py:417-439)


def compressed_sum(mps_list, batchsize=5, temp_m_trunc=None):
    from collections import deque
    from functools import reduce

    assert len(mps_list) != 0
    queue = deque(mps_list)
    if len(queue) == 1:
        new = mps_list[0].canonicalise()
        new.compress(temp_m_trunc=temp_m_trunc)
        return new
    while len(queue) != 1:
        terms = [queue.popleft() for _ in range(min(batchsize, len(queue)))]
        s = reduce(lambda a, b: a.add(b), terms)
        s.canonicalise()
        s.compress(temp_m_trunc=temp_m_trunc)
        queue.append(s)
    return queue[0]


def expand_bond_dimension(mps: Mps, hint_mpo=None, coef=1e-10, include_ex=True):
    """Fill the bond space up to the configured bond dimension with states
    generated by repeated application of ``hint_mpo``, weighted by ``coef``."""
    if hint_mpo is not None and include_ex:
        ex_state = Mps.ground_state(mps.model, False)
        assert mps.model.qn_size == 1
        for _ in range(int(mps.qntot[0])):
            ex_state = Mpo.onsite(mps.model, r"a^\dagger") @ ex_state
        ex_state.compress_config = mps.compress_config
        ex_state.move_qnidx(mps.qnidx)
        ex_state.to_right = mps.to_right
        if mps.on_device:
            from renormalizer_amd.mps.device_setup import upload_complex
            upload_complex(ex_state, mps[0].device)
    else:
        ex_state = None
    return expand_bond_dimension_general(mps, hint_mpo, coef, ex_state)


def expand_bond_dimension_general(mps, hint_mpo=None, coef=1e-10, ex_mps=None):
    """Reference mps.py:1963-2023.  Runs where the state is: a device state
    (complex128 sites) is expanded with the device operations of
    device_setup.py, a host state with the numpy ones of mp.py."""
    if mps.on_device:
        from renormalizer_amd.mps.device_setup import upload_complex
        mps = upload_complex(mps.copy(), mps[0].device)
    mps.compress_config.set_bonddim(len(mps.bond_dims))
    m_target = np.minimum(np.array(mps.compress_config.max_dims) - np.array(mps.bond_dims),
                          mps.bond_dims_exact)
    m_target = np.array(m_target, dtype=int)
    if hint_mpo is None:
        raise NotImplementedError("random expansion (Mps.random) is not provided; pass hint_mpo")
    lastone = mps if ex_mps is None else mps + ex_mps
    expander_list = []
    expander_dims = np.zeros_like(m_target)
    while True:
        lastone = (hint_mpo @ lastone).normalize("mps_and_coeff")
        lastone = lastone.canonicalise().compress(np.max(m_target))
        expander_list.append(lastone)
        expander = compressed_sum(expander_list, temp_m_trunc=m_target)
        if np.all(np.array(expander.bond_dims) >= m_target):
            break
        if np.all(np.array(expander.bond_dims) == expander_dims):
            logger.warning("Expander does not increase anymore. The expand target is too high")
            m_target2 = np.max(m_target - np.array(expander_dims))
            expander2 = (hint_mpo @ lastone).canonicalise().compress(np.maximum(m_target2, 1))
            expander = expander + expander2
            break
        expander_dims = np.array(expander.bond_dims)
        temp_m_trunc = int(np.max(m_target) / np.max(hint_mpo.bond_dims)) + 1
        lastone = lastone.canonicalise().compress(temp_m_trunc)
    return ((mps + expander.scale(coef * mps.norm, inplace=True)).canonicalise()
            .compress(mps.compress_config.max_dims).normalize("mps_norm_to_coeff"))
