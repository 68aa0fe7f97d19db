"""Matrix-product base class: quantum-number bookkeeping shared by Mps and
Mpo (reference mps/mp.py:36-372, 890-1045) and the host-side setup
operations used while preparing an initial state.

Site tensors live either in host memory (numpy; right after construction)
or on the device (torch complex128, after ``to_device``).  The bookkeeping
(qn, qnidx, qntot, to_right) always stays on the host as small int arrays,
exactly as the reference stores it.

The setup algorithms in this file (add, canonicalise, compress) run on
host numpy tensors here and on device tensors in device_setup.py (states
only); the time-evolution and measurement path runs on the device (mps.py,
lib.py).
"""
from typing import List

import numpy as np
import torch

from renormalizer_amd.model import Model
from renormalizer_amd.mps.svd_qn import add_outer, svd_qn_host
from renormalizer_amd.utils import CompressConfig


def _is_dev(t):
    return isinstance(t, torch.Tensor)


class MatrixProduct:
    def __init__(self):
        self._mp: List = []
        self.dtype = np.float64
        self.model: Model = None
        self.compress_config: CompressConfig = CompressConfig()
        self.qn: List[np.ndarray] = []
        self.qnidx: int = None
        self.qntot: np.ndarray = None
        self.to_right: bool = None

    # ---- container -------------------------------------------------------
    def __len__(self):
        return len(self._mp)

    def __getitem__(self, i):
        return self._mp[i]

    def __setitem__(self, i, v):
        self._mp[i] = v
        self._invalidate()

    def __iter__(self):
        return iter(self._mp)

    def append(self, v):
        self._mp.append(v)

    def _invalidate(self):
        pass

    @property
    def site_num(self):
        return len(self._mp)

    @property
    def on_device(self) -> bool:
        return len(self._mp) > 0 and _is_dev(self._mp[0])

    @property
    def is_complex(self):
        return np.issubdtype(np.dtype(self.dtype), np.complexfloating)

    @property
    def bond_dims(self) -> List[int]:
        if not self._mp:
            return []
        return [int(mt.shape[0]) for mt in self._mp] + [int(self._mp[-1].shape[-1])]

    bond_list = vbond_dims = bond_dims

    @property
    def bond_dims_mean(self) -> int:
        return int(round(np.mean(self.bond_dims)))

    @property
    def pbond_dims(self):
        return self.model.pbond_list

    pbond_list = pbond_dims

    @property
    def is_mps(self):
        raise NotImplementedError

    @property
    def is_mpo(self):
        raise NotImplementedError

    @property
    def bond_dims_exact(self) -> np.ndarray:
        """Bond dimensions of an exact factorisation (reference mp.py:130-142)."""
        p = np.array(self.pbond_dims, dtype=float)
        if self.is_mpo or getattr(self, "is_mpdm", False):
            p = p ** 2
        with np.errstate(over="ignore"):
            d1 = [1] + list(np.cumprod(p))
            d2 = ([1] + list(np.cumprod(p[::-1])))[::-1]
        return np.minimum(d1, d2)

    # ---- quantum numbers (reference mp.py:144-352) -----------------------
    def build_empty_qn(self):
        self.qntot = np.array([0] * self.model.qn_size)
        if self.qnidx is None:
            self.qnidx = len(self) - 1
        self.qn = [np.zeros((dim, self.model.qn_size), dtype=int) for dim in self.bond_dims]
        if self.to_right is None:
            self.to_right = False

    def move_qnidx(self, dstidx: int):
        """qn[i] holds L-block numbers for i <= qnidx and R-block numbers
        (qntot - L) for i > qnidx; move that boundary to ``dstidx``."""
        for idx in range(self.qnidx + 1, self.site_num + 1):
            self.qn[idx] = self.qntot - self.qn[idx]
        for idx in range(self.site_num, dstidx, -1):
            self.qn[idx] = self.qntot - self.qn[idx]
        self.qnidx = dstidx

    def iter_idx_list(self, full: bool, stop_idx: int = None):
        if self.to_right:
            last = stop_idx if stop_idx is not None else (self.site_num if full else self.site_num - 1)
            return range(self.qnidx, last)
        last = stop_idx if stop_idx is not None else (-1 if full else 0)
        return range(self.qnidx, last, -1)

    def _switch_direction(self):
        assert self.to_right is not None
        if self.to_right:
            self.qnidx = self.site_num - 1
            self.to_right = False
        else:
            self.qnidx = 0
            self.to_right = True

    def _get_sigmaqn(self, idx):
        raise NotImplementedError

    def _get_big_qn(self, cidx: List[int], with_qnmat: bool = True):
        """Super-block quantum numbers of the active site(s) (reference
        mp.py:308-352).  The third value, the quantum numbers of every
        coefficient (rows x cols), is None unless ``with_qnmat``: at M = 256
        it is a 655 360-entry array, built per site update for nothing on the
        time-evolution path."""
        if len(cidx) == 2:
            cidx = sorted(cidx)
            assert cidx[0] + 1 == cidx[1]
        elif len(cidx) > 2:
            raise AssertionError
        assert self.qnidx in cidx
        sigmaqn = [np.array(self._get_sigmaqn(i)) for i in cidx]
        qnl = np.array(self.qn[cidx[0]])
        qnr = np.array(self.qn[cidx[-1] + 1])
        if len(cidx) == 1:
            if self.to_right:
                qnbigl, qnbigr = add_outer(qnl, sigmaqn[0]), qnr
            else:
                qnbigl, qnbigr = qnl, add_outer(sigmaqn[0], qnr)
        else:
            qnbigl = add_outer(qnl, sigmaqn[0])
            qnbigr = add_outer(sigmaqn[1], qnr)
        return qnbigl, qnbigr, add_outer(qnbigl, qnbigr) if with_qnmat else None

    # ---- copies ----------------------------------------------------------
    def metacopy(self):
        new = self.__class__.__new__(self.__class__)
        new._mp = [None] * len(self)
        new.dtype = self.dtype
        new.model = self.model
        new.compress_config = self.compress_config.copy()
        new.qn = [np.array(q).copy() for q in self.qn]
        new.qnidx = self.qnidx
        new.qntot = None if self.qntot is None else np.array(self.qntot).copy()
        new.to_right = self.to_right
        return new

    def copy(self):
        new = self.metacopy()
        new._mp = [m.clone() if _is_dev(m) else m.copy() for m in self._mp]
        return new

    # ---- host <-> device -------------------------------------------------
    def to_device(self, device=None):
        """Upload the site tensors (complex128 for MPS) in place."""
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self._mp = [m if _is_dev(m) else torch.from_numpy(np.ascontiguousarray(m)).to(device)
                    for m in self._mp]
        return self

    def to_host(self):
        self._mp = [m.cpu().numpy() if _is_dev(m) else m for m in self._mp]
        return self

    def numpy_sites(self) -> List[np.ndarray]:
        return [m.cpu().numpy() if _is_dev(m) else np.asarray(m) for m in self._mp]

    def _require_host(self, what):
        if self.on_device:
            raise RuntimeError(f"{what} is an initial-state preparation step on host tensors; "
                               f"call it before the state is uploaded")

    # ---- setup algorithms (reference mp.py:374-511, 890-1007) -----------
    def add(self, other: "MatrixProduct"):
        """Direct sum of the bond spaces (reference mp.py:374-435)."""
        if self.on_device or other.on_device:
            from renormalizer_amd.mps import device_setup
            return device_setup.add(self, other)
        assert np.all(self.qntot == other.qntot)
        assert self.site_num == other.site_num
        new = self.metacopy()
        if self.is_complex or other.is_complex:
            new.dtype = np.complex128
        new.compress_config.update(self.compress_config)
        dt = new.dtype
        n = self.site_num
        if self.is_mps or getattr(self, "is_mpdm", False):
            new._mp[0] = np.concatenate([self[0], other[0]], axis=2).astype(dt)
            for i in range(1, n - 1):
                a, b = self[i], other[i]
                m = np.zeros((a.shape[0] + b.shape[0], a.shape[1], a.shape[2] + b.shape[2]), dtype=dt)
                m[:a.shape[0], :, :a.shape[2]] = a
                m[a.shape[0]:, :, a.shape[2]:] = b
                new._mp[i] = m
            new._mp[-1] = np.concatenate([self[-1], other[-1]], axis=0).astype(dt)
        else:
            new._mp[0] = np.concatenate([self[0], other[0]], axis=3).astype(dt)
            for i in range(1, n - 1):
                a, b = self[i], other[i]
                m = np.zeros((a.shape[0] + b.shape[0], a.shape[1], a.shape[2], a.shape[3] + b.shape[3]), dtype=dt)
                m[:a.shape[0], :, :, :a.shape[3]] = a
                m[a.shape[0]:, :, :, a.shape[3]:] = b
                new._mp[i] = m
            new._mp[-1] = np.concatenate([self[-1], other[-1]], axis=0).astype(dt)
        new.move_qnidx(other.qnidx)
        new.to_right = other.to_right
        new.qn = [np.concatenate([np.array(q1), np.array(q2)]) for q1, q2 in zip(self.qn, other.qn)]
        new.qn[0] = np.zeros((1, new.qn[0].shape[1]), dtype=int)
        new.qn[-1] = np.zeros((1, new.qn[0].shape[1]), dtype=int)
        return new

    def __add__(self, other):
        return self.add(other)

    def to_complex(self, inplace=False):
        new = self if inplace else self.metacopy()
        new.dtype = np.complex128
        for i, mt in enumerate(self._mp):
            if mt is None:
                continue
            if _is_dev(mt):
                new._mp[i] = mt.to(torch.complex128) if not inplace or not mt.is_complex() else mt
                if not inplace:
                    new._mp[i] = new._mp[i].clone()
            else:
                new._mp[i] = np.array(mt, dtype=np.complex128)
        return new

    def conj(self):
        new = self.metacopy()
        new._mp = [m.conj().resolve_conj() if _is_dev(m) else m.conj() for m in self._mp]
        return new

    def dot_host(self, other) -> complex:
        """<self|other> with ``self`` given as conjugated tensors (host)."""
        e0 = np.eye(1, 1)
        for mt1, mt2 in zip(self._mp, other._mp):
            e0 = np.tensordot(e0, mt2, 1)
            axes = ([0, 1], [0, 1]) if mt1.ndim == 3 else ([0, 1, 2], [0, 1, 2])
            e0 = np.tensordot(e0, mt1, axes).T
        return complex(e0[0, 0])

    # ---- canonical form (reference mp.py:174-228) --------------------------
    @property
    def is_left_canonical(self):
        return self.qnidx == self.site_num - 1

    @property
    def is_right_canonical(self):
        return self.qnidx == 0

    def _site_host(self, i):
        m = self[i]
        return m.cpu().numpy() if _is_dev(m) else np.asarray(m)

    def check_left_canonical(self, rtol=1e-5, atol=1e-8):
        for i in range(len(self) - 1):
            m = self._site_host(i)
            m = m.reshape(-1, m.shape[-1])
            if not np.allclose(m.conj().T @ m, np.eye(m.shape[1]), rtol=rtol, atol=atol):
                return False
        return True

    def check_right_canonical(self, rtol=1e-5, atol=1e-8):
        for i in range(1, len(self)):
            m = self._site_host(i)
            m = m.reshape(m.shape[0], -1)
            if not np.allclose(m @ m.conj().T, np.eye(m.shape[0]), rtol=rtol, atol=atol):
                return False
        return True

    def ensure_left_canonical(self, rtol=1e-5, atol=1e-8):
        if self.to_right or self.qnidx != self.site_num - 1 or not self.check_left_canonical(rtol, atol):
            self.move_qnidx(0)
            self.to_right = True
            return self.canonicalise()
        return self

    def ensure_right_canonical(self, rtol=1e-5, atol=1e-8):
        if (not self.to_right) or self.qnidx != 0 or not self.check_right_canonical(rtol, atol):
            self.move_qnidx(self.site_num - 1)
            self.to_right = False
            return self.canonicalise()
        return self

    def _push_cano(self, idx):
        mt = self[idx]
        assert np.any(mt)
        qnbigl, qnbigr, _ = self._get_big_qn([idx], with_qnmat=False)
        system = "L" if self.to_right else "R"
        u, qnlset, v, qnrset = svd_qn_host(mt, qnbigl, qnbigr, self.qntot, QR=True, system=system,
                                           full_matrices=False)
        self._update_ms(idx, u, v.T, sigma=None, qnlset=qnlset, qnrset=qnrset)

    def canonicalise(self, stop_idx: int = None):
        """QR sweep moving the canonical centre (reference mp.py:890-922)."""
        if self.on_device:
            from renormalizer_amd.mps import device_setup
            return device_setup.canonicalise(self, stop_idx)
        if self.to_right:
            assert self.qnidx == 0
        else:
            assert self.qnidx == self.site_num - 1
        idx = None
        for idx in self.iter_idx_list(full=False, stop_idx=stop_idx):
            self._push_cano(idx)
        if (not self.to_right and idx == 1) or (self.to_right and idx == self.site_num - 2):
            self._switch_direction()
        return self

    def _update_ms(self, idx, u, vt, sigma=None, qnlset=None, qnrset=None, m_trunc=None):
        """Store the factors of a site decomposition (reference mp.py:245-295)."""
        if m_trunc is None:
            m_trunc = u.shape[1]
        u = u[:, :m_trunc]
        vt = vt[:m_trunc, :]
        if sigma is None:
            if self.is_mpo:
                if self.to_right:
                    norm = np.linalg.norm(vt)
                    u = u * norm
                    vt = vt / norm
                else:
                    norm = np.linalg.norm(u)
                    u = u / norm
                    vt = vt * norm
        else:
            sigma = sigma[:m_trunc]
            if (not self.is_mpo and self.to_right) or (self.is_mpo and not self.to_right):
                vt = sigma[:, None] * vt
            else:
                u = u * sigma[None, :]
        pdim = list(self[idx].shape[1:-1])
        pprod = int(np.prod(pdim))
        if self.to_right:
            self._mp[idx + 1] = np.tensordot(vt, self[idx + 1], axes=1)
            ret = u.reshape([u.shape[0] // pprod] + pdim + [m_trunc])
            if qnlset is not None:
                self.qn[idx + 1] = np.array(qnlset[:m_trunc])
                self.qnidx = idx + 1
        else:
            self._mp[idx - 1] = np.tensordot(self[idx - 1], u, axes=1)
            ret = vt.reshape([m_trunc] + pdim + [vt.shape[1] // pprod])
            if qnrset is not None:
                self.qn[idx] = np.array(qnrset[:m_trunc])
                self.qnidx = idx - 1
        assert np.any(ret)
        self._mp[idx] = np.ascontiguousarray(ret)

    def compress(self, temp_m_trunc=None, ret_s=False):
        """SVD truncation sweep (reference mp.py:437-511)."""
        if self.on_device:
            from renormalizer_amd.mps import device_setup
            return device_setup.compress(self, temp_m_trunc, ret_s)
        if self.to_right:
            assert self.qnidx == 0
        else:
            assert self.qnidx == self.site_num - 1
        if self.compress_config.bonddim_should_set:
            self.compress_config.set_bonddim(len(self) + 1)
        system = "L" if self.to_right else "R"
        s_list = []
        for idx in self.iter_idx_list(full=False):
            qnbigl, qnbigr, _ = self._get_big_qn([idx], with_qnmat=False)
            u, sigma, qnlset, v, sigma, qnrset = svd_qn_host(self[idx], qnbigl, qnbigr, self.qntot,
                                                             system=system, full_matrices=False)
            s_list.append(sigma)
            if temp_m_trunc is None:
                m_trunc = self.compress_config.compute_m_trunc(sigma, idx, self.to_right)
            else:
                if isinstance(temp_m_trunc, (list, tuple, np.ndarray)):
                    m_trunc = temp_m_trunc[idx + 1 if self.to_right else idx]
                else:
                    m_trunc = temp_m_trunc
                m_trunc = min(int(m_trunc), len(sigma))
            self._update_ms(idx, u, v.T, sigma, qnlset, qnrset, m_trunc)
        self._switch_direction()
        if not ret_s:
            return self
        width = max(len(s) for s in s_list)
        return self, np.array([np.pad(s, (0, width - len(s))) for s in s_list])
