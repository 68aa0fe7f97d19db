"""Initial-state preparation on the device: direct sum, MPO application,
QR canonicalisation and SVD truncation sweeps of a device MPS (reference
mps/mp.py:374-435, 437-511, 890-922; mps/mpo.py:331-389).

These are the operations ``expand_bond_dimension`` and ``compressed_sum``
are built from (reference mps/mps.py:1934-2023, mps/lib.py:417-439).  The
quantum-number bookkeeping (qn, qnidx, qntot, block plans) stays on the
host exactly as in the host versions in mp.py; every site-tensor operation
runs on the device through the native library: the blocked QR / RQ
(reno_block_qr), the blocked Hermitian eigensolver (reno_block_eigh) for the
SVD branch, and the complex GEMM (reno_zgemm) for MPO application and for
carrying each factor into the neighbouring site.  Torch provides the device
memory and the index permutations between GEMMs.

Only matrix product states (three-leg sites) are handled; the caller keeps
MPOs and density operators on the host path.
"""
from typing import List

import numpy as np
import torch

from renormalizer_amd import native
from renormalizer_amd.mps.svd_qn import add_outer, svd_qn_device, svd_qn_device_svd


def _require_mps(mp, what):
    if not mp.is_mps or getattr(mp, "is_mpdm", False):
        raise NotImplementedError(f"{what} on the device handles matrix product states only")


def _as_z(t: torch.Tensor) -> torch.Tensor:
    return (t if t.dtype == torch.complex128 else t.to(torch.complex128)).contiguous()


def _dev_z(m, device) -> torch.Tensor:
    if isinstance(m, torch.Tensor):
        return _as_z(m.to(device))
    return torch.from_numpy(np.ascontiguousarray(m, dtype=np.complex128)).to(device)


def upload_complex(mp, device=None):
    """Site tensors to the device as complex128 (in place)."""
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    mp._mp = [_dev_z(m, device) for m in mp._mp]
    mp.dtype = np.complex128
    if hasattr(mp, "coeff"):
        mp.coeff = complex(mp.coeff)
    return mp


def _store(mp, idx: int, site: torch.Tensor, rest: torch.Tensor, qn: List, m: int):
    """Keep the first ``m`` bond states of a site decomposition: ``site``
    becomes the (left- or right-) orthonormal tensor of site ``idx`` and
    ``rest`` is multiplied into the neighbour the sweep moves to (reference
    mp.py:245-295 for a state)."""
    pdim = list(mp[idx].shape[1:-1])
    pprod = int(np.prod(pdim))
    mp.dtype = np.complex128
    if hasattr(mp, "coeff"):
        mp.coeff = complex(mp.coeff)
    if mp.to_right:
        nxt = _as_z(mp[idx + 1])
        r = rest[:m].contiguous()
        mp._mp[idx + 1] = native.zgemm(r, nxt.reshape(nxt.shape[0], -1)).reshape((m,) + tuple(nxt.shape[1:]))
        s = site[:, :m].contiguous()
        mp._mp[idx] = s.reshape([s.shape[0] // pprod] + pdim + [m])
        mp.qn[idx + 1] = np.array(qn[:m])
        mp.qnidx = idx + 1
    else:
        prv = _as_z(mp[idx - 1])
        r = rest[:, :m].contiguous()
        mp._mp[idx - 1] = native.zgemm(prv.reshape(-1, prv.shape[-1]), r).reshape(tuple(prv.shape[:-1]) + (m,))
        s = site[:m].contiguous()
        mp._mp[idx] = s.reshape([m] + pdim + [s.shape[1] // pprod])
        mp.qn[idx] = np.array(qn[:m])
        mp.qnidx = idx - 1


def canonicalise(mp, stop_idx: int = None):
    """QR sweep moving the canonical centre (reference mp.py:890-922)."""
    _require_mps(mp, "canonicalise")
    if mp.to_right:
        assert mp.qnidx == 0
    else:
        assert mp.qnidx == mp.site_num - 1
    idx = None
    for idx in mp.iter_idx_list(full=False, stop_idx=stop_idx):
        qnbigl, qnbigr, _ = mp._get_big_qn([idx], with_qnmat=False)
        system = "L" if mp.to_right else "R"
        u, qnl, vt, qnr = svd_qn_device(_as_z(mp[idx]), qnbigl, qnbigr, mp.qntot, system)
        if mp.to_right:
            _store(mp, idx, u, vt, qnl, u.shape[1])
        else:
            _store(mp, idx, vt, u, qnr, vt.shape[0])
    if (not mp.to_right and idx == 1) or (mp.to_right and idx == mp.site_num - 2):
        mp._switch_direction()
    mp._invalidate()
    return mp


def compress(mp, temp_m_trunc=None, ret_s=False):
    """SVD truncation sweep (reference mp.py:437-511)."""
    _require_mps(mp, "compress")
    if mp.to_right:
        assert mp.qnidx == 0
    else:
        assert mp.qnidx == mp.site_num - 1
    if mp.compress_config.bonddim_should_set:
        mp.compress_config.set_bonddim(len(mp) + 1)
    system = "L" if mp.to_right else "R"
    s_list = []
    for idx in mp.iter_idx_list(full=False):
        qnbigl, qnbigr, _ = mp._get_big_qn([idx], with_qnmat=False)
        site, sigma, qn, rest = svd_qn_device_svd(_as_z(mp[idx]), qnbigl, qnbigr, mp.qntot, system)
        s_list.append(sigma)
        if temp_m_trunc is None:
            m_trunc = mp.compress_config.compute_m_trunc(sigma, idx, mp.to_right)
        else:
            if isinstance(temp_m_trunc, (list, tuple, np.ndarray)):
                m_trunc = temp_m_trunc[idx + 1 if mp.to_right else idx]
            else:
                m_trunc = temp_m_trunc
            m_trunc = min(int(m_trunc), len(sigma))
        _store(mp, idx, site, rest, qn, int(m_trunc))
    mp._switch_direction()
    mp._invalidate()
    if not ret_s:
        return mp
    width = max(len(s) for s in s_list)
    return mp, np.array([np.pad(s, (0, width - len(s))) for s in s_list])


def add(a, b):
    """Direct sum of the bond spaces of two device states (reference
    mp.py:374-435)."""
    _require_mps(a, "add")
    assert np.all(a.qntot == b.qntot)
    assert a.site_num == b.site_num
    new = a.metacopy()
    new.dtype = np.complex128
    new.compress_config.update(a.compress_config)
    n = a.site_num
    dev = a[0].device if a.on_device else b[0].device
    sa = [_dev_z(t, dev) for t in a._mp]
    sb = [_dev_z(t, dev) for t in b._mp]
    new._mp[0] = torch.cat([sa[0], sb[0]], dim=2)
    for i in range(1, n - 1):
        x, y = sa[i], sb[i]
        m = torch.zeros((x.shape[0] + y.shape[0], x.shape[1], x.shape[2] + y.shape[2]), dtype=torch.complex128,
                        device=dev)
        m[:x.shape[0], :, :x.shape[2]] = x
        m[x.shape[0]:, :, x.shape[2]:] = y
        new._mp[i] = m
    new._mp[-1] = torch.cat([sa[-1], sb[-1]], dim=0)
    new.move_qnidx(b.qnidx)
    new.to_right = b.to_right
    new.qn = [np.concatenate([np.array(q1), np.array(q2)]) for q1, q2 in zip(a.qn, b.qn)]
    new.qn[0] = np.zeros((1, new.qn[0].shape[1]), dtype=int)
    new.qn[-1] = np.zeros((1, new.qn[0].shape[1]), dtype=int)
    if hasattr(new, "coeff"):
        new.coeff = complex(new.coeff)
    new._invalidate()
    return new


def mpo_apply(mpo, mps):
    """MPO x MPS on the device (reference mpo.py:331-389): per site
    out[(b, l), d, (c, r)] = sum_e W[b, d, e, c] A[l, e, r], one GEMM of the
    (b d c) x e operator matrix with the e x (l r) state matrix."""
    _require_mps(mps, "Mpo.apply")
    assert mpo.site_num == mps.site_num
    new = mps.metacopy()
    new.dtype = np.complex128
    if hasattr(new, "coeff"):
        new.coeff = complex(new.coeff)
    dev = mps[0].device
    for i, (w, a) in enumerate(zip(mpo._mp, mps._mp)):
        w = np.asarray(w)
        b, d, e, c = w.shape
        a = _as_z(a)
        l, e2, r = a.shape
        assert e == e2
        wm = torch.from_numpy(np.ascontiguousarray(w.transpose(0, 1, 3, 2).reshape(b * d * c, e),
                                                   dtype=np.complex128)).to(dev)
        am = a.permute(1, 0, 2).reshape(e, l * r).contiguous()
        p = native.zgemm(wm, am).reshape(b, d, c, l, r)
        new._mp[i] = p.permute(0, 3, 1, 2, 4).reshape(b * l, d, c * r).contiguous()
    orig = new.qnidx
    new.move_qnidx(mpo.qnidx)
    new.qn = [add_outer(np.array(qo), np.array(qm)).reshape(-1, np.array(qo).shape[1])
              for qo, qm in zip(mpo.qn, new.qn)]
    new.qntot = new.qntot + mpo.qntot
    new.move_qnidx(orig)
    new._invalidate()
    return new
