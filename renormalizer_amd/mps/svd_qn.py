"""Quantum-number blocked decompositions (reference mps/svd_qn.py:99-317).

* ``qn_blocks`` - the host-side block bookkeeping: the block order is the
  iteration order of ``set(tuple(t) for t in localqnl)`` exactly as in the
  reference (svd_qn.py:177), so block layout and new bond quantum numbers are
  bit-identical.
* ``svd_qn_device`` - the QR/RQ branch (QR=True, economic) on the device
  through ``reno_block_qr`` (Householder factorisations).
* ``svd_qn_device_svd`` - the SVD branch (full_matrices=False) on the
  device: block QR, then the Hermitian eigensolver on the k x k factor.
* ``svd_qn_host`` - the full routine (SVD and QR branches) in host numpy,
  for states prepared on the host (canonicalise / compress of a starting
  MPS before it is uploaded).
"""
import threading

import numpy as np
import scipy.linalg
import torch

from renormalizer_amd import native


def add_outer(a: np.ndarray, b: np.ndarray):
    """np.add.outer over all but the last (quantum-number) axis."""
    assert a.shape[-1] == b.shape[-1]
    out = np.array([np.add.outer(a[..., i], b[..., i]) for i in range(a.shape[-1])])
    return out.transpose(list(range(1, out.ndim)) + [0])


def get_qn_mask(qnmat: np.ndarray, qntot):
    return np.all(qnmat == np.array(qntot), axis=-1)


def qn_blocks(qnbigl, qnbigr, qntot):
    """[(nl, nr, lset, rset)] in the reference block order."""
    qn_size = len(qntot)
    localqnl = np.asarray(qnbigl).reshape(-1, qn_size)
    localqnr = np.asarray(qnbigr).reshape(-1, qn_size)
    blocks = []
    for nl in set([tuple(t) for t in localqnl]):
        nr = qntot - nl
        rset = np.where(get_qn_mask(localqnr, nr))[0]
        if len(rset) == 0:
            continue
        lset = np.where(get_qn_mask(localqnl, nl))[0]
        blocks.append((nl, nr, lset, rset))
    return blocks


_PLAN_CACHE = {}
_PLAN_CACHE_MAX = 256
_PLAN_LOCK = threading.Lock()


def svd_qn_plan(qnbigl, qnbigr, qntot):
    """Host part of svd_qn_device, independent of the coefficients: the
    block index sets and the new bond quantum numbers, cached by the quantum
    numbers (a sweep meets the same few patterns over and over)."""
    qnbigl = np.ascontiguousarray(qnbigl)
    qnbigr = np.ascontiguousarray(qnbigr)
    key = (qnbigl.shape, qnbigl.tobytes(), qnbigr.shape, qnbigr.tobytes(), tuple(np.asarray(qntot).tolist()))
    with _PLAN_LOCK:
        plan = _PLAN_CACHE.get(key)
    if plan is None:
        rows = int(np.prod(qnbigl.shape[:-1]))
        cols = int(np.prod(qnbigr.shape[:-1]))
        blocks = qn_blocks(qnbigl, qnbigr, qntot)
        if not blocks:
            raise ValueError("Invalid quantum number")
        qnl, qnr = [], []
        for nl, nr, lset, rset in blocks:
            k = min(len(lset), len(rset))
            qnl += [nl] * k
            qnr += [nr] * k
        plan = (rows, cols, [(b[2], b[3]) for b in blocks], qnl, qnr)
        with _PLAN_LOCK:
            if len(_PLAN_CACHE) >= _PLAN_CACHE_MAX:
                _PLAN_CACHE.pop(next(iter(_PLAN_CACHE)))
            _PLAN_CACHE[key] = plan
    return plan


def svd_qn_device(coef, qnbigl, qnbigr, qntot, system, plan=None):
    """QR ("L") / RQ ("R") of the device coefficient tensor.

    Returns (u, qnl, vt, qnr): u (rows x k), vt (k x cols) complex128 device
    tensors with coef = u @ vt, and the new bond quantum numbers as lists of
    tuples (the reference's ``new_qnl``/``new_qnr``).  ``plan`` is the
    result of svd_qn_plan for these quantum numbers, if already at hand.
    """
    if plan is None:
        plan = svd_qn_plan(qnbigl, qnbigr, qntot)
    rows, cols, blocks, qnl, qnr = plan
    u, vt = native.block_qr(system, coef.reshape(rows, cols), blocks, v_transposed=True)
    return u, list(qnl), vt, list(qnr)


def _ct(t):
    return t.conj().transpose(0, 1).contiguous()


def svd_qn_device_svd(coef, qnbigl, qnbigr, qntot, system, plan=None):
    """SVD branch of svd_qn with full_matrices=False (reference
    svd_qn.py:205-240) on the device, for the truncation sweeps of
    ``compress``.

    Per quantum-number block the QR ("L") or RQ ("R") factorisation
    (reno_block_qr) leaves a k x k factor T beside an orthonormal factor Q.
    The Hermitian eigen-decomposition (reno_block_eigh) of T T^H ("L") or
    T^H T ("R") gives the unitary W that carries Q onto the singular vectors:
    U = Q W and sigma V^H = W^H T ("L"), or V^H = W^H Q and U sigma = T W
    ("R").  The singular values are the norms of the rows ("L") or columns
    ("R") of the rotated factor, which holds them to the absolute accuracy
    eps * sigma_max of a bidiagonal SVD.  Singular directions whose sigma
    lies below about sqrt(eps) * sigma_max are resolved only up to a
    rotation among themselves (the Gram matrix squares them); the span they
    share with the larger ones is exact to rounding.

    Returns (site, sigma, qn, rest) in the reference's order (numpy argsort
    of sigma, descending): "L": site = U (rows x k) and rest = sigma V^H
    (k x cols); "R": site = V^H (k x cols) and rest = U sigma (rows x k);
    sigma a host float64 array and qn the new bond quantum numbers (the
    reference's new_qnl for "L", new_qnr for "R")."""
    if plan is None:
        plan = svd_qn_plan(qnbigl, qnbigr, qntot)
    rows, cols, blocks, qnl, qnr = plan
    first, second = native.block_qr(system, coef.reshape(rows, cols), blocks, v_transposed=True)
    sectors = []
    off = 0
    for r, c in blocks:
        k = min(len(r), len(c))
        sectors.append(np.arange(off, off + k, dtype=np.int64))
        off += k
    if system == "L":
        q, t = first, second  # coef = q @ t, t (k x cols)
        w, _ = native.block_eigh(native.zgemm(t, _ct(t)), sectors)
        site = native.zgemm(q, w)
        rest = native.zgemm(_ct(w), t)
        sigma_dev = torch.linalg.vector_norm(rest, dim=1)
        qn = qnl
    elif system == "R":
        t, q = first, second  # coef = t @ q, t (rows x k)
        w, _ = native.block_eigh(native.zgemm(_ct(t), t), sectors)
        site = native.zgemm(_ct(w), q)
        rest = native.zgemm(t, w)
        sigma_dev = torch.linalg.vector_norm(rest, dim=0)
        qn = qnr
    else:
        raise AssertionError("system")
    sigma = sigma_dev.cpu().numpy()
    if not np.all(np.isfinite(sigma)):
        raise FloatingPointError("svd_qn_device_svd: non-finite singular values")
    order = np.argsort(sigma)[::-1].copy()
    idx = torch.from_numpy(order).to(coef.device)
    if system == "L":
        site = site.index_select(1, idx)
        rest = rest.index_select(0, idx)
    else:
        site = site.index_select(0, idx)
        rest = rest.index_select(1, idx)
    return site, sigma[order], np.array(qn)[order].tolist(), rest


def optimized_svd(a, full_matrices):
    try:
        return scipy.linalg.svd(a, full_matrices=full_matrices, lapack_driver="gesdd")
    except scipy.linalg.LinAlgError:
        return scipy.linalg.svd(a, full_matrices=full_matrices, lapack_driver="gesvd")


def svd_qn_host(coef_array, qnbigl, qnbigr, qntot, QR=False, system=None, full_matrices=True):
    """Host version of svd_qn (SVD and QR branches) for initial-state setup.
    ``full_matrices=True`` returns the complete bases of every block."""
    coef_matrix = coef_array.reshape((int(np.prod(qnbigl.shape[:-1])), int(np.prod(qnbigr.shape[:-1]))))
    u_l, u0_l, v_l, v0_l, s_l, su0_l, sv0_l = [], [], [], [], [], [], []
    qnl, qnl0, qnr, qnr0 = [], [], [], []
    for nl, nr, lset, rset in qn_blocks(qnbigl, qnbigr, qntot):
        block = coef_matrix[np.ix_(lset, rset)]
        dim = min(block.shape)
        if QR:
            mode = "full" if full_matrices else "economic"
            if system == "R":
                bu, bvt = scipy.linalg.rq(block, mode=mode)
            elif system == "L":
                bu, bvt = scipy.linalg.qr(block, mode=mode)
            else:
                raise AssertionError("system")
        else:
            bu, bs, bvt = optimized_svd(block, full_matrices)
            s_l.append(bs)
        for src, lst, lst0, ql, ql0, sl0, idx, n, size in (
                (bu, u_l, u0_l, qnl, qnl0, su0_l, lset, nl, coef_matrix.shape[0]),
                (bvt.T, v_l, v0_l, qnr, qnr0, sv0_l, rset, nr, coef_matrix.shape[1])):
            rec = np.zeros((size, dim), dtype=src.dtype)
            rec[idx, :] = src[:, :dim]
            lst.append(rec)
            ql += [n] * dim
            if full_matrices:
                rec0 = np.zeros((size, src.shape[1] - dim), dtype=src.dtype)
                rec0[idx, :] = src[:, dim:]
                lst0.append(rec0)
                ql0 += [n] * (src.shape[1] - dim)
                sl0.append(np.zeros(src.shape[1] - dim))
    if len(u_l) + len(u0_l) == 0 or len(v_l) + len(v0_l) == 0:
        raise ValueError("Invalid quantum number")
    u = np.concatenate(u_l + u0_l, axis=1)
    v = np.concatenate(v_l + v0_l, axis=1)
    new_qnl, new_qnr = qnl + qnl0, qnr + qnr0
    if QR:
        return u, new_qnl, v, new_qnr
    su = np.concatenate(s_l + su0_l)
    sv = np.concatenate(s_l + sv0_l)
    if not full_matrices:
        order = np.argsort(su)[::-1]
        u, v = u[:, order], v[:, order]
        su = sv = su[order]
        new_qnl = np.array(new_qnl)[order].tolist()
        new_qnr = np.array(new_qnr)[order].tolist()
    return u, su, new_qnl, v, sv, new_qnr


def select_basis(vset, sset, qnlist, compset, m_max, percent=0):
    """Renormalised basis selection (reference mps/lib.py:253-322): ``percent``
    of the kept states is shared equally among the quantum-number sectors
    (largest weights first inside each), the rest goes to the largest weights
    overall.  Host bookkeeping only: returns the chosen column indices ``sidx``
    (into vset / compset), the kept quantum numbers and the kept weights."""
    qnlist = [tuple(qn) for qn in qnlist]
    qnset = set(qnlist)
    basdic = {i: [qnlist[i], sset[i]] for i in range(len(qnlist)) if qnlist[i] in qnset}

    def block_select(qn, n):
        block = sorted(((i, v) for i, v in basdic.items() if v[0] == qn), key=lambda x: x[1][1],
                       reverse=True)
        sel = [i for i, _ in block[:min(n, len(block))]]
        for i in sel:
            del basdic[i]
        return sel

    nbasis = min(len(basdic), m_max)
    sidx = []
    if percent != 0:
        nbas_block = int(nbasis * percent / len(qnset))
        for iqn in qnset:
            sidx += block_select(iqn, nbas_block)
    nbasis = nbasis - len(sidx)
    rest = sorted(basdic.items(), key=lambda x: x[1][1], reverse=True)
    sidx += [i for i, _ in rest[:nbasis]]
    assert len(sidx) == len(set(sidx))
    return sidx, np.array([qnlist[i] for i in sidx]), np.array([sset[i] for i in sidx])
