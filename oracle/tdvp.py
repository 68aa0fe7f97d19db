"""numpy complex128 restatement of the reference's one-site TDVP-PS hot path.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  It never runs inside the product path.

Each function states the reference file:line it restates and what pins it
(tests/test_oracle.py holds the pins).  Conventions follow the reference:
MPS site (M_l, d, M_r), MPO site (D_l, d_out, d_in, D_r), environment
(bra, mpo, ket), row-major, quantum numbers as int arrays (dim, qn_size).
"""
import numpy as np
import scipy.linalg
from scipy.linalg import eigh_tridiagonal


# ---------------------------------------------------------------------------
# contractions


def contract_one_site(environ, ms, mo, domain, ms_conj=None):
    """Environment update, reference mps/lib.py:169-250 (3-leg sites).

    Same pairwise order as the reference path lists (lib.py:200-204, 232-236).
    Pinned by: L[i].R[i+1] == <H> for every bond (test_mp.py:51-62 pattern)
    and brute-force einsum on tiny shapes.
    """
    assert domain in ("L", "R")
    if ms_conj is None:
        ms_conj = ms.conj()
    if ms.ndim == 4:
        return _contract_one_site_ancilla(environ, ms, mo, domain, ms_conj)
    if domain == "L":
        assert environ.shape[0] == ms_conj.shape[0]
        assert environ.shape[1] == mo.shape[0]
        assert environ.shape[2] == ms.shape[0]
        t = np.tensordot(environ, ms_conj, axes=([0], [0]))       # abc,adf->bcdf
        t = np.tensordot(t, mo, axes=([0, 2], [0, 1]))             # bcdf,bdeg->cfeg
        t = np.tensordot(t, ms, axes=([0, 2], [0, 1]))             # cfeg,ceh->fgh
        return t
    assert environ.shape[0] == ms_conj.shape[-1]
    assert environ.shape[1] == mo.shape[-1]
    assert environ.shape[2] == ms.shape[-1]
    t = np.tensordot(ms_conj, environ, axes=([2], [0]))            # fda,abc->fdbc
    t = np.tensordot(t, mo, axes=([1, 2], [1, 3]))                 # fdbc,gdeb->fcge
    t = np.tensordot(t, ms, axes=([3, 1], [1, 2]))                 # fcge,hec->fgh
    return t


def _contract_one_site_ancilla(environ, ms, mo, domain, ms_conj):
    """Environment update for a 4-leg density-operator site (M_l, d, d_a,
    M_r), reference mps/lib.py:206-211 (L) and 238-243 (R): the ancilla
    index l is shared by bra and ket and untouched by the operator.
    Pinned by: thermal expectation values vs the dense Gibbs state."""
    if domain == "L":
        t = np.tensordot(environ, ms_conj, axes=([0], [0]))       # abc,adlf->bcdlf
        t = np.tensordot(t, mo, axes=([0, 2], [0, 1]))             # bcdlf,bdeg->clfeg
        t = np.tensordot(t, ms, axes=([0, 3, 1], [0, 1, 2]))       # clfeg,celh->fgh
        return t
    t = np.tensordot(ms_conj, environ, axes=([3], [0]))            # fdla,abc->fdlbc
    t = np.tensordot(t, mo, axes=([1, 3], [1, 3]))                 # fdlbc,gdeb->flcge
    t = np.tensordot(t, ms, axes=([4, 1, 2], [1, 2, 3]))           # flcge,helc->fgh
    return t


def hop1(ltensor, mo, rtensor, c):
    """1-site effective Hamiltonian matvec "abc, bdef, lfk, cek -> adl"
    (reference mps/hop_expr.py:68-79); for a 4-leg density-operator site
    "abc, bdef, lfk, cegk -> adgl" (hop_expr.py:80-91), the ancilla g a
    spectator.  Pinned by: dense H_eff assembled from the Kronecker product
    of L, W, R on tiny shapes; Hermiticity; thermal expectation values vs
    the dense Gibbs state (4-leg)."""
    if c.ndim == 4:
        t = np.tensordot(ltensor, c, axes=([2], [0]))              # abc,cegk->abegk
        t = np.tensordot(t, mo, axes=([1, 2], [0, 2]))             # abegk,bdef->agkdf
        t = np.tensordot(t, rtensor, axes=([2, 4], [2, 1]))        # agkdf,lfk->agdl
        return t.transpose(0, 2, 1, 3)                             # -> adgl
    t = np.tensordot(ltensor, c, axes=([2], [0]))                  # abc,cek->abek
    t = np.tensordot(t, mo, axes=([1, 2], [0, 2]))                 # abek,bdef->akdf
    t = np.tensordot(t, rtensor, axes=([1, 3], [2, 1]))            # akdf,lfk->adl
    return t


def hop0(ltensor, rtensor, s):
    """0-site (bond) matvec "abc, lbk, ck -> al" (reference hop_expr.py:57-66).
    Pinned like hop1."""
    t = np.tensordot(ltensor, s, axes=([2], [0]))                  # abc,ck->abk
    return np.tensordot(t, rtensor, axes=([1, 2], [1, 2]))         # abk,lbk->al


# ---------------------------------------------------------------------------
# Krylov exponential


def _expm_krylov(alpha, beta, V, v_norm, dt):
    """reference lib/krylov/krylov.py:15-24."""
    try:
        w_hess, u_hess = eigh_tridiagonal(alpha, beta)
    except np.linalg.LinAlgError:
        h = np.diag(alpha) + np.diag(beta, k=-1) + np.diag(beta, k=1)
        w_hess, u_hess = np.linalg.eigh(h)
    return V @ (u_hess @ (v_norm * np.exp(dt * w_hess) * u_hess[0]))


def expm_krylov(Afunc, dt, vstart, block_size=50):
    """Lanczos approximation of expm(dt A) v, reference krylov.py:27-84,
    including the stopping rule (breakdown, full space, allclose every even
    j > 3).  Pinned by: dense eigh exponential for N in {1,2,4,10,200,800}
    (reference test_krylov.py:12-39) and the matvec-count sequence."""
    if not np.iscomplex(dt):
        dt = dt.real
    vstart = np.asarray(vstart)
    nrmv = float(np.linalg.norm(vstart))
    assert nrmv > 0
    vstart = vstart / nrmv
    alpha = np.zeros(block_size)
    beta = np.zeros(block_size - 1)
    V = np.empty((block_size, len(vstart)), dtype=vstart.dtype)
    V[0] = vstart
    res = None
    for j in range(len(vstart)):
        w = Afunc(V[j])
        alpha[j] = np.vdot(w, V[j]).real
        if j == len(vstart) - 1:
            return _expm_krylov(alpha[:j + 1], beta[:j], V[:j + 1, :].T, nrmv, dt), j + 1
        if len(V) == j + 1:
            V, old_V = np.empty((len(V) + block_size, len(vstart)), dtype=vstart.dtype), V
            V[:len(old_V)] = old_V
            alpha = np.concatenate([alpha, np.zeros(block_size)])
            beta = np.concatenate([beta, np.zeros(block_size)])
        w = w - (alpha[j] * V[j] + (beta[j - 1] * V[j - 1] if j > 0 else 0))
        beta[j] = np.linalg.norm(w)
        if beta[j] < 100 * len(vstart) * np.finfo(float).eps:
            return _expm_krylov(alpha[:j + 1], beta[:j], V[:j + 1, :].T, nrmv, dt), j + 1
        if 3 < j and j % 2 == 0:
            new_res = _expm_krylov(alpha[:j + 1], beta[:j], V[:j + 1].T, nrmv, dt)
            if res is not None and np.allclose(res, new_res):
                return new_res, j + 1
            res = new_res
        V[j + 1] = w / beta[j]


# ---------------------------------------------------------------------------
# quantum numbers and the blocked factorisations


def add_outer(a, b):
    """reference mps/svd_qn.py:305-313."""
    assert a.shape[-1] == b.shape[-1]
    out = np.array([np.add.outer(a[..., i], b[..., i]) for i in range(a.shape[-1])])
    return out.transpose(list(range(1, out.ndim)) + [0])


def get_qn_mask(qnmat, qntot):
    """reference mps/svd_qn.py:316-317."""
    return np.all(qnmat == np.array(qntot), axis=-1)


def qn_blocks(qnbigl, qnbigr, qntot):
    """Block index sets of svd_qn in the reference's order (svd_qn.py:171-185):
    iteration over ``set(tuple(t) for t in localqnl)``, rows lset, cols rset.
    Returns a list of (nl, nr, lset, rset)."""
    qn_size = len(qntot)
    localqnl = qnbigl.reshape(-1, qn_size)
    localqnr = qnbigr.reshape(-1, qn_size)
    out = []
    for nl in set([tuple(t) for t in localqnl]):
        nr = qntot - nl
        rset = np.where(get_qn_mask(localqnr, nr))[0]
        if len(rset) == 0:
            continue
        lset = np.where(get_qn_mask(localqnl, nl))[0]
        out.append((nl, nr, lset, rset))
    return out


def svd_qn(coef_array, qnbigl, qnbigr, qntot, QR=False, system=None, full_matrices=True):
    """Blocked SVD / QR / RQ, reference mps/svd_qn.py:99-240 (the
    ``opt_full_matrices`` shortcut is not restated: full_matrices=True SVDs
    here return the full bases).  Pinned by: u@v.T reconstructs allowed
    blocks, u/v isometric, exact zeros outside blocks, block order."""
    coef_matrix = coef_array.reshape((np.prod(qnbigl.shape[:-1]), np.prod(qnbigr.shape[:-1])))
    u_list, u0_list, v_list, v0_list = [], [], [], []
    s_list, su0_list, sv0_list = [], [], []
    qnl, qnl0, qnr, qnr0 = [], [], [], []
    for nl, nr, lset, rset in qn_blocks(qnbigl, qnbigr, qntot):
        block = coef_matrix[np.ix_(lset, rset)]
        dim = min(block.shape)
        if not QR:
            try:
                bu, bs, bvt = scipy.linalg.svd(block, full_matrices=full_matrices, lapack_driver="gesdd")
            except scipy.linalg.LinAlgError:
                bu, bs, bvt = scipy.linalg.svd(block, full_matrices=full_matrices, lapack_driver="gesvd")
            s_list.append(bs)
        else:
            mode = "full" if full_matrices else "economic"
            if system == "R":
                bu, bvt = scipy.linalg.rq(block, mode=mode)
            elif system == "L":
                bu, bvt = scipy.linalg.qr(block, mode=mode)
            else:
                raise AssertionError("system")
        bv = bvt.T
        for src, lst, lst0, ql, ql0, sl0, idx, n, size in (
                (bu, u_list, u0_list, qnl, qnl0, su0_list, lset, nl, coef_matrix.shape[0]),
                (bv, v_list, v0_list, qnr, qnr0, sv0_list, rset, nr, coef_matrix.shape[1])):
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
    if len(u_list) + len(u0_list) == 0 or len(v_list) + len(v0_list) == 0:
        raise ValueError("Invalid quantum number")
    u = np.concatenate(u_list + u0_list, axis=1)
    v = np.concatenate(v_list + v0_list, axis=1)
    new_qnl = qnl + qnl0
    new_qnr = qnr + qnr0
    if QR:
        return u, new_qnl, v, new_qnr
    su = np.concatenate(s_list + su0_list)
    sv = np.concatenate(s_list + sv0_list)
    if not full_matrices:
        order = np.argsort(su)[::-1]
        u, v = u[:, order], v[:, order]
        su = sv = su[order]
        new_qnl = np.array(new_qnl)[order].tolist()
        new_qnr = np.array(new_qnr)[order].tolist()
    return u, su, new_qnl, v, sv, new_qnr


# ---------------------------------------------------------------------------
# MPS state for the sweep


class State:
    """Minimal MPS container: the data the TDVP-PS sweep reads and writes
    (mps/mp.py:60-80: _mp, qn, qnidx, qntot, to_right) plus the per-site
    quantum numbers of the local basis (model.basis[i].sigmaqn)."""

    def __init__(self, sites, qn, qnidx, qntot, to_right, sigmaqn):
        self.sites = [np.array(s, dtype=np.complex128) for s in sites]
        self.qn = [np.array(q, dtype=np.int64).reshape(len(q), -1) for q in qn]
        self.qnidx = int(qnidx)
        self.qntot = np.array(qntot, dtype=np.int64).reshape(-1)
        self.to_right = bool(to_right)
        self.sigmaqn = [np.array(s, dtype=np.int64) for s in sigmaqn]

    def copy(self):
        return State(self.sites, self.qn, self.qnidx, self.qntot, self.to_right, self.sigmaqn)

    def __len__(self):
        return len(self.sites)

    # mp.py:230-243
    def iter_idx_list(self, full):
        n = len(self)
        if self.to_right:
            return range(self.qnidx, n if full else n - 1)
        return range(self.qnidx, -1 if full else 0, -1)

    # mp.py:297-306
    def switch_direction(self):
        if self.to_right:
            self.qnidx = len(self) - 1
            self.to_right = False
        else:
            self.qnidx = 0
            self.to_right = True

    # mp.py:308-352, one active site
    def get_big_qn(self, idx):
        assert self.qnidx == idx
        sq = self.sigmaqn[idx]
        qnl = np.array(self.qn[idx])
        qnr = np.array(self.qn[idx + 1])
        if self.to_right:
            return add_outer(qnl, sq), qnr
        return qnl, add_outer(sq, qnr)


class Environ:
    """reference mps/lib.py:12-118 (single MPO, full L and R construction)."""

    def __init__(self, sites, mpo, domain=None, sites_conj=None):
        self.disk = {}
        self.sentinel = np.ones((1, 1, 1))
        if sites_conj is None:
            sites_conj = [s.conj() for s in sites]
        n = len(sites)
        self.disk[("L", -1)] = self.sentinel
        self.disk[("R", n)] = self.sentinel
        for dom in (["L", "R"] if domain is None else [domain]):
            if dom == "L":
                rng = range(0, n - 1)
            else:
                rng = range(n - 1, 0, -1)
            t = self.sentinel
            for idx in rng:
                t = contract_one_site(t, sites[idx], mpo[idx], dom, sites_conj[idx])
                self.disk[(dom, idx)] = t

    def read(self, domain, idx):
        return self.disk[(domain, idx)]

    def system(self, domain, idx, sites, mpo, itensor):
        """GetLR(..., method="System") with a given itensor (lib.py:99-112)."""
        t = contract_one_site(itensor, sites[idx], mpo[idx], domain)
        self.disk[(domain, idx)] = t
        return t


def tdvp_ps(state: State, mpo, evolve_dt):
    """One TDVP-PS time step (two half-sweeps), reference mps/mps.py:1267-1404
    with ivp_solver="krylov" and real ``evolve_dt``, followed by
    normalize("mps_only") as Mps.evolve does for real dt (mps.py:657-661).

    Returns (new_state, local_steps).  Pinned by: exactness at full bond
    dimension vs the dense propagator, closed forms (Delta-only spin,
    g=0 Holstein), norm/energy conservation, the reference's 3-site
    Holstein and 5-mode spin-boson integration targets via exact
    diagonalisation."""
    mps = state.copy()
    environ = Environ(mps.sites, mpo)
    local_steps = []
    n = len(mps)
    for _ in range(2):
        for imps in mps.iter_idx_list(full=True):
            system = "L" if mps.to_right else "R"
            l_array = environ.read("L", imps - 1)
            r_array = environ.read("R", imps + 1)
            shape = mps.sites[imps].shape
            w = mpo[imps]
            mps_t, j = expm_krylov(lambda y: hop1(l_array, w, r_array, y.reshape(shape)).ravel(),
                                   -1j * evolve_dt / 2, mps.sites[imps].ravel())
            local_steps.append(j)
            mps_t = mps_t.reshape(shape)
            qnbigl, qnbigr = mps.get_big_qn(imps)
            u, qnlset, v, qnrset = svd_qn(mps_t, qnbigl, qnbigr, mps.qntot, QR=True,
                                          system=system, full_matrices=False)
            vt = v.T
            if not mps.to_right and imps != 0:
                mps.sites[imps] = vt.reshape([-1] + list(shape[1:]))
                mps.qn[imps] = np.array(qnrset).reshape(len(qnrset), -1)
                mps.qnidx = imps - 1
                r_array = environ.system("R", imps, mps.sites, mpo, r_array)
                shape_u = u.shape
                mps_t, j = expm_krylov(lambda y: hop0(l_array, r_array, y.reshape(shape_u)).ravel(),
                                       1j * evolve_dt / 2, u.ravel())
                local_steps.append(j)
                mps.sites[imps - 1] = np.tensordot(mps.sites[imps - 1], mps_t.reshape(shape_u),
                                                   axes=(-1, 0))
            elif mps.to_right and imps != n - 1:
                mps.sites[imps] = u.reshape(list(shape[:-1]) + [-1])
                mps.qn[imps + 1] = np.array(qnlset).reshape(len(qnlset), -1)
                mps.qnidx = imps + 1
                l_array = environ.system("L", imps, mps.sites, mpo, l_array)
                shape_v = vt.shape
                mps_t, j = expm_krylov(lambda y: hop0(l_array, r_array, y.reshape(shape_v)).ravel(),
                                       1j * evolve_dt / 2, vt.ravel())
                local_steps.append(j)
                mps.sites[imps + 1] = np.tensordot(mps_t.reshape(shape_v), mps.sites[imps + 1],
                                                   axes=(1, 0))
            else:
                mps.sites[imps] = mps_t
        mps.switch_direction()
    normalize_mps_only(mps)
    return mps, local_steps


# ---------------------------------------------------------------------------
# two-site TDVP-PS with SVD truncation


def merge_w(w0, w1):
    """Two MPO sites as one site of physical dimension d0*d1:
    W[b, (d g), (e h), j] = sum_f W0[b,d,e,f] W1[f,g,h,j].  hop1 with it is
    the two-site hop_expr "abc,bdef,fghj,ljk,cehk->adgl" (hop_expr.py:99-103)
    with the (d, g) and (e, h) pairs fused."""
    m = np.einsum("bdef,fghj->bdgehj", w0, w1)
    return m.reshape(w0.shape[0], w0.shape[1] * w1.shape[1], w0.shape[2] * w1.shape[2], w1.shape[3])


def select_basis(vset, sset, qnlist, compset, mmax, percent=0):
    """reference mps/lib.py:253-322: keep the mmax largest weights (a
    percent share first taken equally from every quantum-number block);
    returns (ms, mpsdim, mpsqn, compms) with compms columns scaled by the
    weights."""
    qnlist = [tuple(q) for q in qnlist]
    qnset = set(qnlist)
    basdic = {i: [qnlist[i], sset[i]] for i in range(len(qnlist)) if qnlist[i] in qnset}

    def block_select(qn, cnt):
        blk = sorted(((i, v) for i, v in basdic.items() if v[0] == qn), key=lambda x: x[1][1], reverse=True)
        got = [i for i, _ in blk[:min(cnt, len(blk))]]
        for i in got:
            del basdic[i]
        return got

    nbasis = min(len(basdic), mmax)
    sidx = []
    if percent != 0:
        per = int(nbasis * percent / len(qnset))
        for qn in qnset:
            sidx += block_select(qn, per)
    rest = sorted(basdic.items(), key=lambda x: x[1][1], reverse=True)
    sidx += [i for i, _ in rest[:nbasis - len(sidx)]]
    ms = np.zeros((vset.shape[0], len(sidx)), dtype=vset.dtype)
    comp = None if compset is None else np.zeros((compset.shape[0], len(sidx)), dtype=compset.dtype)
    mpsqn = []
    for k, i in enumerate(sidx):
        ms[:, k] = vset[:, i]
        if comp is not None and i < compset.shape[1]:
            comp[:, k] = compset[:, i] * sset[i]
        mpsqn.append(qnlist[i])
    return ms, len(sidx), np.array(mpsqn), comp


def m_trunc(sigma, bond, max_dim=None, threshold=None):
    """CompressConfig.compute_m_trunc (utils/configs.py:196-219): fixed
    (max_dim at the bond), threshold (normalised weights above it), or the
    smaller of the two."""
    cands = []
    if max_dim is not None:
        cands.append(min(int(max_dim), len(sigma)))
    if threshold is not None:
        cands.append(int(np.sum(sigma / np.linalg.norm(sigma) > threshold)))
    return min(cands)


def truncate_two_site(ms2, qnbigl, qnbigr, qntot, to_right, max_dim=None, threshold=None):
    """The truncation of the two-site TDVP-PS step (reference mp.py:651-780
    with one root and percent 0; mps.py:1406-1517): blocked SVD of the
    (M_l d0) x (d1 M_r) matrix (svd_qn), the number kept by compute_m_trunc
    (m_trunc: fixed, threshold or both), basis selection of the largest
    weights (select_basis).  Moving right the left factor keeps the singular
    vectors and the right one absorbs the weights, moving left the reverse.
    Returns (left site (M_l, d0, m), right site (m, d1, M_r), m, quantum
    numbers of the new bond).  Pinned by: the product left . right is the
    best rank-m approximation of the dense two-site matrix (Eckart-Young:
    the discarded weight equals the norm of its smallest singular values),
    the kept count of the threshold branch, exact zeros outside the quantum
    number blocks."""
    system = "L" if to_right else "R"
    u, su, qnlnew, v, sv, qnrnew = svd_qn(ms2, qnbigl, qnbigr, qntot, system=system)
    lshape = list(qnbigl.shape[:-1])
    rshape = list(qnbigr.shape[:-1])
    if to_right:
        mt = m_trunc(su, None, max_dim, threshold)
        ms, msdim, msqn, comp = select_basis(u, su, qnlnew, v, mt)
        left = ms.reshape(lshape + [msdim])
        right = np.moveaxis(comp.reshape(rshape + [msdim]), -1, 0)
    else:
        mt = m_trunc(sv, None, max_dim, threshold)
        ms, msdim, msqn, comp = select_basis(v, sv, qnrnew, u, mt)
        right = np.moveaxis(ms.reshape(rshape + [msdim]), -1, 0)
        left = comp.reshape(lshape + [msdim])
    return left, right, msdim, msqn


def tdvp_ps2(state: State, mpo, evolve_dt, max_dim=None, threshold=None):
    """One two-site TDVP-PS time step, reference mps/mps.py:1406-1517 with
    ivp_solver="krylov" and real ``evolve_dt``: per bond the two-site
    Krylov step (-i dt/2), SVD truncation (mp.py:651-780, single root,
    percent 0), the environment update, and the one-site step (+i dt/2) of
    the site left behind followed by a QR moving the centre on (mp.py:
    890-908); then normalize("mps_only").  Returns (new_state,
    local_steps).  Pinned by: exactness vs the dense propagator when no
    truncation happens, the 1-site oracle's closed forms, and its
    truncation (truncate_two_site) against a direct SVD of the dense
    two-site tensor."""
    mps = state.copy()
    n = len(mps)
    environ = Environ(mps.sites, mpo)
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
            ms2 = np.tensordot(mps.sites[c0], mps.sites[c1], axes=1)
            shape2 = ms2.shape
            w2 = merge_w(mpo[c0], mpo[c1])
            fused = (shape2[0], shape2[1] * shape2[2], shape2[3])
            mps_t, j = expm_krylov(lambda y: hop1(l_array, w2, r_array, y.reshape(fused)).ravel(),
                                   -1j * evolve_dt / 2, ms2.ravel())
            local_steps.append(j)
            # big quantum numbers of the two-site block (mp.py:308-352)
            qnl = np.array(mps.qn[c0])
            qnr = np.array(mps.qn[c1 + 1])
            qnbigl = add_outer(qnl, mps.sigmaqn[c0])
            qnbigr = add_outer(mps.sigmaqn[c1], qnr)
            left, right, msdim, msqn = truncate_two_site(mps_t.reshape(shape2), qnbigl, qnbigr,
                                                         mps.qntot, mps.to_right, max_dim, threshold)
            mps.sites[c0] = left
            mps.sites[c1] = right
            mps.qnidx = c1 if mps.to_right else c0
            mps.qn[c1] = np.array(msqn).reshape(msdim, -1)
            if imps == last:
                continue
            if mps.to_right:
                l_array = environ.system("L", lidx + 1, mps.sites, mpo, l_array)
            else:
                r_array = environ.system("R", ridx - 1, mps.sites, mpo, r_array)
            shape1 = mps.sites[c2].shape
            w = mpo[c2]
            mps_t, j = expm_krylov(lambda y: hop1(l_array, w, r_array, y.reshape(shape1)).ravel(),
                                   1j * evolve_dt / 2, mps.sites[c2].ravel())
            local_steps.append(j)
            mps.sites[c2] = mps_t.reshape(shape1)
            # _push_cano(c2): QR moving the centre to the next site
            qnbigl1, qnbigr1 = mps.get_big_qn(c2)
            uq, qnlset, vq, qnrset = svd_qn(mps.sites[c2], qnbigl1, qnbigr1, mps.qntot, QR=True,
                                            system="L" if mps.to_right else "R", full_matrices=False)
            if mps.to_right:
                mps.sites[c2 + 1] = np.tensordot(vq.T, mps.sites[c2 + 1], axes=1)
                mps.sites[c2] = uq.reshape(list(shape1[:-1]) + [uq.shape[1]])
                mps.qn[c2 + 1] = np.array(qnlset).reshape(len(qnlset), -1)
                mps.qnidx = c2 + 1
            else:
                mps.sites[c2 - 1] = np.tensordot(mps.sites[c2 - 1], uq, axes=1)
                mps.sites[c2] = vq.T.reshape([vq.shape[1]] + list(shape1[1:]))
                mps.qn[c2] = np.array(qnrset).reshape(len(qnrset), -1)
                mps.qnidx = c2 - 1
        mps.switch_direction()
    normalize_mps_only(mps)
    return mps, local_steps


# ---------------------------------------------------------------------------
# norms and observables


def dot(sites_a, sites_b):
    """<a|b> with a given as already-conjugated sites, reference mp.py:933-956."""
    e0 = np.eye(1, 1)
    for mt1, mt2 in zip(sites_a, sites_b):
        e0 = np.tensordot(e0, mt2, 1)
        if mt1.ndim == 4:   # density-operator site: system and ancilla summed
            e0 = np.tensordot(e0, mt1, ([0, 1, 2], [0, 1, 2])).T
        else:
            e0 = np.tensordot(e0, mt1, ([0, 1], [0, 1])).T
    return complex(e0[0, 0])


def mp_norm(state):
    """reference mp.py:354-372."""
    res = dot([s.conj() for s in state.sites], state.sites).real
    if res < 0:
        assert abs(res) < 1e-8
        res = 0
    return float(np.sqrt(res))


def normalize_mps_only(state):
    """normalize("mps_only"), reference mps.py:2025-2059 + scale mp.py:984-994."""
    nrm = mp_norm(state)
    state.sites[state.qnidx] = state.sites[state.qnidx] * (1.0 / nrm)
    return state


def calc_1site_rdm(state, idx):
    """One-site reduced density matrix, reference mps.py:1547-1598 (identity
    MPO environments).  Pinned by: trace 1, Hermiticity, dense rdm."""
    n = len(state)
    ident = [np.eye(s.shape[1]).reshape(1, s.shape[1], s.shape[1], 1) for s in state.sites]
    env = Environ(state.sites, ident, "R")
    ltensor = np.ones((1, 1, 1))
    for ims in range(idx):
        ltensor = contract_one_site(ltensor, state.sites[ims], ident[ims], "L")
    rtensor = env.read("R", idx + 1) if idx + 1 < n else np.ones((1, 1, 1))
    ms = state.sites[idx]
    lt = ltensor.reshape(ltensor.shape[0], ltensor.shape[-1])
    rt = rtensor.reshape(rtensor.shape[0], rtensor.shape[-1])
    t = np.tensordot(lt, ms.conj(), ([0], [0]))
    t = np.tensordot(t, rt, ([-1], [0]))
    t = np.tensordot(t, ms, ([0, -1], [0, -1]))
    return t


def expectation(state, mpo):
    """<psi|O|psi>, reference mps.py:471-525; for 4-leg density-operator
    sites the ancilla index is traced between bra and ket (mpdm.py:99-122)."""
    conj = [s.conj() for s in state.sites]
    env = Environ(state.sites, mpo, "R", conj)
    r = env.read("R", 1) if len(state) > 1 else np.ones((1, 1, 1))
    if state.sites[0].ndim == 4:
        val = np.einsum("abc,cflh,bdfg,adle,egh->", np.ones((1, 1, 1)), state.sites[0], mpo[0],
                        conj[0], r)
    else:
        val = np.einsum("abc,cfh,bdfg,ade,egh->", np.ones((1, 1, 1)), state.sites[0], mpo[0], conj[0], r)
    if np.isclose(float(val.imag), 0):
        return float(val.real)
    return complex(val)


def energy_bond(state, mpo, bond):
    """L[bond-1] . R[bond] contraction = <H> for a normalised MPS, the
    invariant of reference test_mp.py:51-62."""
    env = Environ(state.sites, mpo)
    l = env.read("L", bond - 1)
    r = env.read("R", bond)
    return complex(np.tensordot(l, r, axes=([0, 1, 2], [0, 1, 2])))
