"""Gradients of a time-evolved expectation value with respect to Hamiltonian
coefficients, by the adjoint (costate) method on top of the device TDVP-PS
step.

For H(c) = H_0 + sum_k c_k V_k, psi(t) = exp(-i H t) psi_0 and an observable
O, Duhamel's formula gives

    d<O>(T) / dc_k = 2 Im int_0^T <lambda(t)| V_k |psi(t)> dt,
    lambda(t) = exp(+i H (T - t)) O psi(T),

so one forward trajectory of psi and one backward trajectory of the costate
lambda give every c_k at once.  Both trajectories are TDVP-PS time steps
(``Mps.evolve``, the device path; lambda runs with -dt); the states of the
forward trajectory are kept every ``checkpoint`` steps and the ones between
recomputed during the backward pass.  The time integral is a Gauss-Legendre
rule on every step: at node t_n + s dt, psi is advanced from psi(t_n) by one
step of s dt and lambda taken back from lambda(t_{n+1}) by one step of
-(1 - s) dt.  The overlaps <lambda| V_k |psi> of product operators V_k (the
spin-boson couplings sigma_z x_k) come from one left and one right chain of
environment updates (``native.env_update``) per node.

This is the continuous adjoint: it differentiates the dynamics the TDVP
steps approximate, not the discrete projector-splitting map itself (whose
derivative through the gauge QR of nearly rank-deficient site tensors is
ill-conditioned).  At full bond dimension the TDVP steps are exact and the
gradient equals the exact one up to the quadrature error, O(dt^(2 nodes)).
``tdvp_observable`` exposes the same computation as a torch.autograd.Function
(forward trajectory in ``forward``, costate pass in ``backward``).  The
reference computes no gradients."""
from typing import List, Sequence, Tuple

import numpy as np
import torch

from renormalizer_amd import native
from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps

# Gauss-Legendre nodes on (0, 1) and weights (summing to one)
_GAUSS = {
    1: ([0.5], [1.0]),
    2: ([0.5 - np.sqrt(3) / 6, 0.5 + np.sqrt(3) / 6], [0.5, 0.5]),
    3: ([0.5 - np.sqrt(15) / 10, 0.5, 0.5 + np.sqrt(15) / 10], [5 / 18, 8 / 18, 5 / 18]),
}

SiteOp = Tuple[int, np.ndarray]


def _apply_site_op(mps: Mps, op: SiteOp) -> Mps:
    """O |mps> for a single-site operator (site, d x d matrix)."""
    site, mat = op
    new = mps.copy()
    t = new._mp[site]
    m = torch.from_numpy(np.ascontiguousarray(mat, dtype=np.complex128)).to(t.device)
    new._mp[site] = torch.tensordot(m, t, dims=([1], [1])).permute(1, 0, 2).contiguous()
    return new


def _as_w(mat, dev):
    d = mat.shape[0]
    return torch.from_numpy(np.ascontiguousarray(np.real(mat), dtype=np.float64).reshape(1, d, d, 1)).to(dev)


def product_overlaps(bra: Mps, ket: Mps, common: SiteOp, ops: Sequence[SiteOp]) -> np.ndarray:
    """<bra| common (x) op_k |ket> for every single-site op_k of ``ops``
    (each on a site after the common one), and <bra| common |ket> as the
    last entry.  Real operator matrices (the device environment update
    takes a real MPO)."""
    n = len(ket)
    dev = ket._mp[0].device
    csite, cmat = common
    assert all(s > csite for s, _ in ops), "the per-k operators must follow the common one"
    for _, m in [common] + list(ops):
        if np.iscomplexobj(m) and np.abs(np.imag(m)).max() > 0:
            raise ValueError("real operator matrices only")
    conj = [b.conj().resolve_conj().contiguous() for b in bra._mp]
    one = torch.ones((1, 1, 1), dtype=torch.complex128, device=dev)
    ident = {}

    def eye(d):
        if d not in ident:
            ident[d] = _as_w(np.eye(d), dev)
        return ident[d]

    # right environments R[j] of sites j .. n-1 (identity operators)
    right = [None] * (n + 1)
    right[n] = one
    for j in range(n - 1, 0, -1):
        right[j] = native.env_update(right[j + 1], ket._mp[j].contiguous(), eye(ket._mp[j].shape[1]), "R",
                                     conj[j])
    # left environments L[j] of sites 0 .. j, the common operator at csite
    wc = _as_w(cmat, dev)
    left = [None] * n
    env = one
    for j in range(n):
        w = wc if j == csite else eye(ket._mp[j].shape[1])
        env = native.env_update(env, ket._mp[j].contiguous(), w, "L", conj[j])
        left[j] = env
    out = []
    for s, mat in ops:
        x = native.env_update(left[s - 1], ket._mp[s].contiguous(), _as_w(mat, dev), "L", conj[s])
        r = right[s + 1]
        out.append(complex((x.reshape(-1) * r.reshape(-1)).sum()))
    out.append(complex((left[n - 1].reshape(-1) * right[n].reshape(-1)).sum()))
    return np.array(out)


def forward_states(mps: Mps, mpo: Mpo, dt: float, nsteps: int, checkpoint: int = 1):
    """The forward TDVP-PS trajectory from ``mps`` (not modified): psi(T) and
    the states kept every ``checkpoint`` steps (and at T), keyed by step."""
    if checkpoint < 1:
        raise ValueError("checkpoint stride must be positive")
    psi = mps.copy()
    if not psi.on_device:
        psi.to_device()
    saved = {0: psi.copy()}
    for n in range(nsteps):
        psi = psi.evolve(mpo, dt)
        if (n + 1) % checkpoint == 0 or n + 1 == nsteps:
            saved[n + 1] = psi.copy()
    return psi, saved


def costate_gradient(psi_t: Mps, saved, mpo: Mpo, dt: float, nsteps: int, obs: SiteOp, common: SiteOp,
                     couplings: List[SiteOp], nodes: int = 3, checkpoint: int = 1, log=None) -> np.ndarray:
    """The backward pass: [d<O>(T)/dc_k] from psi(T) and the checkpoints of
    ``forward_states`` (the same ``checkpoint`` stride), the costate lambda
    running back from O psi(T) and the states between checkpoints
    recomputed segment by segment."""
    if nodes not in _GAUSS:
        raise ValueError(f"nodes must be one of {sorted(_GAUSS)}")
    s_nodes, w_nodes = _GAUSS[nodes]
    lam = _apply_site_op(psi_t, obs)
    grad = np.zeros(len(couplings))
    # backward over the checkpoint segments, each recomputed from its start
    for seg_start in reversed(range(0, nsteps, checkpoint)):
        seg_end = min(seg_start + checkpoint, nsteps)
        states = {seg_start: saved[seg_start]}
        cur = saved[seg_start].copy()
        for n in range(seg_start, seg_end - 1):
            cur = cur.evolve(mpo, dt)
            states[n + 1] = cur.copy()
        for n in range(seg_end - 1, seg_start - 1, -1):
            acc = np.zeros(len(couplings), dtype=np.complex128)
            for s, w in zip(s_nodes, w_nodes):
                psi_s = states[n].copy().evolve(mpo, s * dt)
                lam_s = lam.copy().evolve(mpo, -(1.0 - s) * dt)
                acc += w * product_overlaps(lam_s, psi_s, common, couplings)[:-1]
            grad += 2.0 * dt * acc.imag
            lam = lam.evolve(mpo, -dt)
        if log is not None:
            log(f"adjoint: steps {seg_start}..{seg_end - 1} of {nsteps} done")
    return grad


def observable_gradient(mps: Mps, mpo: Mpo, dt: float, nsteps: int, obs: SiteOp, common: SiteOp,
                        couplings: List[SiteOp], nodes: int = 3, checkpoint: int = 1, log=None):
    """(<O>(T), [d<O>(T)/dc_k]) for T = nsteps * dt, the single-site
    observable O = ``obs`` and the perturbations V_k = common (x)
    couplings[k] (single-site operators; the spin-boson couplings
    sigma_z x_k, whose coefficients are c_k = -omega_k^2 dis_k, reference
    model/model.py:436).  ``mps`` is the device initial state (not
    modified), ``mpo`` the Hamiltonian; the forward states are kept every
    ``checkpoint`` steps; ``log`` (a callable taking a string) hears of
    the progress once per checkpoint segment.  The propagation runs in the
    precision of ``mps.evolve_config``."""
    if nodes not in _GAUSS:
        raise ValueError(f"nodes must be one of {sorted(_GAUSS)}")
    psi, saved = forward_states(mps, mpo, dt, nsteps, checkpoint)
    value = product_overlaps(psi, psi, obs, [])[-1].real
    return value, costate_gradient(psi, saved, mpo, dt, nsteps, obs, common, couplings, nodes, checkpoint, log)


class TdvpObservable(torch.autograd.Function):
    """<O>(T) of the TDVP-PS propagation as a differentiable function of the
    coupling constants c (a 1-D float64 tensor): ``forward`` builds H(c)
    through ``mpo_of`` and runs the forward trajectory, keeping the
    checkpoints; ``backward`` runs the costate pass (costate_gradient) and
    returns grad_output * d<O>(T)/dc."""

    @staticmethod
    def forward(ctx, c, mpo_of, mps, dt, nsteps, obs, common, couplings, nodes, checkpoint):
        mpo = mpo_of(c.detach().cpu().numpy())
        psi, saved = forward_states(mps, mpo, dt, nsteps, checkpoint)
        ctx.state = (psi, saved, mpo, dt, nsteps, obs, common, couplings, nodes, checkpoint)
        return torch.tensor(product_overlaps(psi, psi, obs, [])[-1].real, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_output):
        psi, saved, mpo, dt, nsteps, obs, common, couplings, nodes, checkpoint = ctx.state
        ctx.state = None
        g = costate_gradient(psi, saved, mpo, dt, nsteps, obs, common, couplings, nodes, checkpoint)
        return (grad_output * torch.from_numpy(g),) + (None,) * 9


def tdvp_observable(c: torch.Tensor, mpo_of, mps: Mps, dt: float, nsteps: int, obs: SiteOp, common: SiteOp,
                    couplings: List[SiteOp], nodes: int = 3, checkpoint: int = 1) -> torch.Tensor:
    """<O>(T) with autograd support in the couplings ``c``: H(c) = mpo_of(c)
    must equal H_0 + sum_k c_k common (x) couplings[k] (for the spin-boson
    model, ``spin_boson_mpo_of``)."""
    return TdvpObservable.apply(c, mpo_of, mps, dt, nsteps, obs, common, couplings, nodes, checkpoint)


def spin_boson_mpo_of(model):
    """c -> the Mpo of ``model`` (a SpinBosonModel) with the couplings
    c_k = -omega_k^2 dis_k of sigma_z x_k replaced by c (reference
    model/model.py:436); spin splitting, tunnelling, frequencies and levels
    kept."""
    from renormalizer_amd.model import Phonon, SpinBosonModel
    from renormalizer_amd.utils import Quantity

    def build(c):
        phs = [Phonon.simple_phonon(Quantity(ph.omega[0]), Quantity(-float(ck) / ph.omega[1] ** 2), ph.n_phys_dim)
               for ph, ck in zip(model.ph_list, c)]
        return Mpo(SpinBosonModel(Quantity(model.epsilon), Quantity(model.delta), phs))

    return build


def spin_boson_couplings(model) -> List[SiteOp]:
    """x_k of every mode of a spin-boson model (sites 1 ..)."""
    return [(i, np.asarray(model.basis[i].op_mat("x"), dtype=np.float64)) for i in range(1, len(model.basis))]


def spin_boson_sigma_z(model) -> SiteOp:
    return (0, np.asarray(model.basis[0].op_mat("sigma_z"), dtype=np.float64))
