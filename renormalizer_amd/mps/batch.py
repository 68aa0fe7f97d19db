"""Several independent states evolved in one call (the batched-instance
workload: many small chains, e.g. disorder realisations or parameter scans).

Each state is evolved exactly as ``Mps.evolve`` evolves it alone.  With
``nstreams > 1`` the states are dealt round-robin to that many HIP streams,
one host thread per stream, so the kernels of different states overlap on
the device while each thread's host bookkeeping runs.  What makes this safe:

* every (device, stream) pair has its own native workspace (Krylov basis,
  state block and split-K partials never alias between streams);
* state ``i`` always runs on stream ``i % nstreams``, so the tensors one
  call leaves behind (sites, cached environments) are freed, in the next
  call, on the stream that made them, and torch's allocator hands their
  memory only to later work of that same stream;
* the device index arrays of the QR blocks are cached per stream, and the
  shared host caches (QN plans, identity operators, MPO device copies) are
  locked;
* the caller's stream is synchronised before the threads start and every
  worker stream after they finish.
"""
import ctypes
import threading
from typing import List, Sequence, Union

import numpy as np
import torch

from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps

_STREAMS = {}
_STREAMS_LOCK = threading.Lock()


def _worker_streams(device, n):
    """The same ``n`` streams (default priority) for every call on a device."""
    key = (device.index if device.index is not None else torch.cuda.current_device(), n)
    with _STREAMS_LOCK:
        if key not in _STREAMS:
            _STREAMS[key] = [torch.cuda.Stream(device=device) for _ in range(n)]
        return _STREAMS[key]


def evolve_batch(mps_list: Sequence[Mps], mpo: Union[Mpo, Sequence[Mpo]], evolve_dt,
                 nstreams: int = 1) -> List[Mps]:
    """``[m.evolve(mpo_i, evolve_dt) for m in mps_list]``; ``mpo`` is one
    operator for all states or one per state."""
    n = len(mps_list)
    mpos = list(mpo) if isinstance(mpo, (list, tuple)) else [mpo] * n
    if len(mpos) != n:
        raise ValueError(f"{len(mpos)} operators for {n} states")
    if nstreams < 1:
        raise ValueError("nstreams must be at least 1")
    if nstreams == 1 or n <= 1:
        return [m.evolve(o, evolve_dt) for m, o in zip(mps_list, mpos)]
    device = torch.device("cuda", torch.cuda.current_device())
    streams = _worker_streams(device, nstreams)
    out: List = [None] * n
    errors: List = []
    torch.cuda.current_stream(device).synchronize()

    def work(k):
        try:
            torch.cuda.set_device(device)
            with torch.cuda.stream(streams[k]):
                for i in range(k, n, nstreams):
                    out[i] = mps_list[i].evolve(mpos[i], evolve_dt)
        except BaseException as e:  # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(min(nstreams, n))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in streams:
        s.synchronize()
    if errors:
        raise errors[0]
    return out


# ---------------------------------------------------------------------------
# grouped sweep: every state's site update at one sweep position in a single
# launch (reno_batch_site_step), one workgroup per state


class _BatchSite(ctypes.Structure):
    """reno_batch_site of include/reno_hip.h."""
    _fields_ = ([(f, ctypes.c_int) for f in ("active", "dir", "last", "Ml", "d", "Mr", "Dl", "Dr", "kt",
                                               "nblocks", "nbr_a", "nbr_b", "D0", "kcap", "pad0", "pad1")]
                + [(f, ctypes.c_void_p) for f in ("L", "R", "Wl", "W", "v0", "kout", "site_out", "bond",
                                                  "env_in", "env_out", "L0", "R0", "bond_out", "nbr",
                                                  "carry")]
                + [("brows", ctypes.c_void_p * 4), ("bcols", ctypes.c_void_p * 4),
                   ("bnr", ctypes.c_int * 4), ("bnc", ctypes.c_int * 4),
                   ("scratch", ctypes.c_void_p), ("counts", ctypes.c_void_p)])


def _qn_step_plan(mps):
    """The quantum-number bookkeeping of one TDVP-PS time step without
    tensors (the sweep of Mps._tdvp_ps_sweep): per half-sweep the list of
    (site, direction or None for the last site, svd_qn plan) in sweep order,
    and the state's bond quantum numbers after the step."""
    from renormalizer_amd.mps.svd_qn import svd_qn_plan
    sim = mps.metacopy()
    n = len(sim)
    halves = []
    for _ in range(2):
        seq = []
        for imps in sim.iter_idx_list(full=True):
            qnbigl, qnbigr, _ = sim._get_big_qn([imps], with_qnmat=False)
            if not sim.to_right and imps != 0:
                plan = svd_qn_plan(qnbigl, qnbigr, sim.qntot)
                kt = len(plan[4])
                sim.qn[imps] = np.array(plan[4], dtype=int).reshape(kt, -1)
                sim.qnidx = imps - 1
                seq.append((imps, 0, plan))
            elif sim.to_right and imps != n - 1:
                plan = svd_qn_plan(qnbigl, qnbigr, sim.qntot)
                kt = len(plan[3])
                sim.qn[imps + 1] = np.array(plan[3], dtype=int).reshape(kt, -1)
                sim.qnidx = imps + 1
                seq.append((imps, 1, plan))
            else:
                seq.append((imps, None, None))
        sim._switch_direction()
        halves.append(seq)
    return halves, sim


class BatchEvolution:
    """Many independent states advanced together by one-site TDVP-PS time
    steps, every sweep position of every state in one launch
    (``reno_batch_site_step``: one workgroup per state runs the whole site
    update on the device).  FP64; real time steps.

    The grouped sweep keeps every tensor in place, so it needs states whose
    shapes and bond quantum numbers no longer change from step to step: the
    constructor first advances each state by ordinary ``Mps.evolve`` steps
    (``warmup``, at least one; more while the quantum numbers still move) and
    then checks that one step's quantum-number bookkeeping maps the bonds
    onto themselves with unchanged dimensions.  ``states`` afterwards holds
    the evolving states (device tensors updated in place)."""

    def __init__(self, mps_list: Sequence[Mps], mpo: Union[Mpo, Sequence[Mpo]], evolve_dt, warmup: int = 1,
                 kcap: int = 64, nstreams: int = 1):
        from renormalizer_amd import native
        self.native = native
        n = len(mps_list)
        self.mpos = list(mpo) if isinstance(mpo, (list, tuple)) else [mpo] * n
        if len(self.mpos) != n:
            raise ValueError(f"{len(self.mpos)} operators for {n} states")
        self.dt = float(evolve_dt)
        if np.iscomplex(evolve_dt):
            raise NotImplementedError("the grouped sweep runs real time steps")
        self.kcap = int(kcap)
        states = list(mps_list)
        for m in states:
            if m.evolve_config.precision != "fp64":
                raise NotImplementedError("the grouped sweep runs in FP64")
        for _ in range(max(1, warmup)):
            states = evolve_batch(states, self.mpos, self.dt, nstreams=nstreams)
        plans = []
        for i, m in enumerate(states):
            for _ in range(4):
                halves, after = _qn_step_plan(m)
                if self._static(m, halves, after):
                    break
                m = m.evolve(self.mpos[i], self.dt)
                states[i] = m
            else:
                raise RuntimeError(f"state {i}: bond quantum numbers still change after the warm-up steps")
            plans.append(halves)
        self.states = states
        self._build(plans)
        self.local_steps = None

    @staticmethod
    def _static(m, halves, after):
        if after.qnidx != m.qnidx or after.to_right != m.to_right:
            return False
        for a, b in zip(after.qn, m.qn):
            if not np.array_equal(np.asarray(a).reshape(len(a), -1), np.asarray(b).reshape(len(b), -1)):
                return False
        for seq in halves:
            for site, direction, plan in seq:
                if direction is None:
                    continue
                shape = m._mp[site].shape
                kt = len(plan[4]) if direction == 0 else len(plan[3])
                if kt != (shape[0] if direction == 0 else shape[2]) or len(plan[2]) > 4:
                    return False
        return m._env_cache is not None

    def _build(self, plans):
        native = self.native
        lib = native.lib()
        dev = self.states[0]._mp[0].device
        self.device = dev
        self._keep = []
        nst = len(self.states)
        self.npos = max(len(m) for m in self.states)
        self.counts = torch.zeros((nst, 2, self.npos, 2), dtype=torch.int32, device=dev)
        tables = [[(_BatchSite * nst)() for _ in range(self.npos)] for _ in range(2)]
        centre, lens = [], []
        for b, (m, halves) in enumerate(zip(self.states, plans)):
            n = len(m)
            m._mp = [t.contiguous() for t in m._mp]
            sites = m._mp
            environ = m._env_cache[2]
            w = self.mpos[b].device_tensors(native.PREC_FP64, dev)
            wl = [x.permute(1, 3, 0, 2).reshape(x.shape[1] * x.shape[3], x.shape[0] * x.shape[2]).contiguous()
                  for x in w]
            big = max(t.numel() for t in sites)
            bond = max(t.shape[0] for t in sites) ** 2
            bufs = {k: torch.empty(big if k in ("kout", "carry") else bond, dtype=torch.complex128, device=dev)
                    for k in ("kout", "carry", "bond", "bond_out")}
            scratch_bytes = max(lib.reno_batch_scratch_bytes(t.shape[0], t.shape[1], t.shape[2], x.shape[0],
                                                             x.shape[3], self.kcap) for t, x in zip(sites, w))
            scratch = torch.empty(scratch_bytes, dtype=torch.uint8, device=dev)
            self._keep += [wl, bufs, scratch]
            env = {}
            for key in ("L", "R"):
                rng = range(-1, n - 1) if key == "L" else range(1, n + 1)
                for i in rng:
                    t = environ.read(key, i)
                    if t.dtype != torch.complex128 or not t.is_contiguous():
                        raise RuntimeError("the grouped sweep needs contiguous complex128 environments")
                    env[(key, i)] = t
            for h, seq in enumerate(halves):
                for p, (i, direction, plan) in enumerate(seq):
                    s = tables[h][p][b]
                    site = sites[i]
                    Ml, d, Mr = site.shape
                    Dl, Dr = w[i].shape[0], w[i].shape[3]
                    last = direction is None
                    s.active, s.last, s.kcap = 1, int(last), self.kcap
                    s.dir = 0 if last else direction
                    s.Ml, s.d, s.Mr, s.Dl, s.Dr = Ml, d, Mr, Dl, Dr
                    s.L = env[("L", i - 1)].data_ptr()
                    s.R = env[("R", i + 1)].data_ptr()
                    s.Wl = wl[i].data_ptr()
                    s.W = w[i].data_ptr()
                    s.v0 = site.data_ptr() if p == 0 else bufs["carry"].data_ptr()
                    s.kout = site.data_ptr() if last else bufs["kout"].data_ptr()
                    s.counts = self.counts[b, h, p].data_ptr()
                    s.scratch = scratch.data_ptr()
                    if last:
                        continue
                    rows, cols, blocks, qnl, qnr = plan
                    s.kt = len(qnr) if direction == 0 else len(qnl)
                    s.nblocks = len(blocks)
                    for k, (ri, ci) in enumerate(blocks):
                        rt = torch.from_numpy(np.ascontiguousarray(ri, dtype=np.int64)).to(dev)
                        ct = torch.from_numpy(np.ascontiguousarray(ci, dtype=np.int64)).to(dev)
                        self._keep += [rt, ct]
                        s.brows[k], s.bcols[k] = rt.data_ptr(), ct.data_ptr()
                        s.bnr[k], s.bnc[k] = len(ri), len(ci)
                    s.site_out = site.data_ptr()
                    s.bond = bufs["bond"].data_ptr()
                    s.bond_out = bufs["bond_out"].data_ptr()
                    s.carry = bufs["carry"].data_ptr()
                    if direction == 1:
                        s.env_in = env[("L", i - 1)].data_ptr()
                        s.env_out = env[("L", i)].data_ptr()
                        s.L0, s.R0, s.D0 = env[("L", i)].data_ptr(), env[("R", i + 1)].data_ptr(), Dr
                        nxt = sites[i + 1]
                        s.nbr, s.nbr_b = nxt.data_ptr(), nxt.shape[1] * nxt.shape[2]
                    else:
                        s.env_in = env[("R", i + 1)].data_ptr()
                        s.env_out = env[("R", i)].data_ptr()
                        s.L0, s.R0, s.D0 = env[("L", i - 1)].data_ptr(), env[("R", i)].data_ptr(), Dl
                        prv = sites[i - 1]
                        s.nbr, s.nbr_a = prv.data_ptr(), prv.shape[0] * prv.shape[1]
            c = sites[m.qnidx]
            centre.append(c.data_ptr())
            lens.append(c.numel())
        size = ctypes.sizeof(_BatchSite) * nst
        self.tables = []
        for h in range(2):
            row = []
            for p in range(self.npos):
                raw = np.frombuffer(ctypes.string_at(ctypes.addressof(tables[h][p]), size), dtype=np.uint8)
                row.append(torch.from_numpy(raw.copy()).to(dev))
            self.tables.append(row)
        self.centre = torch.tensor(centre, dtype=torch.int64, device=dev)
        self.centre_len = torch.tensor(lens, dtype=torch.int64, device=dev)
        self.norms = torch.empty(nst, dtype=torch.float64, device=dev)

    def _launch_step(self):
        native = self.native
        lib = native.lib()
        stream = native._stream()
        nst = len(self.states)
        for h in range(2):
            for p in range(self.npos):
                native.check(lib.reno_batch_site_step(nst, ctypes.c_void_p(self.tables[h][p].data_ptr()),
                                                      self.dt, stream), "reno_batch_site_step")
        native.check(lib.reno_batch_normalize(nst, ctypes.c_void_p(self.centre.data_ptr()),
                                              ctypes.c_void_p(self.centre_len.data_ptr()),
                                              ctypes.c_void_p(self.norms.data_ptr()), stream),
                     "reno_batch_normalize")

    def capture(self):
        """Record one time step (all launches of both half-sweeps and the
        normalisation) as a HIP graph; ``step`` then replays it.  Every
        operand lives at a fixed address and the step allocates nothing and
        clears nothing, so each replay advances the states by one step."""
        stream = torch.cuda.Stream(device=self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            self._launch_step()
        torch.cuda.current_stream(self.device).wait_stream(stream)
        self.graph = graph
        return self

    graph = None

    def step(self):
        """One time step of every state: 2 x (longest chain) launches (or one
        graph replay after ``capture``), the centre sites normalised
        (normalize("mps_only")), one wait for the matvec counts at the end."""
        native = self.native
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch_step()
        counts = self.counts.cpu().numpy()
        if (counts < 0).any():
            bad = np.argwhere(counts < 0)[0]
            raise native.RenoNativeError(
                f"state {bad[0]}: Krylov space of {self.kcap} vectors exhausted or zero vector "
                f"(half {bad[1]}, position {bad[2]})")
        self.local_steps = []
        for b, m in enumerate(self.states):
            n = len(m)
            seq = []
            for h in range(2):
                for p in range(n):
                    seq.append(int(counts[b, h, p, 0]))
                    if p != n - 1:
                        seq.append(int(counts[b, h, p, 1]))
            self.local_steps.append(seq)
            m.evolve_config.local_steps = seq
        return self.states
