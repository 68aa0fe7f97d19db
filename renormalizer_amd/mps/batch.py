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
import threading
from typing import List, Sequence, Union

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
