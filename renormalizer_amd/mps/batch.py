"""Several independent states evolved together (the batched-instance
workload: many small chains, e.g. disorder realisations or parameter scans).

Each state is evolved exactly as ``Mps.evolve`` evolves it alone, so the
results are bit-identical to one-by-one evolution.  The intended mode deals
the states to a few HIP streams (default priority), each driven by its own
host thread, so that the latency-bound kernels of small states overlap on
the device.  That mode stopped making progress in its first test on the
device, with no cause found in this code; until one is shown it is switched
off (``nstreams > 1`` raises) and the states run one after another on the
caller's stream.
"""
import threading
from typing import List, Sequence, Union

import torch

from renormalizer_amd import native
from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps


def evolve_batch(mps_list: Sequence[Mps], mpo: Union[Mpo, Sequence[Mpo]], evolve_dt,
                 nstreams: int = 1) -> List[Mps]:
    """``[m.evolve(mpo_i, evolve_dt) for m in mps_list]`` with the work of
    state i on stream i mod nstreams.  ``mpo`` is one operator for all states
    or one per state.  The caller's stream waits for every stream before the
    call returns, and the streams wait for the caller's stream before they
    start, so the states may come from and go back to ordinary stream-ordered
    code."""
    n = len(mps_list)
    if n == 0:
        return []
    mpos = list(mpo) if isinstance(mpo, (list, tuple)) else [mpo] * n
    if len(mpos) != n:
        raise ValueError(f"{len(mpos)} operators for {n} states")
    if nstreams < 1:
        raise ValueError("nstreams must be at least 1")
    if nstreams > 1:
        raise NotImplementedError("concurrent streams are switched off: the multi-threaded mode "
                                  "stopped making progress on the device; use nstreams=1")
    if nstreams == 1:
        return [m.evolve(o, evolve_dt) for m, o in zip(mps_list, mpos)]
    native.lib()  # load once, before the worker threads
    dev = torch.cuda.current_device()
    caller = torch.cuda.current_stream(dev)
    streams = [torch.cuda.Stream(device=dev) for _ in range(min(nstreams, n))]
    for s in streams:
        s.wait_stream(caller)
    out: List[Mps] = [None] * n
    errors = []

    def work(k):
        try:
            torch.cuda.set_device(dev)
            with torch.cuda.stream(streams[k]):
                for i in range(k, n, len(streams)):
                    out[i] = mps_list[i].evolve(mpos[i], evolve_dt)
        except BaseException as e:  # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(streams))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in streams:
        caller.wait_stream(s)
    if errors:
        raise errors[0]
    return out
