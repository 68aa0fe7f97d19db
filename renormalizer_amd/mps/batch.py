"""Several independent states evolved in one call (the batched-instance
workload: many small chains, e.g. disorder realisations or parameter scans).

Each state is evolved exactly as ``Mps.evolve`` evolves it alone, one after
another on the caller's stream, so the results are those of one-by-one
evolution.  Overlapping the states on several HIP streams, one host thread
each, stopped making progress in its first test on the device and no cause
was found in this code; that mode was taken out, and ``nstreams > 1``
raises until it comes back with a shown cause.
"""
from typing import List, Sequence, Union

from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps


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
    if nstreams > 1:
        raise NotImplementedError("evolving states on concurrent streams is not available; "
                                  "use nstreams=1")
    return [m.evolve(o, evolve_dt) for m, o in zip(mps_list, mpos)]
