"""Several independent states evolved on concurrent streams
(renormalizer_amd.mps.batch.evolve_batch): bit-identical to one-by-one
evolution, on instances of the disordered-Holstein workload.  An earlier
version of this mode stopped making progress on the device: all streams
shared one native workspace, so two streams' Krylov calls reset each
other's state block and the convergence test never fired; each stream now
has its own workspace (tests/test_native_cpu.py checks that)."""
import time

import numpy as np
import pytest
import torch

from renormalizer_amd import EvolveConfig, EvolveMethod
from renormalizer_amd.mps.batch import evolve_batch
from tests import helpers

pytestmark = pytest.mark.gpu


def _instances(sizes, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for nmol, m in sizes:
        model = helpers.holstein_model(nmol, 8, j=-1.0, g=0.7, elocal=list(rng.normal(0.0, 0.5, nmol)))
        mps, mpo = helpers.holstein_initial(model, m)
        mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        mps.to_device()
        out.append((mps, mpo))
    return out


def test_batch_is_bit_identical_to_sequential():
    print("[batch evolve: 6 instances, 3 streams]", flush=True)
    inst = _instances([(6, 16), (8, 24), (10, 32), (7, 16), (9, 24), (12, 32)])
    seq = [m.copy() for m, _ in inst]
    bat = [m.copy() for m, _ in inst]
    mpos = [o for _, o in inst]
    for _ in range(3):
        seq = [m.evolve(o, 0.02) for m, o in zip(seq, mpos)]
        bat = evolve_batch(bat, mpos, 0.02, nstreams=3)
    torch.cuda.synchronize()
    for a, b in zip(seq, bat):
        assert list(a.evolve_config.local_steps) == list(b.evolve_config.local_steps)
        for x, y in zip(a.numpy_sites(), b.numpy_sites()):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(a.e_occupations, b.e_occupations)


def test_batch_throughput():
    """16 instances (8 to 20 molecules, M 16 to 48): time per step of all
    instances one by one and on 4 streams (reported, and the batch must not
    be slower)."""
    print("[batch evolve throughput: 16 instances]", flush=True)
    rng = np.random.default_rng(11)
    sizes = [(int(rng.integers(8, 21)), int(rng.choice([16, 32, 48]))) for _ in range(16)]
    inst = _instances(sizes, seed=3)
    mpos = [o for _, o in inst]
    states = [m for m, _ in inst]
    # two steps each way first: budgets and environments in steady state
    for _ in range(2):
        states = [m.evolve(o, 0.02) for m, o in zip(states, mpos)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        states = [m.evolve(o, 0.02) for m, o in zip(states, mpos)]
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / 2
    for _ in range(1):
        states = evolve_batch(states, mpos, 0.02, nstreams=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        states = evolve_batch(states, mpos, 0.02, nstreams=4)
    torch.cuda.synchronize()
    t_bat = (time.perf_counter() - t0) / 2
    print(f"[batch] 16 instances, one step each: sequential {1e3 * t_seq:.1f} ms, "
          f"4 streams {1e3 * t_bat:.1f} ms, x{t_seq / t_bat:.2f}", flush=True)
    assert t_bat < 1.1 * t_seq
