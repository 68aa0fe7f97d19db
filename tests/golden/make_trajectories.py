"""Write the oracle's full trajectories of the C1, C2 and C5 workloads as
fixtures for the device parity tests (tests/test_trajectory_gpu.py).

Each workload starts from the state the tests build (tests/helpers.py,
host expansion of the reference's initial state); the numpy oracle
(oracle/tdvp.py) evolves it for the workload's full length and the file
keeps, per time step, the observables (<sigma_z> of the spin, or the
electronic populations) and the Lanczos matvec count of every local step,
plus fingerprints of the initial state (the norms of its site tensors) so a
test can tell that it started from the same state.

Usage (CPU only, minutes):  python tests/golden/make_trajectories.py [c1 c2 c5]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import tdvp as orc  # noqa: E402
from tests import helpers  # noqa: E402

WORKLOADS = {
    # name: (builder, dt, steps, observable)
    "c1": (lambda: helpers.sbm_initial(helpers.sbm_model(16, 8, alpha=0.5, omega_c=5.0), 16), 0.05, 100, "sz"),
    "c2": (lambda: helpers.holstein_initial(helpers.holstein_model(16, 10, j=-1.0, g=1.0), 64), 0.02, 200,
           "occ"),
    "c5": (lambda: helpers.sbm_initial(helpers.sbm_model(32, 8, alpha=0.4, omega_c=5.0), 64), 0.05, 200, "sz"),
}


def observe(kind, st, model):
    if kind == "sz":
        r = orc.calc_1site_rdm(st, 0)
        return np.array([(r[0, 0] - r[1, 1]).real])
    return helpers.e_occupations_oracle(st, model)


def fingerprint(mps):
    return np.array([np.linalg.norm(s) for s in mps.numpy_sites()])


def make(name):
    build, dt, steps, kind = WORKLOADS[name]
    mps, mpo = build()
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    obs = [observe(kind, st, mps.model)]
    counts = []
    t0 = time.perf_counter()
    for k in range(steps):
        st, ls = orc.tdvp_ps(st, w, dt)
        obs.append(observe(kind, st, mps.model))
        counts.append(ls)
        if (k + 1) % 10 == 0:
            print(f"{name}: {k + 1}/{steps} steps, {time.perf_counter() - t0:.0f} s", flush=True)
    path = os.path.join(HERE, f"trajectory_{name}.npz")
    np.savez_compressed(path, obs=np.array(obs), counts=np.array(counts, dtype=np.int32), dt=dt,
                        steps=steps, kind=kind, fingerprint=fingerprint(mps),
                        bond_dims=np.array(mps.bond_dims))
    print(f"wrote {path}")


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(WORKLOADS)):
        make(name)
