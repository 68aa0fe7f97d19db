"""Full trajectories of the C1, C2 and C5 workloads on the device against the
numpy oracle's trajectories (fixtures written by
tests/golden/make_trajectories.py from the same initial states).

FP64: <sigma_z> / the electronic populations within 1e-8 at every one of
the 100 or 200 steps, the Lanczos matvec counts of every local step as in
_assert_steps_match.  Mixed mode: within the metric's 1e-6 of the oracle
over the whole trajectory."""
import os

import numpy as np
import pytest

from renormalizer_amd import EvolveConfig, EvolveMethod
from tests import helpers
from tests.test_tdvp_gpu import _assert_steps_match

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

BUILD = {
    "c1": lambda: helpers.sbm_initial(helpers.sbm_model(16, 8, alpha=0.5, omega_c=5.0), 16),
    "c2": lambda: helpers.holstein_initial(helpers.holstein_model(16, 10, j=-1.0, g=1.0), 64),
    "c5": lambda: helpers.sbm_initial(helpers.sbm_model(32, 8, alpha=0.4, omega_c=5.0), 64),
}


def _observe(kind, mps):
    if kind == "sz":
        r = mps.calc_1site_rdm(0)[0]
        return np.array([(r[0, 0] - r[1, 1]).real])
    return np.asarray(mps.e_occupations)


def _device_trajectory(name, precision):
    fx = np.load(os.path.join(GOLDEN, f"trajectory_{name}.npz"))
    mps, mpo = BUILD[name]()
    fp = np.array([np.linalg.norm(s) for s in mps.numpy_sites()])
    np.testing.assert_allclose(fp, fx["fingerprint"], rtol=1e-10)
    assert list(mps.bond_dims) == fx["bond_dims"].tolist()
    kind, dt, steps = str(fx["kind"]), float(fx["dt"]), int(fx["steps"])
    dev = mps.copy()
    dev.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, precision=precision)
    dev.to_device()
    obs, counts = [_observe(kind, dev)], []
    for _ in range(steps):
        dev = dev.evolve(mpo, dt)
        obs.append(_observe(kind, dev))
        counts.append(list(dev.evolve_config.local_steps))
    return np.array(obs), counts, fx


@pytest.mark.parametrize("name", ["c1"])
def test_full_trajectory_fp64(name):
    print(f"[trajectory {name} fp64]", flush=True)
    obs, counts, fx = _device_trajectory(name, "fp64")
    err = np.abs(obs - fx["obs"]).max(axis=1)
    print(f"[trajectory {name} fp64] max |device - oracle| {err.max():.2e} over {len(err) - 1} steps",
          flush=True)
    assert err.max() <= 1e-8, (int(err.argmax()), err.max())
    _assert_steps_match(counts, fx["counts"].tolist())


@pytest.mark.parametrize("name", ["c1"])
def test_full_trajectory_mixed(name):
    print(f"[trajectory {name} mixed]", flush=True)
    obs, counts, fx = _device_trajectory(name, "mixed")
    err = np.abs(obs - fx["obs"]).max()
    print(f"[trajectory {name} mixed] max |device - oracle| {err:.2e}", flush=True)
    assert err <= 1e-6
