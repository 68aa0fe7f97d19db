"""TdMpsJob (reference utils/tdmps.py:20-224) on the host: the step plan of
``evolve``, the crash-safe dump (.bak rotation) and dump_mps "one"/"all"
file names, with a stand-in state (no device needed)."""
import os

import numpy as np
import pytest

from renormalizer_amd.utils.tdmps import TdMpsJob, _step_plan


class _State:
    def __init__(self, x):
        self.x = x

    def dump(self, fname):
        np.savez(fname, x=self.x)

    def __str__(self):
        return f"_State({self.x})"


class _Job(TdMpsJob):
    """x(t) = 2 t, stopping once x reaches ``stop``."""

    def __init__(self, stop=None, **kw):
        self.stop = stop
        self.values = []
        super().__init__(**kw)

    def init_mps(self):
        return _State(0.0)

    def process_mps(self, mps):
        self.values.append(mps.x)

    def evolve_single_step(self, dt):
        return _State(self.latest_mps.x + 2 * dt)

    def get_dump_dict(self):
        return {"time series": self.evolve_times, "x": self.values}

    def stop_evolve_criteria(self):
        return self.stop is not None and self.latest_mps.x >= self.stop


def test_step_plan_cases():
    assert _step_plan(0.1, 5, None) == (0.1, 5, True)
    assert _step_plan(0.1, 5, 9.0) == (0.1, 5, True)       # nsteps wins
    dt, n, bounded = _step_plan(None, 4, 1.0)
    assert dt == 0.25 and n == 4 and bounded
    assert _step_plan(0.3, None, 1.0) == (0.3, 4, True)     # 1.0 // 0.3 + 1
    assert _step_plan(0.1, None, None) == (0.1, None, False)
    with pytest.raises(ValueError):
        _step_plan(None, None, 1.0)
    with pytest.raises(ValueError):
        _step_plan(None, 3, None)


def test_evolve_records_times_and_stops():
    job = _Job().evolve(0.5, nsteps=4)
    np.testing.assert_allclose(job.evolve_times, [0, 0.5, 1.0, 1.5, 2.0])
    np.testing.assert_allclose(job.values, [0, 1, 2, 3, 4])
    job = _Job(stop=3.0).evolve(0.5)        # unbounded until the criterion holds
    assert job.latest_mps.x == 3.0 and len(job.evolve_times) == 4


def test_bad_dump_mps_mode():
    with pytest.raises(ValueError):
        _Job(dump_mps="some")
    with pytest.raises(ValueError):
        _Job().dump_dict()                  # no dump dir / job name


@pytest.mark.parametrize("mode", ["one", "all"])
def test_dump_rotation_and_state_files(tmp_path, mode):
    d = str(tmp_path)
    job = _Job(dump_dir=d, job_name="j", dump_mps=mode).evolve(0.25, nsteps=3)
    with np.load(os.path.join(d, "j.npz")) as f:
        np.testing.assert_allclose(f["x"], [0, 0.5, 1.0, 1.5])
        np.testing.assert_allclose(f["time series"], [0, 0.25, 0.5, 0.75])
    assert not os.path.exists(os.path.join(d, "j.npz.bak"))
    if mode == "one":
        with np.load(os.path.join(d, "j_mps.npz")) as f:
            assert float(f["x"]) == 1.5
    else:
        for k, x in ((1, 0.5), (2, 1.0), (3, 1.5)):
            with np.load(os.path.join(d, f"j_mps_{k}.npz")) as f:
                assert float(f["x"]) == x


def test_dump_keeps_previous_file_until_new_one_is_complete(tmp_path, monkeypatch):
    """If writing the new file fails, the previous results survive as .bak."""
    d = str(tmp_path)
    job = _Job(dump_dir=d, job_name="j").evolve(0.25, nsteps=1)

    def broken_savez(*a, **k):
        raise IOError("disk full")

    monkeypatch.setattr(np, "savez", broken_savez)
    job.evolve(0.25, nsteps=1)              # the IOError is logged, the run goes on
    assert len(job.evolve_times) == 3
    monkeypatch.undo()
    with np.load(os.path.join(d, "j.npz.bak")) as f:
        np.testing.assert_allclose(f["x"], [0, 0.5])
