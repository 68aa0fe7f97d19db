"""The tile autotuner (reno_autotune): tuned matvecs give the untuned
results to rounding level, every record is a valid choice no slower than
the planner's measured time, a second call re-times nothing, and a TDVP-PS
trajectory with autotuning follows the untuned one."""
import numpy as np
import pytest
import torch

from renormalizer_amd import EvolveConfig, EvolveMethod, native
from tests import helpers

pytestmark = pytest.mark.gpu


def _ops(dev, Ml, d, Mr, D, prec, seed):
    g = torch.Generator().manual_seed(seed)
    cd = torch.complex128 if prec == native.PREC_FP64 else torch.complex64
    rd = torch.float64 if prec == native.PREC_FP64 else torch.float32

    def c(*shape):
        return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                             torch.randn(*shape, generator=g, dtype=torch.float64))
    L = c(Ml, D, Ml).to(cd).to(dev)
    R = c(Mr, D, Mr).to(cd).to(dev)
    W = torch.randn(D, d, d, D, generator=g, dtype=torch.float64).to(rd).to(dev)
    v1 = c(Ml, d, Mr).to(dev)
    v0 = c(Ml, Mr).to(dev)
    return L, W, R, v1, v0


@pytest.mark.parametrize("prec", [native.PREC_FP64, native.PREC_MIXED])
@pytest.mark.parametrize("shape", [(256, 10, 256, 3), (64, 10, 256, 3), (40, 2, 100, 4)])
def test_tuned_matvec_matches_planner(dev, prec, shape):
    Ml, d, Mr, D = shape
    print(f"[autotune prec={prec} shape={shape}]", flush=True)
    native.autotune_reset()
    L, W, R, v1, v0 = _ops(dev, Ml, d, Mr, D, prec, 7)
    ref1 = native.hop_apply(L, W, R, v1, prec)
    v0 = v0[:Ml, :Ml] if Ml != Mr else v0
    L0, R0 = L, (R if Ml == Mr else L)
    ref0 = native.hop_apply(L0, None, R0, v0.contiguous(), prec)
    recs = native.autotune(L, W, R, v1, prec, reps=4)
    recs += native.autotune(L0, None, R0, v0.contiguous(), prec, reps=4)
    assert len(recs) >= 3
    assert native.autotune_entries() >= len(recs)
    for r in recs:
        assert r["tile"] in (32, 64) and r["nsplit"] >= 1
        assert r["planned_tile"] in (32, 64) and r["planned_nsplit"] >= 1
        assert 0 < r["us_tuned"] <= r["us_planned"]
        assert r["prec"] == prec
    print("[autotune] " + "; ".join(
        f"{r['M']}x{r['N']}x{r['K']}*{r['batch']}: {r['tile']}/{r['nsplit']} {r['us_tuned']:.1f} us "
        f"(planner {r['planned_tile']}/{r['planned_nsplit']} {r['us_planned']:.1f} us)" for r in recs),
        flush=True)
    # already tuned: nothing re-timed
    assert native.autotune(L, W, R, v1, prec) == []
    got1 = native.hop_apply(L, W, R, v1, prec)
    got0 = native.hop_apply(L0, None, R0, v0.contiguous(), prec)
    tol = 1e-12 if prec == native.PREC_FP64 else 2e-5
    for got, ref in ((got1, ref1), (got0, ref0)):
        rel = (torch.linalg.norm(got - ref) / torch.linalg.norm(ref)).item()
        assert rel < tol, rel
    native.autotune_reset()
    assert native.autotune_entries() == 0


def test_tdvp_with_autotune_follows_untuned(dev):
    """Spin-boson, 16 modes x 8 levels, M = 32: 8 asynchronous TDVP-PS steps
    with and without autotuning; <sigma_z> within 1e-10."""
    print("[autotune tdvp sbm 16 modes M=32, 8 steps]", flush=True)
    model = helpers.sbm_model(16, 8)
    mps0, mpo = helpers.sbm_initial(model, 32)
    out = []
    for tune in (False, True):
        native.autotune_reset()
        m = mps0.copy()
        m.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, autotune=tune)
        m.to_device()
        sz = []
        for _ in range(8):
            m = m.evolve(mpo, 0.05)
            r = m.calc_1site_rdm(0)[0]
            sz.append((r[0, 0] - r[1, 1]).real)
        out.append(np.array(sz))
        if tune:
            assert native.autotune_entries() > 0
    native.autotune_reset()
    err = np.abs(out[0] - out[1]).max()
    print(f"[autotune tdvp] max |dsz| {err:.2e}", flush=True)
    assert err < 1e-10
