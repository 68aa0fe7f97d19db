#!/usr/bin/env python
"""TDVP-PS time steps ("sweeps") per second on one MI355X.

Workload (the metric's configuration): Ohmic spin-boson model, 1 spin + 64
harmonic modes, 10 levels each, alpha = 0.3, omega_c = 10, Wang-1
logarithmic discretisation, Delta = 1, eps = 0; MPS bond dimension M = 256;
dt = 0.025.  Initial state: spin up x vacuum expanded along H to M = 256
with coefficient 1e-16 (the reference's SpinBosonDynamics.init_mps), built
on the host before timing.  One step = Mps.evolve = two half-sweeps.

Both device precision modes run from the same initial MPS: "fp64"
(complex128 everywhere) and "mixed" (FP64 environments, vectors, reductions,
exponential and QR; each Lanczos iteration's matvec on FP32 matrix cores
once the step's result weighs the new Lanczos vector below 1e-2).  ``value``
is the rate of the faster mode whose largest <sigma_z> difference from the
FP64 trajectory stays within 1e-6 (the FP64 mode itself when the mixed mode
misses that bound); both rates, their ratio, the difference and the matvec
counts are reported beside it.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

# host BLAS threads for the CPU baseline: all cores we may use, set before
# numpy is imported (the reference does the same with RENO_NUM_THREADS)
_CORES = len(os.sched_getaffinity(0))
os.environ.setdefault("RENO_NUM_THREADS", str(_CORES))
for _k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
    os.environ[_k] = os.environ["RENO_NUM_THREADS"]

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"double": 78.6, "float": 157.3}  # MI355X dense matrix FP64 / FP32
ACCURACY_BOUND = 1e-6  # largest allowed |<sigma_z>_mixed - <sigma_z>_fp64| over the run


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(nmode=64, nlev=10, m=256, cache=None, on_device=True):
    """The C4 model, its MPO and the expanded initial state, expanded on the
    device (``on_device``) or on the host.  With ``cache`` (a .npz path) the
    initial state is loaded from there when the file exists (Mps.load: the
    same tensors and quantum numbers bit for bit) and written there after
    the expansion otherwise."""
    from renormalizer_amd import (CompressConfig, CompressCriteria, EvolveConfig, EvolveMethod, Mpo, Mps,
                                  Phonon, Quantity, SpinBosonModel)
    from renormalizer_amd.sbm import SpectralDensityFunction
    sdf = SpectralDensityFunction(0.3, 10.0)
    w, c2 = sdf.Wang1(nmode)
    omegas, dis = sdf.post_process(w, c2)
    model = SpinBosonModel(Quantity(0.0), Quantity(1.0),
                           [Phonon.simple_phonon(o, d, nlev) for o, d in zip(omegas, dis)])
    mpo = Mpo(model)
    if cache and os.path.exists(cache):
        mps = Mps.load(model, cache)
    else:
        mps = Mps.ground_state(model, False)
        mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=m)
        mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        if on_device:
            mps.to_device()
        mps = mps.expand_bond_dimension(mpo, coef=1e-16, include_ex=False)
        if cache:
            mps.dump(cache)
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=m)
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
    return model, mpo, mps


def sigma_z(mps):
    r = mps.calc_1site_rdm(0)[0]
    return float((r[0, 0] - r[1, 1]).real)


PROF_PERIOD = 16  # event-time a pseudo-random 1 in 16 of the GEMM launches


def run_mode(mps0, mpo, precision, dt, warmup, steps, profile, tau=None, autotune=False):
    """Evolve a copy of mps0 by warmup + steps time steps.  Each timed step is
    bracketed by device synchronisations and only the evolution itself is
    timed; <sigma_z> is taken after each step outside the timed region, and
    only the current state is kept (earlier states are released, so the
    allocator recycles their memory as a long run would)."""
    import torch
    from renormalizer_amd import EvolveConfig, EvolveMethod, Mps, native
    mps = mps0.copy()
    mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps, precision=precision)
    if tau is not None:
        mps.evolve_config.mixed_tau = tau
    mps.evolve_config.autotune = autotune
    mps.to_device()
    sz = [sigma_z(mps)]
    redo0 = Mps.krylov_redo_count
    for _ in range(warmup):
        mps = mps.evolve(mpo, dt)
        sz.append(sigma_z(mps))
    torch.cuda.synchronize()
    redo1 = Mps.krylov_redo_count
    seconds = 0.0
    matvecs = []
    low = []
    for k in range(steps):
        if profile:
            # a new record for the first step, resumed for the later ones
            native.prof_enable(PROF_PERIOD if k == 0 else -PROF_PERIOD)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mps = mps.evolve(mpo, dt)
        torch.cuda.synchronize()
        seconds += time.perf_counter() - t0
        if profile:
            native.prof_enable(0)
        matvecs.append(sum(mps.evolve_config.local_steps))
        low.append(getattr(mps.evolve_config, "fp32_matvecs", 0))
        sz.append(sigma_z(mps))
        if steps > 50 and (k + 1) % 25 == 0:
            log(f"{precision}: {k + 1} of {steps} timed steps, {seconds:.1f} s")
    prof = None
    if profile:
        prof = native.prof_collect()
        for name, p in sorted(prof.items(), key=lambda kv: -kv[1]["flop"]):
            log(f"{precision}: gemm {name}: executed {p['executed']} queued {p['queued']} "
                f"flop {p['flop']:.4g}; timed {p['timed']} in {p['timed_ms']:.1f} ms")
    env = mps._env_cache[2] if mps._env_cache else None
    redone = Mps.krylov_redo_count - redo1
    log(f"{precision}: steps redone after an exhausted Krylov budget: warmup {redo1 - redo0}, "
        f"timed {redone}")
    return {"seconds": seconds, "steps_per_s": steps / seconds, "sz": sz, "prof": prof,
            "matvecs_per_step": matvecs, "fp32_matvecs_per_step": low, "final": mps, "environ": env, "redone": redone}


def torch_baseline(mps0, mpo, dt, steps, ref_sz):
    """The same time steps by the plain-torch FP64 path (the reference's GPU
    algorithm in torch operations: tensordot on the device, rocSOLVER QR,
    Lanczos scalars and the tridiagonal eigenproblem on the host, every
    environment rebuilt per step as the reference does) from the same
    initial state: one untimed step, then ``steps`` timed ones; <sigma_z>
    after each is compared with the HIP FP64 run's."""
    import numpy as np
    import torch
    from renormalizer_amd.mps.torch_sweep import TorchSweepState
    dev = torch.device("cuda", 0)
    ts = TorchSweepState.from_mps(mps0, dev)
    mpo_t = [torch.from_numpy(np.asarray(w)).to(device=dev, dtype=torch.complex128) for w in mpo]

    def sz():
        r = ts.calc_1site_rdm(0)
        return float((r[0, 0] - r[1, 1]).real)

    out_sz = [sz()]
    ts.evolve(mpo_t, dt)
    out_sz.append(sz())
    seconds = 0.0
    matvecs = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        counts = ts.evolve(mpo_t, dt)
        torch.cuda.synchronize()
        seconds += time.perf_counter() - t0
        matvecs.append(sum(counts))
        out_sz.append(sz())
    n = min(len(out_sz), len(ref_sz))
    diff = float(np.max(np.abs(np.array(out_sz[:n]) - np.array(ref_sz[:n]))))
    log(f"torch fp64 baseline: {steps} timed steps in {seconds:.2f} s, matvecs/step {matvecs}")
    return {"sweeps_per_s": round(steps / seconds, 5), "ms_per_step": round(1e3 * seconds / steps, 1),
            "steps": steps, "warmup": 1, "matvecs_per_step": matvecs,
            "max_abs_sigma_z_diff_vs_hip_fp64": diff,
            "what": "plain torch FP64 (rocBLAS zgemm via tensordot, rocSOLVER QR, host Lanczos "
                    "scalars; the reference's CuPy path restated in torch) on the same GPU and start"}


TRAFFIC_PROFILES = [os.path.join(ROOT, "profiles", f"tdvp_c4_m256_{mode}_hbm_traffic.md")
                    for mode in ("fp64", "mixed")]


def measured_traffic(kernel):
    """HBM bytes per executed launch of ``kernel`` from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summaries (lines
    "traffic_bytes_per_launch | <kernel> | <bytes>"), or None."""
    for path in TRAFFIC_PROFILES:
        try:
            with open(path) as f:
                for line in f:
                    parts = [p.strip() for p in line.strip().strip("|").split("|")]
                    if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                        return float(parts[2])
        except OSError:
            pass
    return None


def roofline(prof, seconds, steps):
    """Dominant GEMM instantiation (most executed FLOP): achieved TFLOP/s =
    algorithmic FLOP per event-timed executed launch / its average duration.
    The duration is the event-timed average less the event bracket's own
    time, measured in the same run on sampled launches that were cancelled
    on the device (they return at once): the kernel's execution time, the
    figure a profiler's kernel trace reports.  The uncorrected event figures
    are kept beside it (``*_event``)."""
    name, p = max(prof.items(), key=lambda kv: kv[1]["flop"])
    peak = PEAK_TFLOPS["double" if "<double" in name else "float"]
    event_us = p["timed_ms"] * 1e3 / p["timed"]
    over_us = p["timed_cancelled_ms"] * 1e3 / p["timed_cancelled"] if p["timed_cancelled"] else 0.0
    flop_launch = p["timed_flop"] / p["timed"]
    achieved = flop_launch / ((event_us - over_us) * 1e-6) / 1e12
    achieved_event = flop_launch / (event_us * 1e-6) / 1e12
    total_flop = sum(v["flop"] for v in prof.values())
    # GEMM GPU time of all executed launches, scaled from each sample
    total_ms = sum(v["timed_ms"] * v["executed"] / v["timed"] for v in prof.values() if v["timed"])
    return {
        "bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4), "traffic": measured_traffic(name), "kernel": name,
        "launches": p["executed"], "launches_queued": p["queued"], "launches_timed": p["timed"],
        "avg_launch_us": round(event_us - over_us, 2),
        "avg_launch_us_event": round(event_us, 2),
        "achieved_event": round(achieved_event, 2),
        "frac_event": round(achieved_event / peak, 4),
        # event time of sampled launches cancelled on the device: the
        # bracket's own share (queue latency) of avg_launch_us_event
        "event_overhead_us": round(over_us, 2) if p["timed_cancelled"] else None,
        "flop_per_launch": p["timed_flop"] / p["timed"],
        "gemm_flop_per_step": total_flop / steps,
        "gemm_time_fraction": round(total_ms * 1e-3 / seconds, 3),
        "step_tflops": round(total_flop / seconds / 1e12, 2),
    }


def host_info():
    """CPU model, usable cores, BLAS threads, numpy/scipy versions and BLAS
    library of this host."""
    import numpy as np
    import scipy
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = "unknown"
    try:
        cfg = np.show_config(mode="dicts")
        b = cfg.get("Build Dependencies", {}).get("blas", {})
        blas = f"{b.get('name', '?')} {b.get('version', '')}".strip()
    except Exception:
        pass
    return {"cpu_model": model, "affinity_cores": _CORES,
            "reno_num_threads": int(os.environ["RENO_NUM_THREADS"]),
            "numpy": np.__version__, "scipy": scipy.__version__, "blas": blas}


def qr_flop(m, n):
    """Householder QR of an m x n complex matrix (m >= n) plus forming the
    m x n Q: 2 x 4(2 m n^2 - 2 n^3 / 3) real FLOP."""
    m, n = max(m, n), min(m, n)
    return 2 * 4 * (2 * m * n * n - 2 * n ** 3 / 3)


def site_update_flop(ml, d, mr, dl, dr, k1, k0, mb=None):
    """Algorithmic FLOP of one right-to-left site update: k1 one-site
    matvecs, the RQ of the (ml) x (d mr) site matrix with its Q, the
    environment update, k0 bond matvecs (bond mb x ml) and the bond absorb."""
    mb = ml if mb is None else mb
    f_hop1 = 8 * (ml * ml * dl * d * mr + ml * d * dr * mr * mr) + 4 * ml * mr * dl * dr * d * d
    f_hop0 = 8 * dl * mb * ml * (mb + ml)
    f_absorb = 8 * mb * ml * ml * d
    return k1 * f_hop1 + qr_flop(d * mr, ml) + f_hop1 + k0 * f_hop0 + f_absorb


def cpu_baseline(mps_dev, env, mpo, dt, sample_reps, flop_per_step):
    """The numpy oracle timed on this host on a bounded sample of the
    workload: ``sample_reps`` consecutive right-to-left site updates of the
    evolved state's full-size middle sites (Lanczos exponential of the
    one-site operator, RQ, environment update, bond exponential, bond
    absorb, exactly as one step does them), converted to time steps per
    second by the algorithmic FLOP of the sample (QR included) against the
    FLOP of one device step."""
    import numpy as np

    from oracle import tdvp as orc
    n = len(mps_dev)
    sites = mps_dev.numpy_sites()
    w = [np.asarray(x) for x in mpo]
    site = n // 2 + sample_reps // 2
    R = env.read("R", site + 1).cpu().numpy().astype(np.complex128)
    t = 0.0
    flop = 0.0
    for rep in range(sample_reps):
        i = site - rep
        L = env.read("L", i - 1).cpu().numpy().astype(np.complex128)
        t1 = time.perf_counter()
        c = sites[i]
        Ml, d, Mr = c.shape
        Dl, Dr = w[i].shape[0], w[i].shape[3]
        v, k = orc.expm_krylov(lambda y: orc.hop1(L, w[i], R, y.reshape(c.shape)).ravel(),
                               -1j * dt / 2, c.ravel())
        v = v.reshape(c.shape)
        u, _, vv, _ = orc.svd_qn(v, np.zeros((Ml, 1), dtype=int), np.zeros((d * Mr, 1), dtype=int),
                                 np.zeros(1, dtype=int), QR=True, system="R", full_matrices=False)
        vt = vv.T.reshape(-1, d, Mr)
        R = orc.contract_one_site(R, vt, w[i], "R")
        u_t, k0 = orc.expm_krylov(lambda y: orc.hop0(L, R, y.reshape(u.shape)).ravel(),
                                  1j * dt / 2, u.ravel())
        sites[i - 1] = np.tensordot(sites[i - 1], u_t.reshape(u.shape), axes=(-1, 0))
        t += time.perf_counter() - t1
        flop += site_update_flop(Ml, d, Mr, Dl, Dr, k, k0, mb=u.shape[1])
    rate = flop / t
    info = host_info()
    return {
        "value": round(rate / flop_per_step, 5), "unit": "sweeps/s",
        "cores": info["reno_num_threads"], "kind": "port",
        "sample": (f"{sample_reps} consecutive right-to-left full-size site updates (256x10x256: "
                   f"Lanczos expm, RQ, environment update, bond expm, bond absorb) of the evolved "
                   f"M=256 state by the numpy oracle in {t:.1f} s = {rate / 1e9:.1f} GFLOP/s, "
                   f"scaled by the algorithmic FLOP of one device step ({flop_per_step / 1e12:.2f} "
                   f"TFLOP, QR included)"),
        "host": info,
    }


def step_flop(counts, shapes, mpo_shapes):
    """Algorithmic FLOP of one time step, counted like site_update_flop,
    from the executed matvec counts of its Krylov calls (keys (half, site,
    1) one-site, (half, site, 0) bond; half 0 runs right to left) and the
    site shapes."""
    total = 0.0
    for (half, i, kind), k in counts.items():
        ml, d, mr = shapes[i]
        dl, dr = mpo_shapes[i][0], mpo_shapes[i][3]
        f_hop1 = 8 * (ml * ml * dl * d * mr + ml * d * dr * mr * mr) + 4 * ml * mr * dl * dr * d * d
        if kind == 1:
            total += k * f_hop1
            continue
        if half == 0:   # RQ of (ml) x (d mr); bond (ml x ml) between L[i-1] and the new R[i]
            mb, db, qr = ml, dl, qr_flop(d * mr, ml)
        else:           # QR of (ml d) x (mr); bond (mr x mr)
            mb, db, qr = mr, dr, qr_flop(ml * d, mr)
        total += k * 8 * db * mb * mb * 2 * mb + qr + f_hop1 + 8 * mb * mb * mb * d
    return total


def cpu_trajectory(workload, steps):
    """The numpy oracle alone on one of the workloads, in this process (its
    BLAS thread count was fixed before numpy was imported): median wall
    time per time step after one warm-up step, initial state excluded."""
    import numpy as np

    from oracle import tdvp as orc
    from tests import helpers
    if workload == "c1":
        model = helpers.sbm_model(16, 8, alpha=0.5, omega_c=5.0)
        mps, mpo = helpers.sbm_initial(model, 16)
        dt = 0.05
    elif workload == "c2":
        model = helpers.holstein_model(16, 10, j=-1.0, g=1.0)
        mps, mpo = helpers.holstein_initial(model, 64)
        dt = 0.02
    elif workload == "c4":
        _, mpo, mps = build_workload(64, 10, 256, on_device=False)
        dt = 0.025
    else:
        raise SystemExit(f"unknown workload {workload}")
    st = helpers.to_state(mps)
    w = helpers.mpo_arrays(mpo)
    times = []
    for k in range(steps + 1):
        t0 = time.perf_counter()
        st, _ = orc.tdvp_ps(st, w, dt)
        if k:
            times.append(time.perf_counter() - t0)
        log(f"cpu {workload}: step {k} {time.perf_counter() - t0:.2f} s")
    med = float(np.median(times))
    out = {"workload": workload, "steps_timed": steps, "median_step_s": round(med, 4),
           "sweeps_per_s": round(1.0 / med, 5), "host": host_info()}
    print(json.dumps(out), flush=True)


def cpu_sweep(cores):
    """cpu_trajectory of C1 (100 steps) and C2 (5 steps) at 1, 4 and all
    threads and of C4 (1 step) at all threads, each in a fresh process."""
    import subprocess
    runs = []
    for threads in sorted({1, min(4, cores), cores}):
        for wl, n in (("c1", 100), ("c2", 5)):
            runs.append((wl, n, threads))
    runs.append(("c4", 1, cores))
    for wl, n, threads in runs:
        env = dict(os.environ, RENO_NUM_THREADS=str(threads))
        for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
            env[k] = str(threads)
        log(f"cpu sweep: {wl}, {n} steps, {threads} threads")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-trajectory", wl,
                            "--steps", str(n)], env=env, stdout=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise SystemExit(f"cpu trajectory {wl} at {threads} threads failed ({r.returncode})")
        print(r.stdout.strip().splitlines()[-1], flush=True)


def _c3_instance(args):
    """One instance of the batched disorder workload (host, in a worker
    process): Holstein chain of nmol molecules (scheme 2, 8 levels, omega 1,
    g 0.7, J -1), site energies from N(0, 0.5), Franck-Condon electron in the
    middle, expanded to bond dimension m and canonicalised."""
    nmol, m, elocal = args
    from threadpoolctl import threadpool_limits

    from tests import helpers
    with threadpool_limits(1):
        model = helpers.holstein_model(nmol, 8, j=-1.0, g=0.7, elocal=list(elocal))
        mps, mpo = helpers.holstein_initial(model, m)
    return model, mps, mpo


def c3_site_flop(local_steps, shapes, w_shapes, to_right):
    """Algorithmic FLOP of one grouped time step of one state, from its
    local_steps (per half-sweep and site: one-site matvecs, then the bond
    matvecs) counted like site_update_flop."""
    n = len(shapes)
    first = list(range(n)) if to_right else list(range(n - 1, -1, -1))
    it = iter(local_steps)
    total = 0.0
    for order, right in ((first, to_right), (first[::-1], not to_right)):
        for p, i in enumerate(order):
            ml, d, mr = shapes[i]
            dl, dr = w_shapes[i][0], w_shapes[i][3]
            f_hop1 = 8 * (ml * ml * dl * d * mr + ml * d * dr * mr * mr) + 4 * ml * mr * dl * dr * d * d
            total += next(it) * f_hop1
            if p == n - 1:
                continue
            mb, db = (mr, dr) if right else (ml, dl)
            k0 = next(it)
            qr = qr_flop(ml * d, mr) if right else qr_flop(d * mr, ml)
            total += k0 * 8 * db * mb * mb * 2 * mb + qr + f_hop1 + 8 * mb * mb * mb * d
    return total


def run_c3(instances, steps, warmup, seed=2026, dt=0.02):
    """The batched disorder workload: every state's site update at one sweep
    position in one launch (BatchEvolution), against the same states evolved
    one at a time on four concurrent streams (evolve_batch)."""
    import multiprocessing as mp_

    import numpy as np
    rng = np.random.default_rng(seed)
    args = []
    for i in range(instances):
        nmol = int(rng.integers(8, 33))
        args.append((nmol, 16 * 2 ** (i % 3), rng.normal(0.0, 0.5, nmol)))
    t0 = time.perf_counter()
    ctx = mp_.get_context("fork")
    with ctx.Pool(min(_CORES, 16)) as pool:
        built = pool.map(_c3_instance, args, chunksize=1)
    log(f"c3: {instances} initial states on the host in {time.perf_counter() - t0:.1f} s")
    import torch
    from renormalizer_amd import EvolveConfig, EvolveMethod
    from renormalizer_amd.mps.batch import BatchEvolution, evolve_batch
    torch.cuda.set_device(0)
    states = []
    for _, mps, _ in built:
        mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        mps.to_device()
        states.append(mps)
    mpos = [o for _, _, o in built]
    t0 = time.perf_counter()
    be = BatchEvolution(states, mpos, dt, warmup=max(1, warmup), nstreams=4)
    torch.cuda.synchronize()
    log(f"c3: warm-up (one-state path, 4 streams) and grouped set-up {time.perf_counter() - t0:.1f} s")
    # reference rate: one step of every state through the one-state path
    ref = [m.copy() for m in be.states]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = evolve_batch(ref, mpos, dt, nstreams=4)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    del ref
    be.step()
    # every later step: one replay of a HIP graph holding its launches
    be.capture()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    flop = 0.0
    pops = []
    for _ in range(steps):
        be.step()
        for m, ls, o in zip(be.states, be.local_steps, mpos):
            flop += c3_site_flop(ls, [tuple(t.shape) for t in m._mp], [np.asarray(w).shape for w in o],
                                 m.to_right)
    ev1.record()
    torch.cuda.synchronize()
    t_grouped = (time.perf_counter() - t0) / steps
    gpu_ms = ev0.elapsed_time(ev1) / steps
    launches = 2 * be.npos + 1
    pops = float(np.mean([m.e_occupations.mean() for m in be.states[:8]]))
    out = {
        "metric": "batched disorder: instance time steps per second",
        "value": round(instances / t_grouped, 2), "unit": "instance-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * t_grouped, 2),
        "higher_is_better": True, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"C3: {instances} Holstein chains of 8..32 molecules (scheme 2, 8 levels, "
                               f"omega 1, g 0.7, J -1, site energies N(0, 0.5), seed {seed}), "
                               f"M = 16/32/64, dt {dt}, TDVP-PS one-site, FP64",
                   "launches_per_step": launches, "hip_graph": True},
        "one_state_path_4_streams_ms_per_step": round(1e3 * t_single, 1),
        "speedup_vs_one_state_path": round(t_single / t_grouped, 2),
        "roofline": {"bound": "mfma", "achieved": round(flop / steps / (gpu_ms * 1e-3) / 1e12, 3),
                     "peak": PEAK_TFLOPS["double"], "unit": "TFLOP/s",
                     "frac": round(flop / steps / (gpu_ms * 1e-3) / 1e12 / PEAK_TFLOPS["double"], 4),
                     "traffic": None, "kernel": "k_batch_site(reno_batch_site const*, double)",
                     "flop_per_step": flop / steps, "gpu_ms_per_step": round(gpu_ms, 2),
                     "avg_launch_us": round(1e3 * gpu_ms / launches, 1)},
        "mean_population_check": pops,
    }
    print(json.dumps(out), flush=True)


def copy_calibration(mib=512, reps=3):
    """A plain device-to-device copy of known size, for scaling the PMC
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    torch.cuda.set_device(0)
    a = torch.empty(mib * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": a.numel() * 8, "dispatches": reps}), flush=True)


def _rocprof_csv(directory, suffix):
    import csv
    import glob
    files = glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True)
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def traffic_probe(reps=20):
    """The matvec stages at the C4 full-site shape (M = 256, d = 10, D = 3)
    in both precisions, tuned as the sweep tunes them, each matvec followed
    by a device synchronisation (counter runs: rocprofv3 --pmc)."""
    import torch
    from renormalizer_amd import native
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    gen = torch.Generator().manual_seed(0)
    M, d, D = 256, 10, 3

    def cplx(*shape):
        return torch.complex(torch.randn(*shape, generator=gen, dtype=torch.float64),
                             torch.randn(*shape, generator=gen, dtype=torch.float64))
    for prec in (native.PREC_FP64, native.PREC_MIXED):
        cd = torch.complex128 if prec == native.PREC_FP64 else torch.complex64
        rd = torch.float64 if prec == native.PREC_FP64 else torch.float32
        L, R = cplx(M, D, M).to(cd).to(dev), cplx(M, D, M).to(cd).to(dev)
        W = torch.randn(D, d, d, D, generator=gen, dtype=torch.float64).to(rd).to(dev)
        c1, c0 = cplx(M, d, M).to(dev), cplx(M, M).to(dev)
        native.autotune(L, W, R, c1, prec)
        native.autotune(L, None, R, c0, prec)
        for _ in range(reps):
            native.hop_apply(L, W, R, c1, prec)
            torch.cuda.synchronize()
            native.hop_apply(L, None, R, c0, prec)
            torch.cuda.synchronize()
    log(f"traffic probe: {reps} one-site and bond matvecs per precision at M = {M}, d = {d}, D = {D}")


def qr_probe(reps=10, rows=2560, cols=256):
    """Blocked QR of one random complex rows x cols block (the full-site
    shape) through reno_block_qr, ``reps`` times (kernel-trace and counter
    runs of the Householder kernels)."""
    import numpy as np
    import torch
    from renormalizer_amd import native
    torch.cuda.set_device(0)
    rng = np.random.default_rng(5)
    t = torch.from_numpy(rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).to("cuda:0")
    blocks = [(np.arange(rows), np.arange(cols))]
    for _ in range(reps):
        native.block_qr("L", t, blocks)
    torch.cuda.synchronize()
    log(f"qr probe: {reps} QRs of a {rows} x {cols} block")


def summarize_counters(directory, top=12):
    """Per-kernel table of a rocprofv3 --pmc run: dispatches, dispatches
    whose counter reads under 1 KiB (launches cancelled on the device by the
    Krylov stop flag write nothing), counter sum and a log2(KiB) histogram."""
    import collections
    import math
    agg = collections.defaultdict(list)
    grid = collections.defaultdict(list)
    counter = None
    for r in _rocprof_csv(directory, "counter_collection.csv"):
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        grid[(r["Kernel_Name"], r.get("Grid_Size", "?"))].append(float(r["Counter_Value"]))
        counter = r.get("Counter_Name", counter)
    print(f"counter {counter}")
    print("| kernel | dispatches | under 1 KiB | sum KiB | log2(KiB) histogram |")
    print("|---|---|---|---|---|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        hist = collections.Counter(int(math.log2(x)) if x >= 1 else -1 for x in v)
        h = " ".join(f"{k}:{hist[k]}" for k in sorted(hist))
        print(f"| {name} | {len(v)} | {sum(1 for x in v if x < 1)} | {sum(v):.1f} | {h} |")
    print("\nby grid size (work-items):\n")
    print("| kernel | grid | dispatches | mean KiB | min KiB | max KiB |")
    print("|---|---|---|---|---|---|")
    for (name, gs), v in sorted(grid.items(), key=lambda kv: -sum(kv[1]))[:3 * top]:
        print(f"| {name} | {gs} | {len(v)} | {sum(v) / len(v):.1f} | {min(v):.1f} | {max(v):.1f} |")


def summarize_mfma(directory, top=12):
    """Per kernel and grid size of a rocprofv3 --pmc run of MfmaUtil with
    SQ_INSTS_VALU_MFMA_MOPS_F32 and _F64: over the dispatches that executed
    matrix-core work, the mean MFMA utilisation (percent of SIMD cycles the
    matrix cores are busy while the GPU is active, the profiler's derived
    counter) and the MFMA FLOP per dispatch (MOPS x 512)."""
    import collections
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    meta = {}
    for r in _rocprof_csv(directory, "counter_collection.csv"):
        d = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[d][r["Counter_Name"]] += float(r["Counter_Value"])
        meta[d] = (r["Kernel_Name"], r.get("Grid_Size", "?"))
    groups = collections.defaultdict(list)
    for d, c in per.items():
        flop = 512.0 * (c.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0) + c.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0))
        if flop > 0:
            groups[meta[d]].append((c.get("MfmaUtil", float("nan")), flop))
    print("| kernel | grid (work-items) | dispatches | mean MfmaUtil % | min % | max % | MFMA GFLOP per dispatch |")
    print("|---|---|---|---|---|---|---|")
    for (name, gs), v in sorted(groups.items(), key=lambda kv: -sum(f for _, f in kv[1]))[:top]:
        u = [x for x, _ in v]
        print(f"| {name} | {gs} | {len(v)} | {sum(u) / len(u):.1f} | {min(u):.1f} | {max(u):.1f} | "
              f"{sum(f for _, f in v) / len(v) / 1e9:.3f} |")


def summarize_gaps(directory, last_ms=None):
    """Busy and idle time of the device over a kernel trace: the union of
    the kernel intervals against the span from the first start to the last
    end (restricted to the last ``last_ms`` milliseconds of the trace when
    given, e.g. the timed steps), and the histogram of the gaps between
    consecutive kernels."""
    import collections
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                for r in _rocprof_csv(directory, "kernel_trace.csv"))
    if last_ms is not None:
        t_end = max(e for _, e in iv)
        iv = [(s, e) for s, e in iv if s >= t_end - int(last_ms * 1e6)]
    busy, cur_s, cur_e = 0, None, None
    gaps = []
    for s, e in iv:
        if cur_e is None:
            cur_s, cur_e = s, e
        elif s <= cur_e:
            cur_e = max(cur_e, e)
        else:
            busy += cur_e - cur_s
            gaps.append(s - cur_e)
            cur_s, cur_e = s, e
    busy += cur_e - cur_s
    span = max(e for _, e in iv) - iv[0][0]
    hist = collections.Counter(min(int(g // 1000), 20) for g in gaps)
    print(f"kernels {len(iv)}, span {span / 1e6:.1f} ms, busy {busy / 1e6:.1f} ms, "
          f"idle {(span - busy) / 1e6:.1f} ms ({100 * (span - busy) / span:.1f} %), gaps {len(gaps)}, "
          f"mean gap {sum(gaps) / max(len(gaps), 1) / 1e3:.2f} us")
    print("gap histogram (us: count): " + ", ".join(
        f"{k}{'+' if k == 20 else ''}: {hist[k]}" for k in sorted(hist)))


def summarize_trace(directory, top=12):
    """Per-kernel table of a rocprofv3 --kernel-trace run: dispatches, total
    and average duration, and the number of dispatches shorter than 2 us
    (the launches cancelled on the device end within about a microsecond)."""
    import collections
    dur = collections.defaultdict(list)
    for r in _rocprof_csv(directory, "kernel_trace.csv"):
        dur[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    print("| kernel | dispatches | total ms | avg us | under 2 us |")
    print("|---|---|---|---|---|")
    for name, v in sorted(dur.items(), key=lambda kv: -sum(kv[1]))[:top]:
        print(f"| {name} | {len(v)} | {sum(v) / 1e6:.1f} | {sum(v) / len(v) / 1e3:.2f} | "
              f"{sum(1 for x in v if x < 2000)} |")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dt", type=float, default=0.025)
    ap.add_argument("--bond-dim", type=int, default=256)
    ap.add_argument("--modes", type=int, default=64)
    ap.add_argument("--cpu-reps", type=int, default=8,
                    help="site updates in the CPU baseline sample (about 1.2 s each at M = 256 on 16 cores)")
    ap.add_argument("--mixed-tau", type=float, default=None,
                    help="switch threshold of the mixed mode (EvolveConfig.mixed_tau)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--torch-steps", type=int, default=2,
                    help="timed steps of the plain-torch FP64 baseline (0: skip it)")
    ap.add_argument("--host-start", action=argparse.BooleanOptionalAction, default=True,
                    help="also time both modes from the start built by the host (numpy/LAPACK) "
                         "expansion, the reference's construction (host_start in the output)")
    ap.add_argument("--state-cache", metavar="NPZ", default=None,
                    help="load the initial state from this file if it exists, else build and "
                         "write it there (profiling runs)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="time the tile / split-K candidates of each matvec shape during the "
                         "warm-up (EvolveConfig.autotune) and use the fastest")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="run one precision mode only (profiling runs)")
    ap.add_argument("--workload", choices=["c4", "c3"], default="c4",
                    help="c4: the metric workload (default); c3: the batched disorder workload")
    ap.add_argument("--instances", type=int, default=256, help="c3: number of states")
    ap.add_argument("--cpu-trajectory", choices=["c1", "c2", "c4"],
                    help="time the numpy oracle alone on one workload (no GPU)")
    ap.add_argument("--cpu-sweep", action="store_true",
                    help="cpu-trajectory of C1/C2 at 1, 4 and all threads and C4 at all threads")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the C4-shape matvec stages, synchronously (PMC counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 512 MiB device copy (PMC counter calibration)")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="print the per-kernel table of a rocprofv3 --pmc output directory")
    ap.add_argument("--summarize-gaps", metavar="DIR",
                    help="device busy / idle time and kernel gaps of a rocprofv3 --kernel-trace directory")
    ap.add_argument("--gaps-last-ms", type=float, default=None,
                    help="with --summarize-gaps: only the last this many milliseconds of the trace")
    ap.add_argument("--qr-probe", action="store_true",
                    help="ten blocked QRs of a 2560 x 256 block (profiler runs of the QR kernels) and exit")
    ap.add_argument("--summarize-mfma", metavar="DIR",
                    help="print the per-kernel MFMA utilisation table of a rocprofv3 --pmc MfmaUtil run and exit")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace output directory")
    a = ap.parse_args()
    if a.summarize_counters:
        summarize_counters(a.summarize_counters)
        return
    if a.summarize_gaps:
        summarize_gaps(a.summarize_gaps, a.gaps_last_ms)
        return
    if a.summarize_mfma:
        summarize_mfma(a.summarize_mfma)
        return
    if a.qr_probe:
        qr_probe()
        return
    if a.summarize_trace:
        summarize_trace(a.summarize_trace)
        return
    if a.copy_calibration:
        copy_calibration()
        return
    if a.traffic_probe:
        traffic_probe()
        return
    if a.cpu_trajectory:
        cpu_trajectory(a.cpu_trajectory, a.steps)
        return
    if a.cpu_sweep:
        cpu_sweep(_CORES)
        return
    if a.workload == "c3":
        run_c3(a.instances, a.steps, a.warmup)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the TDVP-PS sweep of a single MPS "
                                   "is sequential along the chain and is measured on one device",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    model, mpo, mps0 = build_workload(a.modes, 10, a.bond_dim, cache=a.state_cache)
    torch.cuda.synchronize()
    init_s = time.perf_counter() - t0
    log(f"initial state (device expansion) {init_s:.1f} s, bond dims max "
        f"{max(mps0.bond_dims)}, sites {len(mps0)}")

    res = {}
    for prec in ("fp64", "mixed"):
        if a.only and prec != a.only:
            continue
        log(f"running {prec}: {a.warmup} warmup + {a.steps} timed steps")
        res[prec] = run_mode(mps0, mpo, prec, a.dt, a.warmup, a.steps, profile=True, tau=a.mixed_tau,
                             autotune=a.autotune)
        log(f"{prec}: {res[prec]['steps_per_s']:.3f} steps/s, matvecs/step "
            f"{res[prec]['matvecs_per_step']}")
    if a.autotune:
        from renormalizer_amd import native
        for t in native.tuned_records:
            log(f"autotune: prec {t['prec']} M {t['M']} N {t['N']} K {t['K']} batch {t['batch']}: "
                f"tile {t['tile']} split {t['nsplit']} {t['us_tuned']:.1f} us (planner: tile "
                f"{t['planned_tile']} split {t['planned_nsplit']} {t['us_planned']:.1f} us)")
    err = None
    if "fp64" in res and "mixed" in res:
        import numpy as np
        err = float(np.max(np.abs(np.array(res["mixed"]["sz"]) - np.array(res["fp64"]["sz"]))))
    if "fp64" not in res:
        main_mode = "mixed"
    elif "mixed" in res and err is not None and err <= ACCURACY_BOUND \
            and res["mixed"]["steps_per_s"] > res["fp64"]["steps_per_s"]:
        main_mode = "mixed"
    else:
        main_mode = "fp64"
    r = res[main_mode]
    rl = roofline(r["prof"], r["seconds"], a.steps)
    out = {
        "metric": "TDVP-PS sweeps/s at M=256, mixed vs FP64 GPU at <=1e-6 population error",
        "value": round(r["steps_per_s"], 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * r["seconds"] / a.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64-mfma+f64" if main_mode == "mixed" else "c128",
        "data": "synthetic (model parameters from the workload definition; no checkpoints)",
        "config": {"workload": f"spin-boson 1 spin + {a.modes} modes x 10 levels, Ohmic alpha=0.3 "
                               f"omega_c=10 Wang1, M={a.bond_dim}, dt={a.dt}, TDVP-PS one-site",
                   "mode": main_mode, "sweep": "1 step = 2 half-sweeps",
                   "autotune": bool(a.autotune)},
        "roofline": rl,
        "krylov_steps_redone": r["redone"],
        "initial_state_s": round(init_s, 2),
    }
    if err is not None:
        out["fp64_sweeps_per_s"] = round(res["fp64"]["steps_per_s"], 4)
        out["mixed_sweeps_per_s"] = round(res["mixed"]["steps_per_s"], 4)
        out["mixed_over_fp64"] = round(res["mixed"]["steps_per_s"] / res["fp64"]["steps_per_s"], 3)
        out["max_abs_sigma_z_err_mixed_vs_fp64"] = err
        out["mixed_within_bound"] = err <= ACCURACY_BOUND
        out["steps_compared"] = len(res["fp64"]["sz"]) - 1
        from renormalizer_amd import EvolveConfig
        out["mixed_tau"] = a.mixed_tau if a.mixed_tau is not None else EvolveConfig().mixed_tau
        mv64, mvmx = sum(res["fp64"]["matvecs_per_step"]), sum(res["mixed"]["matvecs_per_step"])
        out["matvecs_fp64"] = mv64
        out["matvecs_mixed"] = mvmx
        out["mixed_fp32_matvec_share"] = round(sum(res["mixed"]["fp32_matvecs_per_step"]) / mvmx, 3)
        other = "mixed" if main_mode == "fp64" else "fp64"
        out[other + "_roofline"] = roofline(res[other]["prof"], res[other]["seconds"], a.steps)
    if a.host_start and a.state_cache is None:
        # The admixture directions of the 1e-16 expansion are fixed by rounding:
        # the host's LAPACK SVD of the graded matrix draws them from rounding
        # noise, the device's QR + Gram eigensolver keeps them inside the span
        # of H^k psi.  Same physics (<sigma_z> within the hosts' own gesdd /
        # gesvd spread), different Lanczos work per step: both are reported.
        t1 = time.perf_counter()
        _, _, mps_h = build_workload(a.modes, 10, a.bond_dim, on_device=False)
        hs = {"build_s": round(time.perf_counter() - t1, 1)}
        log(f"host-built start in {hs['build_s']} s")
        for prec in res:
            rh = run_mode(mps_h, mpo, prec, a.dt, a.warmup, a.steps, profile=False, tau=a.mixed_tau,
                          autotune=a.autotune)
            hs[prec + "_sweeps_per_s"] = round(rh["steps_per_s"], 4)
            hs[prec + "_matvecs_per_step"] = rh["matvecs_per_step"]
            hs[prec + "_max_abs_sigma_z_diff_vs_device_start"] = float(
                np.max(np.abs(np.array(rh["sz"]) - np.array(res[prec]["sz"]))))
            log(f"host start {prec}: {rh['steps_per_s']:.4f} sweeps/s, matvecs/step {rh['matvecs_per_step']}")
        for prec in res:
            hs["device_start_" + prec + "_matvecs_per_step"] = res[prec]["matvecs_per_step"]
        hs["what"] = ("the same steps from the start expanded on the host (numpy/LAPACK SVD, the reference's "
                      "construction); value above is from the device-expanded start")
        out["host_start"] = hs
    if a.torch_steps > 0 and "fp64" in res:
        tb = torch_baseline(mps0, mpo, a.dt, a.torch_steps, res["fp64"]["sz"])
        out["torch_fp64_baseline"] = tb
        out["speedup_vs_torch_fp64"] = round(out["value"] / tb["sweeps_per_s"], 2)
    if not a.no_cpu:
        ref = res.get("fp64", r)
        log("cpu baseline sample")
        fin = ref["final"]
        counts = fin._env_cache[3][1]
        fstep = step_flop(counts, [tuple(m.shape) for m in fin._mp], [np.asarray(w).shape for w in mpo])
        out["algorithmic_flop_per_step"] = fstep
        out["cpu_baseline"] = cpu_baseline(fin, ref["environ"], mpo, a.dt, a.cpu_reps, fstep)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
