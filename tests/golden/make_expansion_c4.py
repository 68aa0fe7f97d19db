"""Writes expansion_c4.json: the host-path (numpy/LAPACK) spin-boson start
of the benchmark workload (spin + 64 Wang-1 modes x 10 levels, alpha 0.3,
omega_c 10, bond dimension 256, expand_bond_dimension with admixture 1e-16)
and <sigma_z> over 5 FP64 TDVP-PS steps of dt = 0.025 from it, built twice:
with the SVD of the truncation sweeps by LAPACK's gesdd (the default) and by
gesvd.  The two starts differ only by rounding in the 1e-16 admixture
directions; the spread of their trajectories is the conditioning against
which the device-built start is compared.  Run on a GPU (the steps use the
device path; the builds take about a minute each on the host):
    python tests/golden/make_expansion_c4.py"""
import json
import os
import sys
import time

import numpy as np
import scipy.linalg

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))

import renormalizer_amd.mps.svd_qn as svd_qn  # noqa: E402
from renormalizer_amd import (CompressConfig, CompressCriteria, EvolveConfig, EvolveMethod, Mpo, Mps,  # noqa: E402
                              Phonon, Quantity, SpinBosonModel)
from renormalizer_amd.sbm import SpectralDensityFunction  # noqa: E402

OUT = os.path.join(os.path.dirname(__file__), "expansion_c4.json")
DT, STEPS, M = 0.025, 5, 256


def model_and_mpo():
    sdf = SpectralDensityFunction(0.3, 10.0)
    w, c2 = sdf.Wang1(64)
    omegas, dis = sdf.post_process(w, c2)
    model = SpinBosonModel(Quantity(0.0), Quantity(1.0),
                           [Phonon.simple_phonon(o, d, 10) for o, d in zip(omegas, dis)])
    return model, Mpo(model)


def sigma_z(mps):
    r = mps.calc_1site_rdm(0)[0]
    return float((r[0, 0] - r[1, 1]).real)


def main():
    model, mpo = model_and_mpo()
    out = {"dt": DT, "steps": STEPS, "bond_dim": M}
    for driver in ("gesdd", "gesvd"):
        svd_qn.optimized_svd = lambda a, full_matrices, d=driver: scipy.linalg.svd(
            a, full_matrices=full_matrices, lapack_driver=d)
        t = time.time()
        mps = Mps.ground_state(model, False)
        mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=M)
        mps = mps.expand_bond_dimension(mpo, coef=1e-16, include_ex=False)
        print(f"{driver}: host build {time.time() - t:.1f} s", flush=True)
        out["bond_dims"] = [int(b) for b in mps.bond_dims]
        mps.to_device()
        mps.evolve_config = EvolveConfig(EvolveMethod.tdvp_ps)
        zs = []
        for _ in range(STEPS):
            mps = mps.evolve(mpo, DT)
            zs.append(sigma_z(mps))
        out[f"sigma_z_{driver}"] = zs
        print(driver, zs, flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
