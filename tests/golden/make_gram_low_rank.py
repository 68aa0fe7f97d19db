"""Writes gram_low_rank_81.npy: the first 81 x 81 Gram matrix T T^H that
the device SVD branch (svd_qn_device_svd) hands to reno_block_eigh while
expanding the spin-boson start (spin + 64 Wang-1 modes x 10 levels, alpha
0.3, omega_c 10, bond dimension 256, admixture 1e-16) on the device.  It is
real, of numerical rank about 5, with eigenvalues from 1 down to 1e-24 and
diagonal entries down to 1e-47; rocSOLVER's zheevd returned NaN eigenvectors
(info 0) for it.  The committed file was captured while reno_block_eigh still
solved every block with zheevd; with the current library (zheevdj first) the
earlier blocks round differently and this script writes the same matrix to
rounding (its eigenvalues above 1e-12 agree to 1e-9 relative).
Run on a GPU:  python tests/golden/make_gram_low_rank.py"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))

from renormalizer_amd import native  # noqa: E402

OUT = os.path.join(os.path.dirname(__file__), "gram_low_rank_81.npy")


class _Found(Exception):
    pass


def main():
    orig = native.block_eigh

    def capture(dm, sectors):
        if dm.shape[0] == 81:
            a = dm.cpu().numpy()
            assert np.abs(a.imag).max() == 0
            np.save(OUT, np.ascontiguousarray(a.real))
            raise _Found
        return orig(dm, sectors)

    native.block_eigh = capture
    from renormalizer_amd import CompressConfig, CompressCriteria, Mpo, Mps, Phonon, Quantity, SpinBosonModel
    from renormalizer_amd.sbm import SpectralDensityFunction
    sdf = SpectralDensityFunction(0.3, 10.0)
    w, c2 = sdf.Wang1(64)
    omegas, dis = sdf.post_process(w, c2)
    model = SpinBosonModel(Quantity(0.0), Quantity(1.0),
                           [Phonon.simple_phonon(o, d, 10) for o, d in zip(omegas, dis)])
    mps = Mps.ground_state(model, False)
    mps.compress_config = CompressConfig(CompressCriteria.fixed, max_bonddim=256)
    mps.to_device()
    try:
        mps.expand_bond_dimension(Mpo(model), coef=1e-16, include_ex=False)
    except _Found:
        print("wrote", OUT)
        return
    raise SystemExit("no 81 x 81 Gram matrix met")


if __name__ == "__main__":
    main()
