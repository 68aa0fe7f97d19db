"""MI355X-native TDVP-PS / DMRG local-update path of Renormalizer.

Public names mirror the reference package at the hot path: models and
operators (``Model``, ``HolsteinModel``, ``SpinBosonModel``, ``Op``,
``OpSum``, bases), ``Mpo``, ``Mps`` (``evolve`` with ``EvolveMethod.tdvp_ps``,
observables), ``optimize_mps`` and the configuration classes.  The compute
path is ``libreno_hip.so`` (see ``renormalizer_amd.native``).
"""
import os as _os

# thread count for host-side numpy work, as the reference's RENO_NUM_THREADS
_n = _os.environ.get("RENO_NUM_THREADS")
if _n is not None:
    for _k in ("MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS", "OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS"):
        _os.environ.setdefault(_k, _n)

from renormalizer_amd.utils import (Quantity, CompressCriteria, CompressConfig, OptimizeConfig,  # noqa: E402
                                    EvolveMethod, EvolveConfig)
from renormalizer_amd.model import (Op, OpSum, BasisSet, BasisSHO, BasisHalfSpin,  # noqa: E402
                                    BasisSimpleElectron, Phonon, Mol, Model, HolsteinModel,
                                    SpinBosonModel)
from renormalizer_amd.mps import (Mpo, Mps, MpDm, ThermalProp, construct_mps_mpo,  # noqa: E402
                                  optimize_mps)
