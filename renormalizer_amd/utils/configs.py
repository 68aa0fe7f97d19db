"""Configuration objects of the hot path (reference utils/configs.py:41-416).

Names, defaults and semantics follow the reference.  One field is new:
``EvolveConfig.precision`` selects the device arithmetic of the TDVP-PS
sweep: "fp64" (complex128 everywhere); "mixed" (complex128 environments,
QR, Krylov vectors, reductions and tridiagonal exponential; each Lanczos
iteration's matvec on FP64 matrix cores while the step's result still
weighs the Lanczos vectors it creates by at least ``mixed_tau``, on FP32
matrix cores with complex64 copies of the environments after that);
"fp32" (every contraction on FP32 matrix cores with complex64
environments; kept for comparison, it misses the 1e-6 population bound).
``EvolveConfig.autotune`` picks the GEMM tile and split-K of each matvec
shape by timing the candidates on the device (off by default: the choice
depends on measured times, so results can differ between runs at rounding
level).
"""
from enum import Enum
from typing import Union

import numpy as np
import scipy.linalg


class CompressCriteria(Enum):
    threshold = "threshold"
    fixed = "fixed"
    both = "both"


class CompressConfig:
    """MPS compression configuration (reference configs.py:41-264).
    Variational-compression and disk-dump options are accepted for API
    compatibility; only SVD compression is implemented."""

    def __init__(self, criteria: Union[CompressCriteria, str] = CompressCriteria.threshold,
                 threshold: float = 1e-3, max_bonddim: int = 32, vmethod: str = "2site",
                 vprocedure=None, vrtol=1e-5, vguess_m=(5, 5), dump_matrix_size=np.inf,
                 dump_matrix_dir="./", ofs=None, ofs_swap_jw: bool = False):
        if isinstance(criteria, str):
            criteria = getattr(CompressCriteria, criteria)
        self.criteria: CompressCriteria = criteria
        self._threshold = None
        self.threshold = threshold
        self.bond_dim_max_value = max_bonddim
        self.max_dims: np.ndarray = None
        self.vmethod = vmethod
        if vprocedure is None:
            if vmethod == "1site":
                vprocedure = [[max_bonddim, 1.0], [max_bonddim, 0.7], [max_bonddim, 0.5],
                              [max_bonddim, 0.3], [max_bonddim, 0.1]] + [[max_bonddim, 0]] * 10
            else:
                vprocedure = [[max_bonddim, 0.5], [max_bonddim, 0.3],
                              [max_bonddim, 0.1]] + [[max_bonddim, 0]] * 10
        self.vprocedure = vprocedure
        self.vrtol = vrtol
        self.vguess_m = vguess_m
        self.dump_matrix_size = dump_matrix_size
        self.dump_matrix_dir = dump_matrix_dir
        if ofs is not None:
            raise NotImplementedError("on-the-fly swapping is outside the TDVP-PS hot path")
        self.ofs = ofs
        self.ofs_swap_jw = ofs_swap_jw

    @property
    def threshold(self):
        return self._threshold

    @threshold.setter
    def threshold(self, v):
        if v <= 0:
            raise ValueError("non-positive threshold")
        if v == 1:
            raise ValueError("1 is an ambiguous threshold")
        if v > 1:
            raise ValueError("Can't set threshold to be larger than 1")
        self._threshold = v

    def set_bonddim(self, length):
        if self.max_dims is None:
            self.max_dims = np.full(length, self.bond_dim_max_value, dtype=int)

    def _threshold_m_trunc(self, sigma):
        assert 0 < self.threshold < 1
        normed = sigma / scipy.linalg.norm(sigma)
        return int(np.sum(normed > self.threshold))

    def _fixed_m_trunc(self, sigma, idx, left):
        assert self.max_dims is not None
        bond_idx = idx + 1 if left else idx
        return min(self.max_dims[bond_idx], len(sigma))

    def compute_m_trunc(self, sigma, idx, left) -> int:
        if self.criteria is CompressCriteria.threshold:
            return self._threshold_m_trunc(sigma)
        if self.criteria is CompressCriteria.fixed:
            return self._fixed_m_trunc(sigma, idx, left)
        if self.criteria is CompressCriteria.both:
            return min(self._threshold_m_trunc(sigma), self._fixed_m_trunc(sigma, idx, left))
        raise AssertionError

    def update(self, other: "CompressConfig"):
        if self.criteria != other.criteria:
            raise ValueError("Can't update configs with different standard")
        self.threshold = min(self.threshold, other.threshold)
        if self.max_dims is None:
            self.max_dims = other.max_dims
        elif other.max_dims is not None:
            self.max_dims = np.maximum(self.max_dims, other.max_dims)

    def relax(self):
        self.threshold = min(self.threshold * 3, 0.9)
        if self.max_dims is not None:
            self.max_dims = np.maximum(np.int64(self.max_dims * 0.8), np.full_like(self.max_dims, 2))

    def copy(self) -> "CompressConfig":
        new = self.__class__.__new__(self.__class__)
        new.__dict__ = self.__dict__.copy()
        if self.max_dims is not None:
            new.max_dims = self.max_dims.copy()
        return new

    @property
    def bonddim_should_set(self):
        return self.criteria is not CompressCriteria.threshold and self.max_dims is None

    def __str__(self):
        return f"\ncriteria: {self.criteria}\nthreshold: {self.threshold}"


class OptimizeConfig:
    """DMRG configuration (reference configs.py:267-300)."""

    def __init__(self, procedure=None):
        self.procedure = procedure if procedure is not None else \
            [[10, 0.4], [20, 0.2], [30, 0.1], [40, 0], [40, 0]]
        self.method = "2site"
        self.algo = "davidson"
        self.nroots = 1
        self.e_rtol = 1e-6
        self.e_atol = 1e-8
        self.inverse = 1.0

    def copy(self):
        new = self.__class__.__new__(self.__class__)
        new.__dict__ = self.__dict__.copy()
        new.procedure = self.procedure.copy()
        return new


class EvolveMethod(Enum):
    prop_and_compress = "P&C"
    prop_and_compress_tdrk4 = "P&C TD RK4"
    prop_and_compress_tdrk = "P&C TD RK"
    tdvp_ps = "TDVP PS one-site"
    tdvp_ps2 = "TDVP PS two-site"
    tdvp_vmf = "TDVP Variable Mean Field"
    tdvp_mu_cmf = "TDVP Matrix Unfolding Constant Mean Field"
    tdvp_mu_vmf = "TDVP Matrix Unfolding Variable Mean Field"


PRECISIONS = ("fp64", "mixed", "fp32")


class EvolveConfig:
    """Time-evolution configuration (reference configs.py:342-416)."""

    def __init__(self, method: Union[EvolveMethod, str] = EvolveMethod.prop_and_compress,
                 adaptive=False, guess_dt=1e-1, adaptive_rtol=5e-4, taylor_order: int = None,
                 rk_solver="C_RK4", reg_epsilon=1e-10, ivp_rtol=1e-5, ivp_atol=1e-8,
                 ivp_solver="krylov", force_ovlp=True, precision: str = "fp64",
                 krylov_async: bool = True, mixed_tau: float = 1e-2, autotune: bool = False):
        if isinstance(method, str):
            method = getattr(EvolveMethod, method)
        self.method = method
        self.adaptive = adaptive
        self.rk_solver = rk_solver
        if taylor_order is None:
            taylor_order = 5 if adaptive else 4
        self.taylor_order = taylor_order
        self.guess_dt: complex = guess_dt
        self.adaptive_rtol = adaptive_rtol
        self.tdvp_cmf_midpoint = True
        self.tdvp_cmf_c_trapz = False
        self.reg_epsilon: float = reg_epsilon
        self.ivp_rtol: float = ivp_rtol
        self.ivp_atol: float = ivp_atol
        self.ivp_solver: str = ivp_solver
        self.force_ovlp: bool = force_ovlp
        self.vmf_auto_switch: bool = True
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self.precision = precision
        # device sweep: queue the Lanczos iterations of a whole time step
        # without waiting for the device (iteration budgets from the previous
        # step; the step is redone with waits when a budget proves too small)
        self.krylov_async = bool(krylov_async)
        # mixed mode: a Lanczos iteration runs its matvec in FP32 once the
        # coefficient of the previous Lanczos vector in the step's result is
        # below mixed_tau times the norm of the starting vector
        self.mixed_tau = float(mixed_tau)
        # asynchronous device sweep: time the tile / split-K candidates of
        # each new matvec shape once (native.autotune) and use the fastest
        self.autotune = bool(autotune)
        self.stat = None

    @property
    def is_tdvp(self):
        return self.method not in (EvolveMethod.prop_and_compress,
                                   EvolveMethod.prop_and_compress_tdrk4,
                                   EvolveMethod.prop_and_compress_tdrk)

    def check_valid_dt(self, evolve_dt: complex):
        info = f"in config: {self.guess_dt}, in arg: {evolve_dt}"
        if np.iscomplex(evolve_dt) ^ np.iscomplex(self.guess_dt):
            raise ValueError("real and imag not compatible. " + info)
        if (np.iscomplex(evolve_dt) and evolve_dt.imag * self.guess_dt.imag < 0) or \
                (not np.iscomplex(evolve_dt) and evolve_dt * self.guess_dt < 0):
            raise ValueError("evolve into wrong direction. " + info)

    def copy(self):
        new = self.__class__.__new__(self.__class__)
        new.__dict__ = self.__dict__.copy()
        return new

    def __str__(self):
        return "".join(f"\n{k}: {v}" for k, v in self.__dict__.items())
