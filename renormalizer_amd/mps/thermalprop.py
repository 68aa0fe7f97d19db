"""Imaginary-time thermal propagation of a purified density operator
(reference renormalizer/mps/thermalprop.py:13-168).

rho(beta) is proportional to exp(-beta H / 2) rho(0) exp(-beta H / 2); in the
purification picture the state is exp(-beta H / 2) |rho(0)^(1/2)>, obtained by
evolving the MpDm in imaginary time tau = beta / 2 (``evolve_dt`` = tau / 1j,
negative imaginary part) with the device TDVP-PS sweep.  Each step shifts the
Hamiltonian by the current energy, as the reference does, so the norm stays
of order one.  The time-stepping driver keeps the reference's bookkeeping:
``energies``, electron / phonon occupations and bond entropies after every
step.
"""
import logging

import numpy as np

from renormalizer_amd.mps.mpdm import MpDm
from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.utils import EvolveConfig, Quantity

logger = logging.getLogger(__name__)


class ThermalProp:
    """reference thermalprop.py:13-148.  ``exact=True`` (propagation by
    the exact local propagator, Mpo.exact_propagator) is not on the device
    path and raises NotImplementedError."""

    def __init__(self, init_mpdm: MpDm, h_mpo_model=None, exact: bool = False, space: str = "GS",
                 evolve_config: EvolveConfig = None, dump_mps=None, dump_dir: str = None,
                 job_name: str = None, properties=None, auto_expand: bool = True):
        if exact:
            raise NotImplementedError("exact thermal propagation (Mpo.exact_propagator) is not on "
                                      "the device path; use TDVP-PS")
        if dump_mps is not None or dump_dir is not None or properties is not None:
            raise NotImplementedError("job dumping and Property hooks are not provided")
        self.init_mpdm: MpDm = init_mpdm.canonicalise()
        if h_mpo_model is None:
            h_mpo_model = self.init_mpdm.model
        self.h_mpo = Mpo(h_mpo_model)
        self.exact = exact
        assert space in ["GS", "EX"]
        self.space = space
        self.energies = []
        self._e_occupations_array = []
        self._ph_occupations_array = []
        self._vn_entropy_array = []
        self.auto_expand = auto_expand
        self.evolve_config = EvolveConfig() if evolve_config is None else evolve_config
        self.job_name = job_name
        self.evolve_times = []
        self.latest_mps = None

    def init_mps(self) -> MpDm:
        self.init_mpdm.evolve_config = self.evolve_config
        if self.evolve_config.is_tdvp and self.auto_expand:
            self.init_mpdm = self.init_mpdm.expand_bond_dimension(self.h_mpo)
        return self.init_mpdm

    def process_mps(self, mps: MpDm):
        """reference thermalprop.py:77-99."""
        new_energy = mps.expectation(self.h_mpo)
        self.energies.append(new_energy)
        for attr_str in ["e_occupations", "ph_occupations"]:
            attr = getattr(mps, attr_str)
            logger.info(f"{attr_str}: {attr}")
            getattr(self, f"_{attr_str}_array").append(attr)
        vn_entropy = mps.calc_bond_entropy()
        self._vn_entropy_array.append(vn_entropy)
        logger.info(f"Energy: {new_energy}, total electron: {self._e_occupations_array[-1].sum()}")

    def evolve_prop(self, old_mpdm: MpDm, evolve_dt):
        h_mpo = Mpo(self.h_mpo.model, offset=Quantity(self.energies[-1]))
        return old_mpdm.evolve(h_mpo, evolve_dt)

    def evolve_single_step(self, evolve_dt):
        return self.evolve_prop(self.latest_mps, evolve_dt)

    def evolve(self, evolve_dt=None, nsteps=None, evolve_time=None):
        """reference thermalprop.py:123-128 and the TdMpsJob loop
        (utils/tdmps.py:68-160): ``nsteps`` steps of ``evolve_dt``, or
        ``evolve_time`` split into ``nsteps`` steps."""
        if evolve_dt is not None:
            assert np.iscomplex(evolve_dt) and evolve_dt.imag < 0
        if evolve_time is not None:
            assert np.iscomplex(evolve_time) and evolve_time.imag < 0
        if evolve_dt is None:
            if evolve_time is None or nsteps is None:
                raise ValueError("give evolve_dt and nsteps, or evolve_time and nsteps")
            evolve_dt = evolve_time / nsteps
        elif nsteps is None:
            if evolve_time is None:
                raise ValueError("give nsteps or evolve_time")
            nsteps = int(round((evolve_time / evolve_dt).real))
        if self.latest_mps is None:
            mps = self.init_mps()
            self.evolve_times.append(0)
            self.process_mps(mps)
            self.latest_mps = mps
        for _ in range(nsteps):
            new = self.evolve_single_step(evolve_dt)
            self.evolve_times.append(self.evolve_times[-1] + evolve_dt)
            self.process_mps(new)
            self.latest_mps = new
        return self.latest_mps

    @property
    def e_occupations_array(self):
        return np.array(self._e_occupations_array)

    @property
    def ph_occupations_array(self):
        return np.array(self._ph_occupations_array)

    @property
    def vn_entropy_array(self):
        return np.array(self._vn_entropy_array)

    def get_dump_dict(self):
        """reference thermalprop.py:142-156 (without Property results)."""
        return {
            "time series": [-complex(t).imag for t in self.evolve_times],
            "energies": self.energies,
            "electron occupations array": self.e_occupations_array.tolist(),
            "phonon occupations array": self.ph_occupations_array.tolist(),
            "vn entropy array": self.vn_entropy_array.tolist(),
        }


def load_thermal_state(model, path: str):
    """reference thermalprop.py:159-168: the MpDm stored at ``path``, or
    None when there is no such file."""
    try:
        return MpDm.load(model, path)
    except FileNotFoundError:
        return None
