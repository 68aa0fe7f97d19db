"""Spin-boson dynamics job (the driver of reference sbm/sbm.py:13-92).

Spin up x vibrational vacuum, expanded along the Hamiltonian with a 1e-16
admixture to the compress config's bond dimension when the method is TDVP;
after every step the spin's reduced density matrix, <sigma_z>, <sigma_x>
and the von Neumann entropy of every bond are recorded.
"""
import logging

from renormalizer_amd.model import Model
from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps
from renormalizer_amd.utils import CompressConfig
from renormalizer_amd.utils.tdmps import TdMpsJob

logger = logging.getLogger(__name__)


class SpinBosonDynamics(TdMpsJob):
    def __init__(self, model: Model, auto_expand: bool = True, compress_config=None, evolve_config=None,
                 dump_dir=None, dump_mps=None, job_name=None, bond_entropy: bool = True):
        self.model = model
        self.h_mpo = Mpo(model)
        self.auto_expand = auto_expand
        self.compress_config = CompressConfig() if compress_config is None else compress_config
        # the entropy of every bond is a host SVD sweep of the whole state
        # per step; bond_entropy=False skips it for long large-M runs
        self.record_bond_entropy = bond_entropy
        self.sigma_x, self.sigma_z, self.rho, self.bond_entropy = [], [], [], []
        super().__init__(evolve_config=evolve_config, dump_dir=dump_dir, dump_mps=dump_mps,
                         job_name=job_name)

    def init_mps(self):
        mps = Mps.ground_state(self.model, False)
        mps.compress_config = self.compress_config
        mps.evolve_config = self.evolve_config
        if self.evolve_config.is_tdvp and self.auto_expand:
            mps = mps.expand_bond_dimension(self.h_mpo, coef=1e-16, include_ex=False)
        return mps

    def process_mps(self, mps):
        spin = next(i for i, b in enumerate(self.model.basis) if b.is_spin)
        rho = mps.calc_1site_rdm(idx=spin)[spin]
        self.rho.append(rho)
        self.sigma_z.append(float((rho[0, 0] - rho[1, 1]).real))
        self.sigma_x.append(float((rho[0, 1] + rho[1, 0]).real))
        if self.record_bond_entropy:
            self.bond_entropy.append(mps.calc_entropy("bond"))

    def evolve_single_step(self, evolve_dt):
        return self.latest_mps.evolve(self.h_mpo, evolve_dt)

    def get_dump_dict(self):
        return {"time series": self.evolve_times, "sigma_x": self.sigma_x, "sigma_z": self.sigma_z,
                "rho": self.rho, "bond_entropy": self.bond_entropy}
