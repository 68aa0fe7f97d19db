"""Harmonic mode with a displaced excited-state surface (reference
model/phonon.py:22-154, the parts used by the Holstein / spin-boson models)."""
from typing import List

import numpy as np

from renormalizer_amd.utils import Quantity


class Phonon:
    @classmethod
    def simple_phonon(cls, omega: Quantity, displacement: Quantity, n_phys_dim: int):
        return cls([omega, omega], [Quantity(0), displacement], n_phys_dim)

    def __init__(self, omega: List[Quantity], displacement: List[Quantity], n_phys_dim: int = None):
        self.omega = [o.as_au() for o in omega]
        self.dis = [d.as_au() for d in displacement]
        self.n_phys_dim = n_phys_dim

    @property
    def pbond(self):
        return self.n_phys_dim

    nlevels = pbond

    @property
    def reorganization_energy(self):
        return Quantity(0.5 * (self.dis[1] - self.dis[0]) ** 2 * self.omega[1] ** 2)

    @property
    def e0(self):
        return self.reorganization_energy

    @property
    def is_simple(self):
        return self.omega[0] == self.omega[1]

    @property
    def coupling_constant(self):
        """g = sqrt(lambda / omega)."""
        return float(np.sqrt(self.reorganization_energy.as_au() / self.omega[0]))

    @property
    def term10(self):
        return self.omega[1] ** 2 / np.sqrt(2.0 * self.omega[0]) * (-self.dis[1])

    def __eq__(self, other):
        return self.__dict__ == other.__dict__
