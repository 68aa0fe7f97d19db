"""Spectral densities and their discretisation into harmonic modes
(reference sbm/lib.py:38-136): Ohmic J(w) = pi/2 alpha w^s w_c^(1-s) e^(-w/w_c),
Wang's first (logarithmic) scheme and the trapezoid grid."""
from typing import Union

import numpy as np

from renormalizer_amd.utils import Quantity


class DebyeSpectralDensityFunction:
    """J(w) = 2 lambda w w_c / (w^2 + w_c^2)."""

    def __init__(self, lamb, omega_c):
        self.lamb = lamb
        self.omega_c = omega_c

    def func(self, omega_value):
        return 2.0 * self.lamb * omega_value * self.omega_c / (omega_value ** 2 + self.omega_c ** 2)


DebyeSDF = DebyeSpectralDensityFunction


class SpectralDensityFunction:
    def __init__(self, alpha: float, omega_c: Union[Quantity, float], s: float = 1):
        self.alpha = alpha
        self.omega_c = omega_c.as_au() if isinstance(omega_c, Quantity) else omega_c
        self.s = s

    def func(self, omega_value):
        return (np.pi / 2.0 * self.alpha * omega_value ** self.s * self.omega_c ** (1 - self.s)
                * np.exp(-omega_value / self.omega_c))

    @staticmethod
    def post_process(omega_value, c_j2, ifsort=True):
        """Frequencies and displacements D_j = sqrt(c_j^2) / w_j^2, sorted by
        c_j^2 / w_j descending (the reference's mode order)."""
        displacement = np.sqrt(c_j2) / omega_value ** 2
        idx = np.argsort(c_j2 / omega_value)[::-1] if ifsort else np.arange(len(omega_value))
        return ([Quantity(omega_value[i]) for i in idx], [Quantity(displacement[i]) for i in idx])

    def _dos_Wang1(self, nb, omega_value):
        return (nb + 1) / self.omega_c * np.exp(-omega_value / self.omega_c)

    def Wang1(self, nb):
        """w_j = -w_c ln(1 - j/(nb+1)), c_j^2 = 2/pi w_j J(w_j) / rho(w_j)."""
        omega_value = np.array([-np.log(-float(j) / (nb + 1) + 1.0) * self.omega_c for j in range(1, nb + 1)])
        c_j2 = 2.0 / np.pi * omega_value * self.func(omega_value) / self._dos_Wang1(nb, omega_value)
        return omega_value, c_j2

    def trapz(self, nb, x0, x1):
        dw = (x1 - x0) / float(nb)
        xs = [x0 + i * dw for i in range(nb + 1)]
        omega_value = np.array([(xs[i] + xs[i + 1]) / 2.0 for i in range(nb)])
        c_j2 = (np.array([(self.func(xs[i]) + self.func(xs[i + 1])) / 2 for i in range(nb)])
                * 2.0 / np.pi * omega_value * dw)
        return omega_value, c_j2


OhmicSDF = SpectralDensityFunction
