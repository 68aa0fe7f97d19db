"""Physical quantity with a unit, converted to atomic units on demand
(reference utils/quantity.py:35-103)."""
import math

from renormalizer_amd.utils import constant

_AU_PER_UNIT = {
    "meV": constant.au2ev * 1e3,
    "eV": constant.au2ev,
    "cm^{-1}": 1 / constant.cm2au,
    "cm-1": 1 / constant.cm2au,
    "K": constant.au2K,
    "a.u.": 1,
    "au": 1,
    "fs": constant.au2fs,
}
_AU_PER_UNIT.update({k.lower(): v for k, v in _AU_PER_UNIT.items()})


class Quantity:
    def __init__(self, value, unit="a.u."):
        self.value = float(value)
        if unit not in _AU_PER_UNIT:
            raise ValueError(f"Unit not in {set(_AU_PER_UNIT)}, got {unit}.")
        self.unit = unit

    def as_au(self):
        return self.value / _AU_PER_UNIT[self.unit]

    def as_unit(self, unit):
        return Quantity(self.as_au() * _AU_PER_UNIT[unit], unit)

    def to_beta(self):
        return math.inf if self.value == 0 else 1.0 / self.as_au()

    def __neg__(self):
        return Quantity(-self.value, self.unit)

    def __add__(self, other):
        assert isinstance(other, Quantity)
        return Quantity(self.as_au() + other.as_au())

    def __sub__(self, other):
        assert isinstance(other, Quantity)
        return Quantity(self.as_au() - other.as_au())

    def __mul__(self, other):
        assert not isinstance(other, Quantity)
        return Quantity(self.as_au() * other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        assert not isinstance(other, Quantity)
        return Quantity(self.as_au() / other)

    def __eq__(self, other):
        if hasattr(other, "as_au"):
            return self.as_au() == other.as_au()
        if other == 0:
            return self.value == 0
        raise TypeError(f"Quantity can only compare with Quantity or 0, not {other.__class__}")

    def __ne__(self, other):
        return not self == other

    def __str__(self):
        return "%g %s" % (self.value, self.unit)

    __repr__ = __str__
