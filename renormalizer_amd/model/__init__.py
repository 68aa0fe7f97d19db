from renormalizer_amd.model.op import Op, OpSum
from renormalizer_amd.model.basis import BasisSet, BasisSHO, BasisHalfSpin, BasisSimpleElectron
from renormalizer_amd.model.phonon import Phonon
from renormalizer_amd.model.mol import Mol
from renormalizer_amd.model.model import Model, HolsteinModel, SpinBosonModel
