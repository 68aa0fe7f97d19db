from renormalizer_amd.mps.mpo import Mpo
from renormalizer_amd.mps.mps import Mps, expand_bond_dimension
from renormalizer_amd.mps.lib import Environ, contract_one_site
from renormalizer_amd.mps.gs import construct_mps_mpo, optimize_mps
from renormalizer_amd.mps.mpdm import MpDm
from renormalizer_amd.mps.thermalprop import ThermalProp, load_thermal_state
