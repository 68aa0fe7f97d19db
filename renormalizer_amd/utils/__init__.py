from renormalizer_amd.utils.quantity import Quantity
from renormalizer_amd.utils.configs import (CompressCriteria, CompressConfig, OptimizeConfig,
                                            EvolveMethod, EvolveConfig)
from renormalizer_amd.utils.tdmps import TdMpsJob
