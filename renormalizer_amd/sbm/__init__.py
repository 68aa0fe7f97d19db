from renormalizer_amd.sbm.lib import (SpectralDensityFunction, OhmicSDF, DebyeSpectralDensityFunction,
                                      DebyeSDF)
from renormalizer_amd.sbm.sbm import SpinBosonDynamics
