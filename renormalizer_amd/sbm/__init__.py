from renormalizer_amd.sbm.lib import (SpectralDensityFunction, OhmicSDF, DebyeSpectralDensityFunction,
                                      DebyeSDF)
