"""MI355X-native hip platform for the alsfvm finite-volume timestep and the
alsuq sample statistics: gfx950 HIP kernels behind a C ABI
(include/alsvinn_hip.h), with this thin Python host layer over it."""
from . import _lib
from ._lib import (EQ_BURGERS, EQ_EULER1, EQ_EULER2, EQ_EULER3, FLUX_HLL, FLUX_HLL3,
                   FLUX_GODUNOV, FLUX_CENTRAL, FLUX_RUSANOV, FLUX_ROE, RECON_NONE, RECON_WENO2,
                   RECON_ENO2, RECON_ENO3, RECON_ENO4, RECON_MC, RECON_MM, RECON_WENOF2,
                   RECON_WENO3, INT_FE, INT_RK2, INT_RK3, INT_RK4, BC_NEUMANN, BC_PERIODIC, FP64,
                   FP32, HipError)
from .ensemble import Ensemble, Pool, make_scheme, make_sample, state_elements

__all__ = ["Ensemble", "Pool", "make_scheme", "make_sample", "state_elements", "HipError"]
