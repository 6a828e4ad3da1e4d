"""fluidsim_amd: MI355X (gfx950) HIP implementation of fluidsim's plasma-transport
hot path (SG fluxes, continuity/energy RHS, RK-Merson, semi-implicit Poisson).

The compute lives in libfsim_hip.so (C ABI: include/fsim.h).  This package is
the Python host; the Fortran host is under fluidsim_amd/fortran.
"""
from .configs import config0, config1, config2  # noqa: F401
from .discharge import Discharge, constants, eval_dfdph, eval_fits, eval_flux  # noqa: F401

__all__ = ["Discharge", "constants", "eval_fits", "eval_flux", "eval_dfdph", "config0",
           "config1", "config2"]
