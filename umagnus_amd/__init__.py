"""umagnus_amd -- MI355X (gfx950) implementation of the uMagNUS effective-field
and time-stepping path behind a C ABI (include/umhip.h, libumhip.so).

    umagnus_amd.build       builds libumhip.so in-tree (hipcc, gfx950)
    umagnus_amd.umhip       ctypes binding: opencl-package mirror + Sim handle
    umagnus_amd.params      host-side exchange/DMI region tables
"""
from . import params, umhip
from .params import exchange_lut, symmidx
from .umhip import (F32, F64, SOLVER_HEUN, SOLVER_RK4, SOLVER_RK45DP, DemagConvolution, Sim, UmError,
                    load)

__all__ = ["umhip", "params", "Sim", "DemagConvolution", "exchange_lut", "symmidx", "load", "UmError",
           "F32", "F64", "SOLVER_HEUN", "SOLVER_RK4", "SOLVER_RK45DP"]
