"""numericalearth_hip — MI355X (gfx950) drop-in for NumericalEarth.jl's coupled surface time step.

The HIP kernels sit behind the C ABI declared in include/numericalearth_hip.h
(built into numericalearth.jl_amd/lib/libnumericalearth_hip.so). This package is
the host-side mirror of the reference interface (OceanOnlyModel, time_step!, the
flux-formulation types) over that ABI; it holds no numerical fallback.
"""
from ._ctypes_abi import *  # noqa: F401,F403
from .params import *  # noqa: F401,F403

__version__ = "0.1.0"
