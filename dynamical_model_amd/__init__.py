"""MI355X-native MOLOCH non-hydrostatic dynamical core.

A drop-in for the timestep ``moloch(istep, lprint)`` of the reference
(nompi/mod_moloch.F90:272-421): hand-written HIP kernels for gfx950 behind
the C ABI of include/moloch_hip.h, with a Fortran module
(fortran/mod_moloch.F90) and this Python mirror of the reference's
mod_moloch interface on top.
"""
from .core import (  # noqa: F401
    MLC_FP32,
    MLC_FP64,
    HostState,
    Moloch,
    build_case,
    canonical_shape,
    case_config,
    case_params,
    default_params,
    interior,
    run_case,
)
from ._lib import FIELDS, LIB_PATH, MolochError  # noqa: F401
