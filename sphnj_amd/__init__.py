"""sphnj_amd -- MI355X (gfx950) implementation of SPHNJ's per-step particle pipeline.

The compute path is libsphnj.so (HIP kernels behind the C ABI of include/sphnj.h); this
package holds its build script and a ctypes host mirror of the reference's operator API.
"""
from ._abi import Params  # noqa: F401
from .sim import (  # noqa: F401
    Case,
    Simulation,
    SphnjError,
    default_params,
    grid,
    library,
    params_from_ini,
    validate_params,
)
