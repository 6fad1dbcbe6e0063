"""MI355X-native implementation of ProjecTracer's per-timestep hot path.

The compute path is libprojectracer_gpu.so (HIP, gfx950) behind the C ABI
of include/projectracer_gpu.h; `Simulation` is a thin host-side mirror of
the reference's subroutine interface over that ABI.
"""
from ._abi import (PtParams, PtSolveInfo, default_params, load_library, FIELD,  # noqa: F401
                   PT_PREC_FP64, PT_PREC_MIXED, PT_GUESS_WARM, PT_GUESS_LEARNED)
from .simulation import Simulation, PtError  # noqa: F401
