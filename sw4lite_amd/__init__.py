"""sw4lite_amd -- MI355X (gfx950) implementation of the sw4lite elastic-wave
time-step hot path behind the reference's EW / kernel interfaces.

The compute lives in ``libsw4hip.so`` (HIP kernels + C ABI, include/sw4hip.h);
this package holds the build script and thin Python bindings.
"""
from .lib import (BC_DIRICHLET, BC_FREE, BC_PERIODIC, BC_SUPERGRID, TF_C6SMOOTHBUMP,  # noqa: F401
                  TF_GAUSSIAN, TF_RICKER, Sw4HipError, load)
from .ew import EW, Batch  # noqa: F401
