"""mom5_amd -- MI355X (gfx950) implementation of the MOM5 ocean hot path.

The compute runs in the HIP library ``libmom5_gpu.so`` (C-ABI in
``include/mom5_gpu.h``); this package holds its ctypes binding
(``_lib``), the host-side mirror of the reference interface (``ocean``),
the seeded synthetic-state generator (``synth``) and the build recipe
(``build``).
"""
from .ocean import (ADVECT_DST_LINEAR, ADVECT_MDFL_SWEBY, ADVECT_QUICKER, DeviceArray,
                    MomGpuError, OceanGPU)

__all__ = ["OceanGPU", "DeviceArray", "MomGpuError", "ADVECT_QUICKER", "ADVECT_MDFL_SWEBY",
           "ADVECT_DST_LINEAR"]
