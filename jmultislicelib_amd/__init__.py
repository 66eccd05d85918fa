"""jmultislicelib_amd -- MI355X-native multislice slice loop.

The library (libjmultislice_hip.so, built from ``csrc/``) implements the
reference's CJMultiSlice GPU multislice path with hand-written HIP kernels
for gfx950; this package is its Python binding:

  * ``JMultiSlice``  -- the reference's flat C interface (setup chain,
    OffsetIncomingWave, GPUMultislice, GetResult, batched scans)
  * ``Engine``       -- the batched engine ABI (include/jms_hip.h)
  * ``configs``      -- benchmark workloads (synthetic perovskite specimens)
"""
from ._lib import EXPORTED, LIB_PATH, lib  # noqa: F401
from .jms import Engine, Group, JMSError, JMultiSlice, fft2d, grid_supported  # noqa: F401

__all__ = ["Engine", "Group", "JMultiSlice", "JMSError", "fft2d", "grid_supported", "lib", "LIB_PATH"]
