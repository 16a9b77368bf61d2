"""MI355X-native implementation of RMT3D's per-timestep fluid-structure step.

The compute path is the HIP library librmt3d.so behind the C ABI of
include/rmt3d.h; this package only loads it (rmt3d_amd.lib) and mirrors its
structs (rmt3d_amd._abi).
"""
from . import _abi  # noqa: F401
