"""gray_amd: MI355X-native GRay2 geodesic engine.

The compute path is the gfx950 HIP library libgrayhip.so behind the C ABI of
include/grayhip.h; ``gray_amd.grayhip`` is its ctypes binding.
"""
from . import grayhip  # noqa: F401

__all__ = ["grayhip"]
