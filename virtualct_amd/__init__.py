"""virtualct_amd -- MI355X-native MBFPA forward projection and fan-beam FBP.

The compute path is libvct.so (hand-written HIP for gfx950 behind the C ABI
in include/vct.h).  This package binds it with ctypes and uses PyTorch only
for device memory and streams.
"""
from ._lib import (DENS_F32, DENS_U16, FILTER_RAMLAK, FILTER_SHEPP_LOGAN, Geometry, VctError,
                   default_geometry, lib)

__all__ = ["DENS_F32", "DENS_U16", "FILTER_RAMLAK", "FILTER_SHEPP_LOGAN", "Geometry", "VctError",
           "default_geometry", "lib"]
__version__ = "0.1.0"
