"""WaterCube SPH step on AMD Instinct MI355X (gfx950).

The GPU work lives in ``libwatercube.so`` (hand-written HIP kernels behind the C ABI of
``include/watercube.h``).  This package is its Python host side: ctypes bindings, the
``Fluid`` / ``Sort`` stage classes that mirror the reference's ``src/core`` classes, and scene
generators.
"""
from . import _lib
from ._lib import PARTICLE_DTYPE, WaterCubeError, preset, wc_params
from .fluid import Fluid, Sort
from . import scenes

__all__ = ["Fluid", "Sort", "PARTICLE_DTYPE", "WaterCubeError", "preset", "wc_params", "scenes", "_lib"]
