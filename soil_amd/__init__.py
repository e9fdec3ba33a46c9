"""soil_amd — MI355X (gfx950) HIP implementation of SOIL's per-step particle
substrate: binning, CSR half neighbor lists with skin, periodic and
Lees–Edwards ghost images, the displacement rebuild trigger and velocity
Verlet, behind the C ABI of include/soil_hip.h.
"""
from . import _ffi  # noqa: F401
from .substrate import (Batch, Box, Context, FULL_REBUILD, COMMUNICATE_ONLY, device_count, gen_arrivals,  # noqa: F401
                        gen_box, gen_custom, make_desc, desc_to_dict)

__all__ = ["Batch", "Box", "Context", "device_count", "gen_box", "gen_custom", "gen_arrivals", "make_desc"]
