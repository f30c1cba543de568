"""MI355X-native progressive-refinement radiosity (hemicube shooting).

The compute path is libradhip.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/radhip.h); this package only binds it with ctypes.
"""
from ._lib import (RAD_DENSE_ROWS, RAD_DENSE_SHOOT, RAD_FACES_QUADS,  # noqa: F401
                   RAD_FACES_TRIANGLES, RAD_MODE_PHYSICAL, RAD_MODE_REFERENCE, RAD_SELECT_POWER,
                   RAD_SELECT_REFERENCE, RadError, Scene, Solver, cornell_faces, lib, obj_parse)
