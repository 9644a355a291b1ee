"""amssbssn -- MI355X (gfx950) BSSN right-hand side + RK4 of AMSS-NCKU.

The compute path is libamssbssn.so (hand-written HIP kernels behind a C ABI,
see include/amss_bssn.h); this package is the Python driver around it.
"""
from ._lib import AmssError, lib, last_error, LIB_PATH  # noqa: F401
from .grid import (FIELDS, NF, Config, PatchGeom, config0, config1, config2, cube_patch,  # noqa: F401
                   puncture_state, uneven_level)
from .level import (BC_COPY_LAYER, BC_FROZEN_GHOSTS, BC_NONE, BC_SOMMERFELD, LIST_COR, LIST_OLD, LIST_PRE,  # noqa: F401
                    LIST_RES, LIST_RHS, LIST_STATE, MODE_BATCHED, MODE_GRAPH, MODE_PLAIN, RES_EVERY_STEP,
                    RES_LAST_STEP, RES_NONE, Level)
from .amr import Hierarchy, nested_box  # noqa: F401,E402
