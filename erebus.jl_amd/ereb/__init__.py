"""ereb -- host side of the MI355X marker-in-cell transfer library for
Erebus.jl (ctypes over the C ABI of libereb_hip.so)."""
from .constants import (Geometry, Hydration, Materials, prod_constants, test_constants,
                        DT_LONGEST)
from .lib import (Context, EREBError, EREBModeError, FP64, MIXED, GRID_FIELDS,
                  MARKER_FIELDS, SUM_FIELDS, SYS_STOKES_DARCY, SYS_THERMAL, load, LIB_PATH)
from . import api, synth

__all__ = ["Geometry", "Hydration", "Materials", "prod_constants", "test_constants", "DT_LONGEST",
           "Context", "EREBError", "EREBModeError", "FP64", "MIXED", "GRID_FIELDS",
           "MARKER_FIELDS", "SUM_FIELDS", "SYS_STOKES_DARCY", "SYS_THERMAL", "load",
           "LIB_PATH", "api", "synth"]
