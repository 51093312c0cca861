"""Python host-side binding of libge_physics.so (the gfx950 physics step).

The product is the C ABI in include/ge_physics.h; this package is a thin
ctypes layer used by the tests, the smoke entry and bench.py.  It never falls
back to a CPU path: loading fails loudly when the shared library is missing.
"""
from .lib import (GE_FLAG_DARK, GE_FLAG_DRAWN, GE_FLAG_PINNED, GE_FLAG_SPH,
                  GeError, Context, StepParams, default_params, library_path,
                  load_library, exported_symbols)
from .lib3d import Context3D, Step3Params, default_params3, load_library3

__all__ = ["GE_FLAG_DARK", "GE_FLAG_DRAWN", "GE_FLAG_PINNED", "GE_FLAG_SPH",
           "GeError", "Context", "StepParams", "default_params", "library_path",
           "load_library", "exported_symbols", "Context3D", "Step3Params", "default_params3",
           "load_library3"]
