"""Python host-side binding of libge_physics.so (the gfx950 physics step).

The product is the C ABI in include/ge_physics.h; this package is a thin
ctypes layer used by the tests, the smoke entry and bench.py.  It never falls
back to a CPU path: loading fails loudly when the shared library is missing.
"""
from .lib import (GE_FLAG_DARK, GE_FLAG_DRAWN, GE_FLAG_PINNED, GE_FLAG_SPH,
                  GeError, Context, StepParams, default_params, library_path,
                  load_library, exported_symbols)

__all__ = ["GE_FLAG_DARK", "GE_FLAG_DRAWN", "GE_FLAG_PINNED", "GE_FLAG_SPH",
           "GeError", "Context", "StepParams", "default_params", "library_path",
           "load_library", "exported_symbols"]
