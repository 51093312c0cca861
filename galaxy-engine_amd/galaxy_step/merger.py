"""ctypes mirror of include/ge_merger.h: UpdateNeighbors + mergerSolver on a
Context created with render=True."""
import ctypes as C

from .lib import GE_OK, GeError, load_library

MERGER_EXPORTS = ["ge_default_merger_params", "ge_merger_cell_size", "ge_merge"]


class MergerParams(C.Structure):
    _fields_ = [("density_radius", C.c_float), ("cell_size", C.c_float), ("texture_half_size", C.c_float),
                ("size_multiplier", C.c_float), ("reserved", C.c_uint32 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_ready = False


def load_merger_library():
    global _ready
    lib = load_library()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge_default_merger_params": (None, [P(MergerParams)]),
        "ge_merger_cell_size": (C.c_int, [vp, P(MergerParams), P(C.c_float)]),
        "ge_merge": (C.c_int, [vp, P(MergerParams), P(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ready = True
    return lib


def default_merger_params(**kw):
    lib = load_merger_library()
    m = MergerParams()
    lib.ge_default_merger_params(C.byref(m))
    for k, v in kw.items():
        if not hasattr(m, k):
            raise KeyError(k)
        setattr(m, k, v)
    return m


def _check(ctx, rc):
    if rc != GE_OK:
        raise GeError(rc, ctx.lib.ge_last_error(ctx.h).decode())


def cell_size(ctx, m=None):
    lib = load_merger_library()
    m = m if m is not None else default_merger_params()
    v = C.c_float()
    _check(ctx, lib.ge_merger_cell_size(ctx.h, C.byref(m), C.byref(v)))
    return v.value


def merge(ctx, m=None):
    """Runs the merger; returns the number of particles merged away."""
    lib = load_merger_library()
    m = m if m is not None else default_merger_params()
    r = C.c_uint32()
    _check(ctx, lib.ge_merge(ctx.h, C.byref(m), C.byref(r)))
    return r.value
