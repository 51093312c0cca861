"""ctypes mirror of include/ge_constraints.h: the particle constraints of the
2D frame on a Context created with max_constraints > 0."""
import ctypes as C

import numpy as np

from .lib import GE_OK, GeError, load_library

CONSTRAINT_EXPORTS = [
    "ge_default_constraint_params", "ge_constraints_capacity", "ge_set_constraint_params", "ge_constraints_upload", "ge_constraints_download",
    "ge_constraints_clear", "ge_constraints_solve", "ge_constraints_order", "ge_create_constraints",
]
# ParticleConstraint, the 36-byte record of the save files
CONSTRAINT_DTYPE = np.dtype([("id1", "<u4"), ("id2", "<u4"), ("rest_length", "<f4"), ("original_length", "<f4"),
                             ("stiffness", "<f4"), ("resistance", "<f4"), ("displacement", "<f4"),
                             ("plasticity_point", "<f4"), ("is_broken", "u1"), ("is_plastic", "u1"),
                             ("pad", "u1", (2,))])


class Constraint(C.Structure):
    _fields_ = [("id1", C.c_uint32), ("id2", C.c_uint32), ("rest_length", C.c_float),
                ("original_length", C.c_float), ("stiffness", C.c_float), ("resistance", C.c_float),
                ("displacement", C.c_float), ("plasticity_point", C.c_float), ("is_broken", C.c_uint8),
                ("is_plastic", C.c_uint8), ("pad", C.c_uint8 * 2)]


class ConstraintParams(C.Structure):
    _fields_ = [("stiffness_mult", C.c_float), ("resistance_mult", C.c_float), ("unbreakable", C.c_int32),
                ("forces", C.c_int32), ("damping", C.c_float), ("correction_ratio", C.c_float),
                ("reserved", C.c_uint32 * 2)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_ready = False


def load_constraint_library():
    global _ready
    lib = load_library()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    u32p = P(C.c_uint32)
    sig = {
        "ge_default_constraint_params": (None, [P(ConstraintParams)]),
        "ge_constraints_capacity": (C.c_int, [vp, u32p]),
        "ge_set_constraint_params": (C.c_int, [vp, P(ConstraintParams)]),
        "ge_constraints_upload": (C.c_int, [vp, vp, C.c_uint32]),
        "ge_constraints_download": (C.c_int, [vp, vp, C.c_uint32, u32p]),
        "ge_constraints_clear": (C.c_int, [vp]),
        "ge_constraints_solve": (C.c_int, [vp, vp]),
        "ge_constraints_order": (C.c_int, [vp, u32p, C.c_uint32, u32p, C.c_uint32, u32p]),
        "ge_create_constraints": (C.c_int, [vp, vp, u32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ready = True
    return lib


def _check(ctx, rc):
    if rc != GE_OK:
        raise GeError(rc, ctx.lib.ge_last_error(ctx.h).decode())


def default_constraint_params(**kw):
    lib = load_constraint_library()
    p = ConstraintParams()
    lib.ge_default_constraint_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def set_params(ctx, p):
    _check(ctx, load_constraint_library().ge_set_constraint_params(ctx.h, C.byref(p)))


def upload(ctx, cons):
    lib = load_constraint_library()
    a = np.ascontiguousarray(cons, dtype=CONSTRAINT_DTYPE)
    _check(ctx, lib.ge_constraints_upload(ctx.h, a.ctypes.data if len(a) else None, len(a)))


def download(ctx):
    lib = load_constraint_library()
    cnt = C.c_uint32()
    _check(ctx, lib.ge_constraints_download(ctx.h, None, 0, C.byref(cnt)))
    out = np.zeros(cnt.value, CONSTRAINT_DTYPE)
    if cnt.value:
        c2 = C.c_uint32()
        _check(ctx, lib.ge_constraints_download(ctx.h, out.ctypes.data, cnt.value, C.byref(c2)))
    return out


def clear(ctx):
    _check(ctx, load_constraint_library().ge_constraints_clear(ctx.h))


def solve(ctx, p):
    _check(ctx, load_constraint_library().ge_constraints_solve(ctx.h, C.byref(p)))


def create(ctx, params):
    """ge_create_constraints: the solidified particles' constraints on the
    current state; returns the number appended."""
    n = C.c_uint32()
    _check(ctx, load_constraint_library().ge_create_constraints(ctx.h, C.byref(params), C.byref(n)))
    return n.value


def colors(ctx, count):
    """The colour of each of the current `count` constraints (from the substep order)."""
    lib = load_constraint_library()
    order = np.zeros(max(count, 1), np.uint32)
    start = np.zeros(72, np.uint32)
    nc = C.c_uint32()
    u32p = C.POINTER(C.c_uint32)
    _check(ctx, lib.ge_constraints_order(ctx.h, order.ctypes.data_as(u32p), count, start.ctypes.data_as(u32p), 72,
                                         C.byref(nc)))
    col = np.full(count, 0xFFFFFFFF, np.uint32)
    for q in range(nc.value):
        col[order[start[q]:start[q + 1]]] = q
    return col, nc.value
