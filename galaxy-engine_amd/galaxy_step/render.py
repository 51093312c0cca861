"""ctypes mirror of include/ge_render.h: render fields, the post-step
consumers (sizeByDensity and the colour modes), PLY export and the galaxy
spawner, on Context / Context3D objects."""
import ctypes as C

import numpy as np

from .lib import GE_OK, GeError, load_library
from .lib3d import load_library3

RENDER_EXPORTS = [
    "ge_render_enabled", "ge3_render_enabled", "ge_render_upload", "ge_render_download", "ge3_render_upload", "ge3_render_download",
    "ge_default_visual_params", "ge_visuals", "ge3_visuals", "ge_ply_vertices", "ge3_ply_vertices",
    "ge_ply_write", "ge3_ply_write", "ge_ply_format", "ge_default_galaxy_params", "ge_spawn_galaxy",
]
RENDER_FIELDS = ["color", "pcolor", "scolor", "size", "neighbors", "previous_size"]
_RDT = {"color": np.uint32, "pcolor": np.uint32, "scolor": np.uint32, "size": np.float32,
        "neighbors": np.int32, "previous_size": np.float32}


class RenderFields(C.Structure):
    _fields_ = [("color", C.c_void_p), ("pcolor", C.c_void_p), ("scolor", C.c_void_p), ("size", C.c_void_p),
                ("neighbors", C.c_void_p), ("previous_size", C.c_void_p)]


class VisualParams(C.Structure):
    _fields_ = [("count_neighbors", C.c_int32), ("search", C.c_int32), ("cell", C.c_float),
                ("radius", C.c_float), ("density_size", C.c_int32), ("force_size", C.c_int32),
                ("min_size", C.c_float), ("max_size", C.c_float), ("size_acc", C.c_float),
                ("size_max_neighbors", C.c_int32), ("size_multiplier", C.c_float),
                ("solid_color", C.c_int32), ("density_color", C.c_int32), ("force_color", C.c_int32),
                ("p_color", C.c_uint32), ("s_color", C.c_uint32), ("color_max_neighbors", C.c_int32),
                ("min_color_acc", C.c_float), ("max_color_acc", C.c_float), ("reserved", C.c_uint32 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PlyVertex(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("r", C.c_uint8), ("g", C.c_uint8),
                ("b", C.c_uint8), ("pad", C.c_uint8), ("radius", C.c_float)]


PLY_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1"), ("g", "u1"), ("b", "u1"),
                      ("pad", "u1"), ("radius", "<f4")])


class GalaxyParams(C.Structure):
    _fields_ = [("center_x", C.c_float), ("center_y", C.c_float), ("extra_vx", C.c_float),
                ("extra_vy", C.c_float), ("visible", C.c_uint32), ("dark", C.c_uint32),
                ("outer_radius", C.c_float), ("scale_length", C.c_float), ("outer_radius_dm", C.c_float),
                ("radius_core_dm", C.c_float), ("mass_visible", C.c_float), ("mass_dark", C.c_float),
                ("mass_scatter", C.c_float), ("first_id", C.c_uint32), ("reserved", C.c_uint32 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_ready = False


def load_render_library():
    global _ready
    lib = load_library()
    load_library3()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge_render_enabled": (C.c_int, [vp, P(C.c_int32)]),
        "ge3_render_enabled": (C.c_int, [vp, P(C.c_int32)]),
        "ge_render_upload": (C.c_int, [vp, P(RenderFields)]),
        "ge_render_download": (C.c_int, [vp, P(RenderFields)]),
        "ge3_render_upload": (C.c_int, [vp, P(RenderFields)]),
        "ge3_render_download": (C.c_int, [vp, P(RenderFields)]),
        "ge_default_visual_params": (None, [P(VisualParams)]),
        "ge_visuals": (C.c_int, [vp, P(VisualParams)]),
        "ge3_visuals": (C.c_int, [vp, P(VisualParams)]),
        "ge_ply_vertices": (C.c_int, [vp, C.c_float, C.c_float, vp, C.c_uint32, P(C.c_uint32)]),
        "ge3_ply_vertices": (C.c_int, [vp, vp, C.c_uint32, P(C.c_uint32)]),
        "ge_ply_write": (C.c_int, [vp, C.c_float, C.c_float, C.c_char_p]),
        "ge3_ply_write": (C.c_int, [vp, C.c_char_p]),
        "ge_ply_format": (C.c_int, [vp, C.c_uint32, vp, C.c_uint64, P(C.c_uint64)]),
        "ge_default_galaxy_params": (None, [P(GalaxyParams)]),
        "ge_spawn_galaxy": (C.c_int, [vp, P(GalaxyParams), C.c_uint64, vp, P(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ready = True
    return lib


def default_visual_params(**kw):
    lib = load_render_library()
    v = VisualParams()
    lib.ge_default_visual_params(C.byref(v))
    for k, val in kw.items():
        if not hasattr(v, k):
            raise KeyError(k)
        setattr(v, k, val)
    return v


def default_galaxy_params(**kw):
    lib = load_render_library()
    g = GalaxyParams()
    lib.ge_default_galaxy_params(C.byref(g))
    for k, val in kw.items():
        if not hasattr(g, k):
            raise KeyError(k)
        setattr(g, k, val)
    return g


def rgba(r, g, b, a):
    """raylib Color {r, g, b, a} as the packed RGBA8 word of the render fields."""
    return int(r) | int(g) << 8 | int(b) << 16 | int(a) << 24


def _is3d(ctx):
    return type(ctx).__name__ == "Context3D"


def _check(ctx, lib, rc):
    if rc != GE_OK:
        err = lib.ge3_last_error(ctx.h) if _is3d(ctx) else lib.ge_last_error(ctx.h)
        raise GeError(rc, err.decode())


def _count(ctx):
    lib = ctx.lib
    n = C.c_uint32()
    rc = lib.ge3_count(ctx.h, C.byref(n)) if _is3d(ctx) else lib.ge_count(ctx.h, C.byref(n))
    _check(ctx, lib, rc)
    return n.value


def render_upload(ctx, fields):
    """fields: dict with any of color, pcolor, scolor, size, neighbors (length = particle count)."""
    lib = load_render_library()
    n = _count(ctx)
    rf = RenderFields()
    hold = []
    for f in RENDER_FIELDS:
        if fields.get(f) is not None:
            a = np.ascontiguousarray(fields[f], dtype=_RDT[f])
            assert len(a) == n, f
            hold.append(a)
            setattr(rf, f, a.ctypes.data)
    fn = lib.ge3_render_upload if _is3d(ctx) else lib.ge_render_upload
    _check(ctx, lib, fn(ctx.h, C.byref(rf)))


def render_download(ctx, fields=None):
    lib = load_render_library()
    n = _count(ctx)
    out = {f: np.zeros(n, _RDT[f]) for f in (fields or RENDER_FIELDS)}
    rf = RenderFields()
    for f, a in out.items():
        setattr(rf, f, a.ctypes.data)
    fn = lib.ge3_render_download if _is3d(ctx) else lib.ge_render_download
    _check(ctx, lib, fn(ctx.h, C.byref(rf)))
    return out


def visuals(ctx, v=None):
    lib = load_render_library()
    v = v if v is not None else default_visual_params()
    fn = lib.ge3_visuals if _is3d(ctx) else lib.ge_visuals
    _check(ctx, lib, fn(ctx.h, C.byref(v)))


def ply_vertices(ctx, domain_w=3840.0, domain_h=2160.0):
    """The compacted PLY vertices as a structured array (PLY_DTYPE)."""
    lib = load_render_library()
    cnt = C.c_uint32()
    if _is3d(ctx):
        _check(ctx, lib, lib.ge3_ply_vertices(ctx.h, None, 0, C.byref(cnt)))
    else:
        _check(ctx, lib, lib.ge_ply_vertices(ctx.h, domain_w, domain_h, None, 0, C.byref(cnt)))
    out = np.zeros(cnt.value, PLY_DTYPE)
    if cnt.value:
        c2 = C.c_uint32()
        if _is3d(ctx):
            _check(ctx, lib, lib.ge3_ply_vertices(ctx.h, out.ctypes.data, cnt.value, C.byref(c2)))
        else:
            _check(ctx, lib, lib.ge_ply_vertices(ctx.h, domain_w, domain_h, out.ctypes.data, cnt.value,
                                                 C.byref(c2)))
        assert c2.value == cnt.value
    return out


def ply_write(ctx, path, domain_w=3840.0, domain_h=2160.0):
    lib = load_render_library()
    if _is3d(ctx):
        _check(ctx, lib, lib.ge3_ply_write(ctx.h, str(path).encode()))
    else:
        _check(ctx, lib, lib.ge_ply_write(ctx.h, domain_w, domain_h, str(path).encode()))


def ply_format(vertices):
    """The PLY text (bytes) the export writes for a PLY_DTYPE array (host code)."""
    lib = load_render_library()
    v = np.ascontiguousarray(vertices, dtype=PLY_DTYPE)
    n = C.c_uint64()
    ptr = v.ctypes.data if len(v) else None
    rc = lib.ge_ply_format(ptr, len(v), None, 0, C.byref(n))
    if rc != GE_OK:
        raise GeError(rc, "ge_ply_format failed")
    buf = C.create_string_buffer(n.value)
    rc = lib.ge_ply_format(ptr, len(v), buf, n.value, C.byref(n))
    if rc != GE_OK:
        raise GeError(rc, "ge_ply_format failed")
    return buf.raw[:n.value]


def spawn_galaxy(ctx, g=None, seed=0, uniforms=None):
    """Appends a galaxy (ge_spawn_galaxy); returns the index of its first particle."""
    lib = load_render_library()
    g = g if g is not None else default_galaxy_params()
    first = C.c_uint32()
    hold = None
    ptr = None
    if uniforms is not None:
        hold = np.ascontiguousarray(uniforms, dtype=np.float32)
        assert len(hold) == 3 * g.visible + 5 * g.dark
        ptr = hold.ctypes.data
    _check(ctx, lib, lib.ge_spawn_galaxy(ctx.h, C.byref(g), C.c_uint64(seed), ptr, C.byref(first)))
    return first.value
