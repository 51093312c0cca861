"""ctypes mirror of include/ge_thermal.h: the temperature coupling of the 2D
frame (node temperatures, the walk's heat transfer, temperatureCalculation
with the SPH materials) on a Context created with temperature=True."""
import ctypes as C

import numpy as np

from .lib import GE_OK, GeError, load_library

THERMAL_EXPORTS = [
    "ge_default_thermal_params", "ge_thermal_enabled", "ge_default_materials", "ge_set_thermal", "ge_thermal_upload",
    "ge_thermal_download", "ge_temperature_calculation", "ge_get_node_temps",
]
THERMAL_FIELDS = ["temp", "label", "ke", "prev_ke"]
_TDT = {"temp": np.float32, "label": np.uint32, "ke": np.float32, "prev_ke": np.float32}
MATERIAL_FLOATS = ("mass_mult", "rest_dens", "stiff", "visc", "cohesion",
                   "hot_point", "hot_rest_dens", "hot_mass_mult", "hot_stiff", "hot_visc", "hot_cohesion",
                   "cold_point", "cold_rest_dens", "cold_mass_mult", "cold_stiff", "cold_visc", "cold_cohesion",
                   "heat_conductivity", "constraint_resistance", "constraint_plastic_point",
                   "constraint_plastic_point_mult", "constraint_stiffness")


class Material(C.Structure):
    _fields_ = [("id", C.c_uint32)] + [(f, C.c_float) for f in MATERIAL_FLOATS] + [("is_plastic", C.c_int32)]


class ThermalParams(C.Structure):
    _fields_ = [("heat_conductivity", C.c_float), ("ambient_temp", C.c_float), ("ambient_heat_rate", C.c_float),
                ("base_mass", C.c_float), ("reserved", C.c_uint32 * 4)]


class ThermalFields(C.Structure):
    _fields_ = [("temp", C.c_void_p), ("label", C.c_void_p), ("ke", C.c_void_p), ("prev_ke", C.c_void_p)]


_ready = False


def load_thermal_library():
    global _ready
    lib = load_library()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge_default_thermal_params": (None, [P(ThermalParams)]),
        "ge_default_materials": (C.c_int, [P(Material), C.c_uint32, P(C.c_uint32)]),
        "ge_thermal_enabled": (C.c_int, [vp, P(C.c_int32)]),
        "ge_set_thermal": (C.c_int, [vp, P(ThermalParams), P(Material), C.c_uint32]),
        "ge_thermal_upload": (C.c_int, [vp, P(ThermalFields)]),
        "ge_thermal_download": (C.c_int, [vp, P(ThermalFields)]),
        "ge_temperature_calculation": (C.c_int, [vp, vp]),
        "ge_get_node_temps": (C.c_int, [vp, P(C.c_float), C.c_uint32]),
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


def default_thermal_params(**kw):
    lib = load_thermal_library()
    t = ThermalParams()
    lib.ge_default_thermal_params(C.byref(t))
    for k, v in kw.items():
        if not hasattr(t, k):
            raise KeyError(k)
        setattr(t, k, v)
    return t


def default_materials():
    lib = load_thermal_library()
    arr = (Material * 64)()
    n = C.c_uint32()
    rc = lib.ge_default_materials(arr, 64, C.byref(n))
    if rc != GE_OK:
        raise GeError(rc, "ge_default_materials failed")
    return [arr[k] for k in range(n.value)]


def set_thermal(ctx, params=None, materials=None):
    lib = load_thermal_library()
    params = params if params is not None else default_thermal_params()
    materials = default_materials() if materials is None else list(materials)
    arr = (Material * max(len(materials), 1))(*materials)
    _check(ctx, lib.ge_set_thermal(ctx.h, C.byref(params), arr, len(materials)))


def thermal_upload(ctx, fields):
    lib = load_thermal_library()
    n = C.c_uint32()
    _check(ctx, lib.ge_count(ctx.h, C.byref(n)))
    tf = ThermalFields()
    hold = []
    for f in THERMAL_FIELDS:
        if fields.get(f) is not None:
            a = np.ascontiguousarray(fields[f], dtype=_TDT[f])
            assert len(a) == n.value, f
            hold.append(a)
            setattr(tf, f, a.ctypes.data)
    _check(ctx, lib.ge_thermal_upload(ctx.h, C.byref(tf)))


def thermal_download(ctx, fields=None):
    lib = load_thermal_library()
    n = C.c_uint32()
    _check(ctx, lib.ge_count(ctx.h, C.byref(n)))
    out = {f: np.zeros(n.value, _TDT[f]) for f in (fields or THERMAL_FIELDS)}
    tf = ThermalFields()
    for f, a in out.items():
        setattr(tf, f, a.ctypes.data)
    _check(ctx, lib.ge_thermal_download(ctx.h, C.byref(tf)))
    return out


def temperature_calculation(ctx, p):
    lib = load_thermal_library()
    _check(ctx, lib.ge_temperature_calculation(ctx.h, C.byref(p)))


def node_temps(ctx):
    lib = load_thermal_library()
    cnt = C.c_uint32()
    _check(ctx, lib.ge_get_node_count(ctx.h, C.byref(cnt)))
    out = np.zeros(max(cnt.value, 1), np.float32)
    _check(ctx, lib.ge_get_node_temps(ctx.h, out.ctypes.data_as(C.POINTER(C.c_float)), cnt.value))
    return out[:cnt.value]
