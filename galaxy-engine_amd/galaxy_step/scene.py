"""ctypes mirror of include/ge_scene.h: Galaxy Engine save files (version 190)
to and from device contexts (SaveSystem, GE/src/UX/saveSystem.cpp:3-614,
GE/include/UX/saveSystem.h:30-611)."""
import ctypes as C

from .lib import GE_OK, GeError, StepParams, load_library
from .lib3d import Step3Params, load_library3
from .constraints import ConstraintParams, load_constraint_library
from .thermal import ThermalParams, load_thermal_library

SCENE_EXPORTS = [
    "ge_scene_load", "ge_scene_new", "ge_scene_free", "ge_scene_last_error", "ge_scene_get_info",
    "ge_scene_value", "ge_scene_set_value", "ge_scene_records", "ge_scene_params", "ge_scene_params3",
    "ge_scene_set_params", "ge_scene_set_params3", "ge_scene_upload", "ge_scene_upload3", "ge_scene_download",
    "ge_scene_download3", "ge_scene_save", "ge_scene_thermal_params", "ge_scene_set_thermal_params",
    "ge_scene_constraint_params", "ge_scene_set_constraint_params",
]


class SceneInfo(C.Structure):
    _fields_ = [("is3d", C.c_int32), ("version", C.c_uint32), ("count", C.c_uint32),
                ("record_bytes", C.c_uint32), ("global_id", C.c_uint32), ("global_shape_id", C.c_uint32),
                ("global_wall_id", C.c_uint32), ("reserved", C.c_uint32), ("tail_bytes", C.c_uint64)]


_ready = False


def load_scene_library():
    global _ready
    lib = load_library()
    load_library3()
    load_thermal_library()
    load_constraint_library()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge_scene_load": (C.c_int, [C.c_char_p, P(vp)]),
        "ge_scene_new": (C.c_int, [C.c_int32, P(vp)]),
        "ge_scene_free": (None, [vp]),
        "ge_scene_last_error": (C.c_char_p, [vp]),
        "ge_scene_get_info": (C.c_int, [vp, P(SceneInfo)]),
        "ge_scene_value": (C.c_char_p, [vp, C.c_char_p]),
        "ge_scene_set_value": (C.c_int, [vp, C.c_char_p, C.c_char_p]),
        "ge_scene_records": (C.c_int, [vp, P(vp), P(C.c_uint64)]),
        "ge_scene_params": (C.c_int, [vp, P(StepParams)]),
        "ge_scene_params3": (C.c_int, [vp, P(Step3Params)]),
        "ge_scene_set_params": (C.c_int, [vp, P(StepParams)]),
        "ge_scene_set_params3": (C.c_int, [vp, P(Step3Params)]),
        "ge_scene_upload": (C.c_int, [vp, vp]),
        "ge_scene_upload3": (C.c_int, [vp, vp]),
        "ge_scene_download": (C.c_int, [vp, vp]),
        "ge_scene_download3": (C.c_int, [vp, vp]),
        "ge_scene_save": (C.c_int, [vp, C.c_char_p]),
        "ge_scene_thermal_params": (C.c_int, [vp, P(ThermalParams)]),
        "ge_scene_set_thermal_params": (C.c_int, [vp, P(ThermalParams)]),
        "ge_scene_constraint_params": (C.c_int, [vp, P(ConstraintParams), P(C.c_int32)]),
        "ge_scene_set_constraint_params": (C.c_int, [vp, P(ConstraintParams), C.c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ready = True
    return lib


class Scene:
    """One scene file in host memory (ge_scene)."""

    def __init__(self, handle):
        self.lib = load_scene_library()
        self.h = handle

    @classmethod
    def load(cls, path):
        lib = load_scene_library()
        h = C.c_void_p()
        rc = lib.ge_scene_load(str(path).encode(), C.byref(h))
        if rc != GE_OK:
            raise GeError(rc, lib.ge_scene_last_error(None).decode())
        return cls(h)

    @classmethod
    def new(cls, is3d=False):
        lib = load_scene_library()
        h = C.c_void_p()
        rc = lib.ge_scene_new(1 if is3d else 0, C.byref(h))
        if rc != GE_OK:
            raise GeError(rc, "ge_scene_new failed")
        return cls(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ge_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != GE_OK:
            raise GeError(rc, self.lib.ge_scene_last_error(self.h).decode())

    def info(self):
        i = SceneInfo()
        self._check(self.lib.ge_scene_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in SceneInfo._fields_ if f != "reserved"}

    def value(self, key):
        v = self.lib.ge_scene_value(self.h, key.encode())
        return None if v is None else v.decode()

    def set_value(self, key, value):
        self._check(self.lib.ge_scene_set_value(self.h, key.encode(), str(value).encode()))

    def records(self):
        """The packed particle records as bytes (count * record_bytes)."""
        p = C.c_void_p()
        n = C.c_uint64()
        self._check(self.lib.ge_scene_records(self.h, C.byref(p), C.byref(n)))
        return C.string_at(p, n.value) if n.value else b""

    def params(self):
        p = StepParams()
        self._check(self.lib.ge_scene_params(self.h, C.byref(p)))
        return p

    def params3(self):
        p = Step3Params()
        self._check(self.lib.ge_scene_params3(self.h, C.byref(p)))
        return p

    def thermal_params(self):
        t = ThermalParams()
        self._check(self.lib.ge_scene_thermal_params(self.h, C.byref(t)))
        return t

    def constraint_params(self):
        """(ConstraintParams, ParticleConstraints enabled)."""
        c = ConstraintParams()
        on = C.c_int32()
        self._check(self.lib.ge_scene_constraint_params(self.h, C.byref(c), C.byref(on)))
        return c, bool(on.value)

    def set_constraint_params(self, c, enabled):
        self._check(self.lib.ge_scene_set_constraint_params(self.h, C.byref(c), 1 if enabled else 0))

    def set_thermal_params(self, t):
        self._check(self.lib.ge_scene_set_thermal_params(self.h, C.byref(t)))

    def set_params(self, p):
        fn = self.lib.ge_scene_set_params3 if isinstance(p, Step3Params) else self.lib.ge_scene_set_params
        self._check(fn(self.h, C.byref(p)))

    def upload(self, ctx):
        """Particles into a Context (2D) or Context3D."""
        is3d = type(ctx).__name__ == "Context3D"
        fn = self.lib.ge_scene_upload3 if is3d else self.lib.ge_scene_upload
        self._check(fn(self.h, ctx.h))

    def download(self, ctx):
        is3d = type(ctx).__name__ == "Context3D"
        fn = self.lib.ge_scene_download3 if is3d else self.lib.ge_scene_download
        self._check(fn(self.h, ctx.h))

    def save(self, path):
        self._check(self.lib.ge_scene_save(self.h, str(path).encode()))

