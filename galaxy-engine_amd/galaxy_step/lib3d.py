"""ctypes mirror of include/ge_physics3d.h.

Context3D wraps one ge3_ctx: upload/download of SoA numpy arrays, ge3_step
and the stage entry points that replace the reference's Morton3D / Octree /
Physics3D / NeighborSearch3DV2AVX2 / SPH3D member functions
(GE/src/Physics/morton.cpp:64-138, quadtree.cpp:109-292,
physics3D.cpp:131-470, 740-1168, GE/include/Particles/neighborSearch.h:
978-1193, GE/src/Physics/SPH3D.cpp:98-366).
"""
import ctypes as C

import numpy as np

from .lib import GE_OK, GeError, load_library

EXPORTS3 = [
    "ge3_default_config", "ge3_default_params", "ge3_create", "ge3_destroy", "ge3_last_error", "ge3_sync",
    "ge3_count", "ge3_upload", "ge3_download", "ge3_bounding_box", "ge3_compute_morton_keys",
    "ge3_get_morton_keys", "ge3_sort_by_morton_key", "ge3_get_permutation", "ge3_build_octree",
    "ge3_get_node_count", "ge3_get_nodes", "ge3_compute_gravity", "ge3_integrate_start", "ge3_integrate_end",
    "ge3_step", "ge3_direct_fp64", "ge3_get_interaction_counts", "ge3_sph_new_grid", "ge3_sph_pcisph",
    "ge3_get_neighbors", "ge3_neighbor_counts",
]

FIELDS3_F = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "pvx", "pvy", "pvz", "mass"]
FIELDS3_U = ["flags", "id"]
FIELDS3_SPH = ["sph_mass", "rest_dens", "stiff", "visc", "cohesion", "dens", "press", "press_tmp", "pred_dens",
               "ppx", "ppy", "ppz", "pfx", "pfy", "pfz"]
PARTICLE3_ORDER = FIELDS3_F + FIELDS3_U + FIELDS3_SPH
GRAVITY3_FIELDS = FIELDS3_F + FIELDS3_U


class Ge3Config(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("device", C.c_int32), ("enable_sph", C.c_int32),
                ("enable_render", C.c_int32), ("reserved", C.c_uint32 * 4)]


class Ge3Particles(C.Structure):
    _fields_ = [(f, C.POINTER(C.c_uint32) if f in FIELDS3_U else C.POINTER(C.c_float)) for f in PARTICLE3_ORDER]


class Step3Params(C.Structure):
    _fields_ = [("G", C.c_float), ("theta", C.c_float), ("softening", C.c_float), ("dt_start", C.c_float),
                ("dt", C.c_float), ("naive", C.c_int32), ("direct_below", C.c_int32), ("periodic", C.c_int32),
                ("domain_w", C.c_float), ("domain_h", C.c_float), ("domain_d", C.c_float), ("sph", C.c_int32),
                ("temperature", C.c_int32), ("compensated", C.c_int32), ("leaf_capacity", C.c_uint32),
                ("walk_block", C.c_uint32), ("sph_h", C.c_float), ("sph_cell", C.c_float),
                ("sph_unit_mass", C.c_float), ("viscosity", C.c_float), ("cohesion_coef", C.c_float),
                ("delta", C.c_float), ("stiff_mult", C.c_float), ("sph_max_vel", C.c_float),
                ("sph_ground", C.c_int32), ("vertical_gravity", C.c_float), ("boundary_friction", C.c_float),
                ("reserved", C.c_uint32 * 4)]


class Ge3Node(C.Structure):
    _fields_ = [("posx", C.c_float), ("posy", C.c_float), ("posz", C.c_float), ("size", C.c_float),
                ("comx", C.c_float), ("comy", C.c_float), ("comz", C.c_float), ("mass", C.c_float),
                ("start", C.c_uint32), ("end", C.c_uint32), ("next", C.c_uint32), ("child", C.c_uint32 * 8)]


NODE3_DTYPE = np.dtype([("posx", "<f4"), ("posy", "<f4"), ("posz", "<f4"), ("size", "<f4"),
                        ("comx", "<f4"), ("comy", "<f4"), ("comz", "<f4"), ("mass", "<f4"),
                        ("start", "<u4"), ("end", "<u4"), ("next", "<u4"), ("child", "<u4", (8,))])
assert NODE3_DTYPE.itemsize == C.sizeof(Ge3Node)

_ready = False


def load_library3():
    """The 3D entry points of libge_physics.so (raises if the library is missing)."""
    global _ready
    lib = load_library()
    if _ready:
        return lib
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge3_default_config": (None, [P(Ge3Config)]),
        "ge3_default_params": (None, [P(Step3Params)]),
        "ge3_create": (C.c_int, [P(vp), P(Ge3Config)]),
        "ge3_destroy": (C.c_int, [vp]),
        "ge3_last_error": (C.c_char_p, [vp]),
        "ge3_sync": (C.c_int, [vp]),
        "ge3_count": (C.c_int, [vp, P(C.c_uint32)]),
        "ge3_upload": (C.c_int, [vp, C.c_uint32, P(Ge3Particles)]),
        "ge3_download": (C.c_int, [vp, P(Ge3Particles)]),
        "ge3_bounding_box": (C.c_int, [vp, P(C.c_float)]),
        "ge3_compute_morton_keys": (C.c_int, [vp, P(C.c_float)]),
        "ge3_get_morton_keys": (C.c_int, [vp, P(C.c_uint64)]),
        "ge3_sort_by_morton_key": (C.c_int, [vp]),
        "ge3_get_permutation": (C.c_int, [vp, P(C.c_uint32)]),
        "ge3_build_octree": (C.c_int, [vp, P(C.c_float)]),
        "ge3_get_node_count": (C.c_int, [vp, P(C.c_uint32)]),
        "ge3_get_nodes": (C.c_int, [vp, vp, C.c_uint32]),
        "ge3_compute_gravity": (C.c_int, [vp, P(Step3Params)]),
        "ge3_integrate_start": (C.c_int, [vp, P(Step3Params)]),
        "ge3_integrate_end": (C.c_int, [vp, P(Step3Params)]),
        "ge3_step": (C.c_int, [vp, P(Step3Params), C.c_int32]),
        "ge3_direct_fp64": (C.c_int, [vp, C.c_double, C.c_double, C.c_uint32, P(C.c_uint32), P(C.c_double),
                                      P(C.c_double), P(C.c_double)]),
        "ge3_get_interaction_counts": (C.c_int, [vp, P(C.c_uint64)]),
        "ge3_sph_new_grid": (C.c_int, [vp, P(Step3Params)]),
        "ge3_sph_pcisph": (C.c_int, [vp, P(Step3Params)]),
        "ge3_get_neighbors": (C.c_int, [vp, C.c_int32, P(C.c_int64), P(C.c_uint32), C.c_int64, P(C.c_int64)]),
        "ge3_neighbor_counts": (C.c_int, [vp, P(Step3Params), P(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ready = True
    return lib


def default_params3(**kw):
    lib = load_library3()
    p = Step3Params()
    lib.ge3_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Context3D:
    """One 3D device context (ge3_ctx)."""

    def __init__(self, capacity, device=0, sph=False, render=False):
        self.lib = load_library3()
        cfg = Ge3Config()
        self.lib.ge3_default_config(C.byref(cfg))
        cfg.capacity = int(capacity)
        cfg.device = int(device)
        cfg.enable_sph = 1 if sph else 0
        cfg.enable_render = 1 if render else 0
        self.sph = bool(sph)
        self.render = bool(render)
        h = C.c_void_p()
        rc = self.lib.ge3_create(C.byref(h), C.byref(cfg))
        if rc != GE_OK:
            raise GeError(rc, "ge3_create failed")
        self.h = h
        self.capacity = int(capacity)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ge3_destroy(self.h)
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
            raise GeError(rc, self.lib.ge3_last_error(self.h).decode())

    def upload(self, state):
        n = len(state["x"])
        gp = Ge3Particles()
        keep = []
        for f in PARTICLE3_ORDER:
            if f in state and state[f] is not None:
                a = np.ascontiguousarray(state[f], dtype=np.uint32 if f in FIELDS3_U else np.float32)
                assert len(a) == n, f
                keep.append(a)
                setattr(gp, f, _uptr(a) if f in FIELDS3_U else _fptr(a))
        self._check(self.lib.ge3_upload(self.h, n, C.byref(gp)))

    def count(self):
        n = C.c_uint32()
        self._check(self.lib.ge3_count(self.h, C.byref(n)))
        return n.value

    def download(self, fields=None):
        n = self.count()
        fields = fields or (PARTICLE3_ORDER if self.sph else GRAVITY3_FIELDS)
        out = {}
        gp = Ge3Particles()
        for f in fields:
            a = np.empty(n, dtype=np.uint32 if f in FIELDS3_U else np.float32)
            out[f] = a
            setattr(gp, f, _uptr(a) if f in FIELDS3_U else _fptr(a))
        self._check(self.lib.ge3_download(self.h, C.byref(gp)))
        return out

    def sync(self):
        self._check(self.lib.ge3_sync(self.h))

    def step(self, params, frames=1):
        self._check(self.lib.ge3_step(self.h, C.byref(params), int(frames)))

    def bounding_box(self):
        bb = (C.c_float * 4)()
        self._check(self.lib.ge3_bounding_box(self.h, bb))
        return np.array(list(bb), dtype=np.float32)

    def compute_morton_keys(self, bb):
        self._check(self.lib.ge3_compute_morton_keys(self.h, (C.c_float * 4)(*[float(v) for v in bb])))

    def keys(self):
        a = np.empty(self.count(), np.uint64)
        self._check(self.lib.ge3_get_morton_keys(self.h, a.ctypes.data_as(C.POINTER(C.c_uint64))))
        return a

    def sort_by_morton_key(self):
        self._check(self.lib.ge3_sort_by_morton_key(self.h))

    def perm(self):
        a = np.empty(self.count(), np.uint32)
        self._check(self.lib.ge3_get_permutation(self.h, _uptr(a)))
        return a

    def build_octree(self, bb):
        self._check(self.lib.ge3_build_octree(self.h, (C.c_float * 4)(*[float(v) for v in bb])))

    def node_count(self):
        c = C.c_uint32()
        self._check(self.lib.ge3_get_node_count(self.h, C.byref(c)))
        return c.value

    def nodes(self):
        nn = self.node_count()
        a = np.zeros(nn, dtype=NODE3_DTYPE)
        self._check(self.lib.ge3_get_nodes(self.h, C.c_void_p(a.ctypes.data), nn))
        return a

    def compute_gravity(self, params):
        self._check(self.lib.ge3_compute_gravity(self.h, C.byref(params)))

    def integrate_start(self, params):
        self._check(self.lib.ge3_integrate_start(self.h, C.byref(params)))

    def integrate_end(self, params):
        self._check(self.lib.ge3_integrate_end(self.h, C.byref(params)))

    def direct_fp64(self, targets, G, softening):
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        out = [np.zeros(len(t), np.float64) for _ in range(3)]
        dp = C.POINTER(C.c_double)
        self._check(self.lib.ge3_direct_fp64(self.h, G, softening, len(t), _uptr(t),
                                             *[o.ctypes.data_as(dp) for o in out]))
        return out

    def interaction_counts(self):
        a = (C.c_uint64 * 2)()
        self._check(self.lib.ge3_get_interaction_counts(self.h, a))
        return a[0], a[1]

    def sph_new_grid(self, params):
        self._check(self.lib.ge3_sph_new_grid(self.h, C.byref(params)))

    def sph_pcisph(self, params):
        self._check(self.lib.ge3_sph_pcisph(self.h, C.byref(params)))

    def neighbors(self, which=0):
        """Candidate sets of the current grid at pos (0) or predPos (1): CSR (off, idx)."""
        n = self.count()
        off = np.zeros(n + 1, np.int64)
        total = C.c_int64()
        op = off.ctypes.data_as(C.POINTER(C.c_int64))
        self._check(self.lib.ge3_get_neighbors(self.h, int(which), op, None, 0, C.byref(total)))
        idx = np.zeros(max(total.value, 1), np.uint32)
        self._check(self.lib.ge3_get_neighbors(self.h, int(which), op, _uptr(idx), total.value, C.byref(total)))
        return off, idx[:total.value]

    def neighbor_counts(self, params):
        a = np.zeros(self.count(), np.uint32)
        self._check(self.lib.ge3_neighbor_counts(self.h, C.byref(params), _uptr(a)))
        return a
