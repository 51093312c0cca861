"""ctypes mirror of include/ge_physics.h.

Context wraps one ge_ctx: upload/download of SoA numpy arrays, ge_step and the
stage entry points that replace the reference's Morton / Quadtree / Physics /
SPH member functions (GE/src/Physics/*.cpp).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)

GE_OK = 0
GE_FLAG_PINNED = 1
GE_FLAG_SPH = 2
GE_FLAG_DRAWN = 4
GE_FLAG_DARK = 8
GE_FLAG_UNIQUE_COLOR = 16
GE_FLAG_SOLID = 32
GE_FLAG_HOT_POINT = 64
GE_FLAG_SOLIDIFIED = 128

# every symbol include/ge_physics.h declares
EXPORTS = [
    "ge_default_config", "ge_default_params", "ge_create", "ge_destroy", "ge_last_error",
    "ge_version", "ge_upload", "ge_download", "ge_count", "ge_sync", "ge_step",
    "ge_bounding_box", "ge_compute_morton_keys", "ge_sort_by_morton_key", "ge_build_quadtree",
    "ge_compute_gravity", "ge_integrate_start", "ge_integrate_end", "ge_sph_new_grid",
    "ge_sph_pcisph", "ge_get_keys", "ge_get_perm", "ge_get_node_count", "ge_get_nodes",
    "ge_get_interactions", "ge_get_neighbors", "ge_profile_enable", "ge_profile_read",
    "ge_upload_systems", "ge_system_offsets", "ge_systems_update", "ge_download_system", "ge_poll", "ge_memory_bytes", "ge_diagnostics", "ge_diagnostics_noalloc", "ge_direct_fp64", "ge_get_interaction_counts", "ge_get_walk_stats", "ge_upload_aos", "ge_download_aos", "ge_neighbor_counts",
]
PHASES = ["integrate_bbox", "keys", "sort", "gather", "tree", "gravity", "sph_grid", "sph",
          "integrate_end", "multipole"]

FIELDS_F = ["x", "y", "vx", "vy", "ax", "ay", "pvx", "pvy", "mass"]
FIELDS_U = ["flags", "id"]
FIELDS_SPH = ["sph_mass", "rest_dens", "stiff", "visc", "cohesion", "dens", "press",
              "press_tmp", "pred_dens", "ppx", "ppy", "pfx", "pfy"]
PARTICLE_ORDER = ["x", "y", "vx", "vy", "ax", "ay", "pvx", "pvy", "mass", "flags", "id",
                  "sph_mass", "rest_dens", "stiff", "visc", "cohesion", "dens", "press",
                  "press_tmp", "pred_dens", "ppx", "ppy", "pfx", "pfy"]


class GeConfig(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("device", C.c_int32), ("enable_sph", C.c_int32),
                ("max_systems", C.c_uint32), ("enable_render", C.c_int32), ("enable_temperature", C.c_int32),
                ("max_constraints", C.c_uint32), ("precise_positions", C.c_int32)]


class GeParticles(C.Structure):
    _fields_ = [(f, C.POINTER(C.c_uint32) if f in FIELDS_U else C.POINTER(C.c_float))
                for f in PARTICLE_ORDER]


class GeDiag(C.Structure):
    _fields_ = [("mass", C.c_double), ("px", C.c_double), ("py", C.c_double), ("lz", C.c_double),
                ("kinetic", C.c_double), ("potential", C.c_double), ("comx", C.c_double),
                ("comy", C.c_double)]


class AosLayout(C.Structure):
    _fields_ = [("stride", C.c_uint32), ("off", C.c_int32 * 24)]

    @classmethod
    def from_dtype(cls, dtype):
        """Layout of a numpy structured dtype whose field names are ge_particles
        field names (others are ignored)."""
        lay = cls()
        lay.stride = dtype.itemsize
        for k, f in enumerate(PARTICLE_ORDER):
            lay.off[k] = dtype.fields[f][1] if f in dtype.names else -1
        return lay


class StepParams(C.Structure):
    _fields_ = [("G", C.c_float), ("theta", C.c_float), ("softening", C.c_float),
                ("dt_start", C.c_float), ("dt", C.c_float), ("naive", C.c_int32),
                ("direct_below", C.c_int32), ("sph", C.c_int32), ("sph_h", C.c_float),
                ("sph_cell", C.c_float), ("sph_unit_mass", C.c_float), ("viscosity", C.c_float),
                ("cohesion_coef", C.c_float), ("delta", C.c_float), ("stiff_mult", C.c_float),
                ("sph_max_vel", C.c_float), ("periodic", C.c_int32), ("temperature", C.c_int32),
                ("sph_ground", C.c_int32), ("compensated", C.c_int32), ("accuracy", C.c_int32),
                ("leaf_capacity", C.c_uint32), ("walk_block", C.c_uint32), ("use_graph", C.c_int32),
                ("sph_search", C.c_int32), ("sph_search_radius", C.c_float),
                ("domain_w", C.c_float), ("domain_h", C.c_float), ("vertical_gravity", C.c_float),
                ("boundary_friction", C.c_float), ("constraints", C.c_int32), ("reserved", C.c_uint32 * 1)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class GeNode(C.Structure):
    _fields_ = [("posx", C.c_float), ("posy", C.c_float), ("comx", C.c_float),
                ("comy", C.c_float), ("size", C.c_float), ("mass", C.c_float),
                ("start", C.c_uint32), ("end", C.c_uint32), ("next", C.c_uint32),
                ("child", C.c_uint32 * 4)]


NODE_DTYPE = np.dtype([("posx", "<f4"), ("posy", "<f4"), ("comx", "<f4"), ("comy", "<f4"),
                       ("size", "<f4"), ("mass", "<f4"), ("start", "<u4"), ("end", "<u4"),
                       ("next", "<u4"), ("child", "<u4", (4,))])
assert NODE_DTYPE.itemsize == C.sizeof(GeNode)


class GeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ge error {code}: {msg}")
        self.code = code


def library_path():
    return os.environ.get("GE_PHYSICS_LIB", os.path.join(_PKG, "libge_physics.so"))


_lib = None


def load_library():
    """Load libge_physics.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run __graft_entry__.build() or make -C galaxy-engine_amd")
    lib = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ge_default_config": (None, [P(GeConfig)]),
        "ge_default_params": (None, [P(StepParams)]),
        "ge_create": (C.c_int, [P(vp), P(GeConfig)]),
        "ge_destroy": (C.c_int, [vp]),
        "ge_last_error": (C.c_char_p, [vp]),
        "ge_version": (C.c_char_p, []),
        "ge_upload": (C.c_int, [vp, C.c_uint32, P(GeParticles)]),
        "ge_download": (C.c_int, [vp, P(GeParticles)]),
        "ge_count": (C.c_int, [vp, P(C.c_uint32)]),
        "ge_sync": (C.c_int, [vp]),
        "ge_step": (C.c_int, [vp, P(StepParams), C.c_int32]),
        "ge_bounding_box": (C.c_int, [vp, P(C.c_float)]),
        "ge_compute_morton_keys": (C.c_int, [vp, P(C.c_float)]),
        "ge_sort_by_morton_key": (C.c_int, [vp]),
        "ge_build_quadtree": (C.c_int, [vp, P(C.c_float)]),
        "ge_compute_gravity": (C.c_int, [vp, P(StepParams)]),
        "ge_integrate_start": (C.c_int, [vp, P(StepParams)]),
        "ge_integrate_end": (C.c_int, [vp, P(StepParams)]),
        "ge_sph_new_grid": (C.c_int, [vp, P(StepParams)]),
        "ge_sph_pcisph": (C.c_int, [vp, P(StepParams)]),
        "ge_get_keys": (C.c_int, [vp, P(C.c_uint64)]),
        "ge_get_perm": (C.c_int, [vp, P(C.c_uint32)]),
        "ge_get_node_count": (C.c_int, [vp, P(C.c_uint32)]),
        "ge_get_nodes": (C.c_int, [vp, vp, C.c_uint32]),
        "ge_get_interactions": (C.c_int, [vp, P(C.c_uint64)]),
        "ge_get_interaction_counts": (C.c_int, [vp, P(C.c_uint64), P(C.c_uint64)]),
        "ge_get_walk_stats": (C.c_int, [vp, P(C.c_uint64)]),
        "ge_upload_aos": (C.c_int, [vp, C.c_uint32, vp, P(AosLayout), P(C.c_uint32)]),
        "ge_neighbor_counts": (C.c_int, [vp, P(StepParams), P(C.c_uint32)]),
        "ge_download_aos": (C.c_int, [vp, vp, P(AosLayout), P(C.c_uint32)]),
        "ge_get_neighbors": (C.c_int, [vp, C.c_int32, P(C.c_int64), P(C.c_uint32), C.c_int64,
                                       P(C.c_int64)]),
        "ge_profile_enable": (C.c_int, [vp, C.c_int]),
        "ge_profile_read": (C.c_int, [vp, P(C.c_double), P(C.c_uint32)]),
        "ge_upload_systems": (C.c_int, [vp, C.c_uint32, P(C.c_uint32), P(GeParticles)]),
        "ge_system_offsets": (C.c_int, [vp, P(C.c_uint32), P(C.c_uint32), C.c_uint32]),
        "ge_systems_update": (C.c_int, [vp, C.c_uint32, P(C.c_uint32), C.c_uint32, P(C.c_uint32),
                                        P(GeParticles)]),
        "ge_download_system": (C.c_int, [vp, C.c_uint32, P(GeParticles)]),
        "ge_poll": (C.c_int, [vp, P(C.c_int32)]),
        "ge_memory_bytes": (C.c_int, [vp, P(C.c_uint64)]),
        "ge_diagnostics": (C.c_int, [vp, C.c_uint32, P(StepParams), C.c_int32, P(GeDiag)]),
        "ge_diagnostics_noalloc": (C.c_int, [vp, C.c_uint32, P(StepParams), C.c_int32, P(GeDiag)]),
        "ge_direct_fp64": (C.c_int, [vp, C.c_uint32, P(C.c_uint32), P(StepParams), P(C.c_double),
                                     P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names from EXPORTS that the loaded library resolves."""
    lib = load_library()
    out = []
    for name in EXPORTS:
        try:
            getattr(lib, name)
            out.append(name)
        except AttributeError:
            pass
    return out


def default_params(**kw):
    lib = load_library()
    p = StepParams()
    lib.ge_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Context:
    """One device context (ge_ctx)."""

    def __init__(self, capacity, sph=False, device=0, max_systems=1, render=False, temperature=False,
                 max_constraints=0, precise_positions=False):
        self.lib = load_library()
        cfg = GeConfig()
        self.lib.ge_default_config(C.byref(cfg))
        cfg.capacity = int(capacity)
        cfg.device = int(device)
        cfg.enable_sph = 1 if sph else 0
        cfg.max_systems = int(max_systems)
        cfg.enable_render = 1 if render else 0
        cfg.enable_temperature = 1 if temperature else 0
        cfg.max_constraints = int(max_constraints)
        cfg.precise_positions = 1 if precise_positions else 0
        self.render = bool(render)
        self.temperature = bool(temperature)
        h = C.c_void_p()
        rc = self.lib.ge_create(C.byref(h), C.byref(cfg))
        if rc != GE_OK:
            raise GeError(rc, "ge_create failed")
        self.h = h
        self.sph = sph
        self.capacity = int(capacity)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ge_destroy(self.h)
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
            raise GeError(rc, self.lib.ge_last_error(self.h).decode())

    # ---- state transfer
    @staticmethod
    def _particles(state, n):
        keep = []
        gp = GeParticles()
        for f in PARTICLE_ORDER:
            if f in state and state[f] is not None:
                dt = np.uint32 if f in FIELDS_U else np.float32
                a = np.ascontiguousarray(state[f], dtype=dt)
                assert len(a) == n, f
                keep.append(a)
                setattr(gp, f, _uptr(a) if f in FIELDS_U else _fptr(a))
        return gp, keep

    def upload(self, state, counts=None):
        """state: dict of numpy arrays (x, y, mass required).  counts: particles
        per system when the arrays hold several systems back to back."""
        n = len(state["x"])
        gp, keep = self._particles(state, n)
        if counts is None:
            self._check(self.lib.ge_upload(self.h, n, C.byref(gp)))
        else:
            cnt = np.ascontiguousarray(counts, dtype=np.uint32)
            assert int(cnt.sum()) == n
            self._check(self.lib.ge_upload_systems(self.h, len(cnt), _uptr(cnt), C.byref(gp)))

    def systems_update(self, keep, state=None, counts=()):
        """Keep systems `keep` (ascending current indices), then append the
        systems of `state` (counts per system, concatenated)."""
        kp = np.ascontiguousarray(keep, dtype=np.uint32)
        cnt = np.ascontiguousarray(counts, dtype=np.uint32)
        if state is not None and len(cnt):
            gp, hold = self._particles(state, int(cnt.sum()))
            gpp = C.byref(gp)
        else:
            gpp, hold = None, []
        self._check(self.lib.ge_systems_update(self.h, len(kp), _uptr(kp) if len(kp) else None, len(cnt),
                                               _uptr(cnt) if len(cnt) else None, gpp))
        del hold

    def download_system(self, sys, fields=None):
        off = self.system_offsets()
        n = int(off[sys + 1] - off[sys])
        if fields is None:
            fields = FIELDS_F + FIELDS_U + (FIELDS_SPH if self.sph else [])
        out = {}
        gp = GeParticles()
        for f in fields:
            a = np.empty(n, dtype=np.uint32 if f in FIELDS_U else np.float32)
            out[f] = a
            setattr(gp, f, _uptr(a) if f in FIELDS_U else _fptr(a))
        self._check(self.lib.ge_download_system(self.h, int(sys), C.byref(gp)))
        return out

    def poll(self):
        d = C.c_int32()
        self._check(self.lib.ge_poll(self.h, C.byref(d)))
        return bool(d.value)

    def diagnostics(self, sys=0, params=None, potential=False, noalloc=False):
        """ge_diagnostics, or ge_diagnostics_noalloc (the entry point that
        allocates nothing) with noalloc=True."""
        d = GeDiag()
        fn = self.lib.ge_diagnostics_noalloc if noalloc else self.lib.ge_diagnostics
        self._check(fn(self.h, int(sys), C.byref(params) if params is not None else None,
                       1 if potential else 0, C.byref(d)))
        return {k: getattr(d, k) for k, _ in d._fields_}

    def direct_fp64(self, targets, params):
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        ax = np.zeros(len(t), np.float64)
        ay = np.zeros(len(t), np.float64)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.ge_direct_fp64(self.h, len(t), _uptr(t), C.byref(params),
                                            ax.ctypes.data_as(dp), ay.ctypes.data_as(dp)))
        return ax, ay

    def memory_bytes(self):
        b = C.c_uint64()
        self._check(self.lib.ge_memory_bytes(self.h, C.byref(b)))
        return b.value

    def upload_aos(self, records, flags=None):
        """records: numpy structured array (ge_particles field names)."""
        lay = AosLayout.from_dtype(records.dtype)
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        self._check(self.lib.ge_upload_aos(self.h, len(records), C.c_void_p(records.ctypes.data), C.byref(lay),
                                           _uptr(fl) if fl is not None else None))

    def download_aos(self, dtype, with_flags=False):
        n = self.count()
        rec = np.zeros(n, dtype=dtype)
        fl = np.zeros(n, np.uint32) if with_flags else None
        lay = AosLayout.from_dtype(rec.dtype)
        self._check(self.lib.ge_download_aos(self.h, C.c_void_p(rec.ctypes.data), C.byref(lay),
                                             _uptr(fl) if fl is not None else None))
        return (rec, fl) if with_flags else rec

    def system_offsets(self):
        ns = C.c_uint32()
        self._check(self.lib.ge_system_offsets(self.h, C.byref(ns), None, 0))
        off = np.zeros(ns.value + 1, np.uint32)
        self._check(self.lib.ge_system_offsets(self.h, C.byref(ns), _uptr(off), len(off)))
        return off

    def count(self):
        n = C.c_uint32()
        self._check(self.lib.ge_count(self.h, C.byref(n)))
        return n.value

    def download(self, fields=None):
        n = self.count()
        if fields is None:
            fields = FIELDS_F + FIELDS_U + (FIELDS_SPH if self.sph else [])
        out = {}
        gp = GeParticles()
        for f in fields:
            a = np.empty(n, dtype=np.uint32 if f in FIELDS_U else np.float32)
            out[f] = a
            setattr(gp, f, _uptr(a) if f in FIELDS_U else _fptr(a))
        self._check(self.lib.ge_download(self.h, C.byref(gp)))
        return out

    def sync(self):
        self._check(self.lib.ge_sync(self.h))

    # ---- frames and stages
    def step(self, params, nsteps=1):
        self._check(self.lib.ge_step(self.h, C.byref(params), int(nsteps)))

    def bounding_box(self):
        bb = (C.c_float * 3)()
        self._check(self.lib.ge_bounding_box(self.h, bb))
        return np.array(list(bb), dtype=np.float32)

    def compute_morton_keys(self, bb):
        b = (C.c_float * 3)(*[float(v) for v in bb])
        self._check(self.lib.ge_compute_morton_keys(self.h, b))

    def sort_by_morton_key(self):
        self._check(self.lib.ge_sort_by_morton_key(self.h))

    def build_quadtree(self, bb):
        b = (C.c_float * 3)(*[float(v) for v in bb])
        self._check(self.lib.ge_build_quadtree(self.h, b))

    def compute_gravity(self, params):
        self._check(self.lib.ge_compute_gravity(self.h, C.byref(params)))

    def integrate_start(self, params):
        self._check(self.lib.ge_integrate_start(self.h, C.byref(params)))

    def integrate_end(self, params):
        self._check(self.lib.ge_integrate_end(self.h, C.byref(params)))

    def sph_new_grid(self, params):
        self._check(self.lib.ge_sph_new_grid(self.h, C.byref(params)))

    def sph_pcisph(self, params):
        self._check(self.lib.ge_sph_pcisph(self.h, C.byref(params)))

    # ---- parity hooks
    def keys(self):
        a = np.empty(self.count(), dtype=np.uint64)
        self._check(self.lib.ge_get_keys(self.h, a.ctypes.data_as(C.POINTER(C.c_uint64))))
        return a

    def perm(self):
        a = np.empty(self.count(), dtype=np.uint32)
        self._check(self.lib.ge_get_perm(self.h, _uptr(a)))
        return a

    def node_count(self):
        c = C.c_uint32()
        self._check(self.lib.ge_get_node_count(self.h, C.byref(c)))
        return c.value

    def nodes(self):
        nn = self.node_count()
        a = np.zeros(nn, dtype=NODE_DTYPE)
        self._check(self.lib.ge_get_nodes(self.h, C.c_void_p(a.ctypes.data), nn))
        return a

    def interactions(self):
        c = C.c_uint64()
        self._check(self.lib.ge_get_interactions(self.h, C.byref(c)))
        return c.value

    def interaction_counts(self):
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.ge_get_interaction_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def walk_stats(self):
        a = (C.c_uint64 * 4)()
        self._check(self.lib.ge_get_walk_stats(self.h, a))
        return {"interactions": a[0], "nodes": a[1], "visits": a[2], "leaf_trips": a[3]}

    def profile_enable(self, on=True):
        self._check(self.lib.ge_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """{phase: (total_ms, launches)} since the last enable/read."""
        ms = (C.c_double * len(PHASES))()
        cnt = (C.c_uint32 * len(PHASES))()
        self._check(self.lib.ge_profile_read(self.h, ms, cnt))
        return {ph: (ms[k], cnt[k]) for k, ph in enumerate(PHASES)}

    def neighbor_counts(self, params):
        out = np.zeros(self.count(), np.uint32)
        self._check(self.lib.ge_neighbor_counts(self.h, C.byref(params), _uptr(out)))
        return out

    def neighbors(self, which=0):
        n = self.count()
        off = np.zeros(n + 1, dtype=np.int64)
        tot = C.c_int64()
        P64 = C.POINTER(C.c_int64)
        self._check(self.lib.ge_get_neighbors(self.h, which, off.ctypes.data_as(P64), None, 0,
                                              C.byref(tot)))
        idx = np.zeros(max(tot.value, 1), dtype=np.uint32)
        self._check(self.lib.ge_get_neighbors(self.h, which, off.ctypes.data_as(P64), _uptr(idx),
                                              tot.value, C.byref(tot)))
        return off, idx[:tot.value]
