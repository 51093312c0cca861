"""ctypes wrappers of the 3D CPU restatement (ge_oracle3d.c in libge_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the bench's cpu_baseline leg,
never by the product library or its Python binding.
"""
import ctypes as C

import numpy as np

import oracle as _o

NODE3_DTYPE = np.dtype([("posx", "<f4"), ("posy", "<f4"), ("posz", "<f4"),
                        ("comx", "<f4"), ("comy", "<f4"), ("comz", "<f4"),
                        ("size", "<f4"), ("mass", "<f4"), ("start", "<u4"), ("end", "<u4"),
                        ("next", "<u4"), ("child", "<u4", (8,))])

STATE3_F = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "pvx", "pvy", "pvz", "mass"]
SPH3_F = ["sph_mass", "rest_dens", "stiff", "visc", "cohesion", "dens", "press", "press_tmp", "pred_dens",
          "ppx", "ppy", "ppz", "pfx", "pfy", "pfz"]

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class _State3(C.Structure):
    _fields_ = ([("n", C.c_int32)] + [(f, f32p) for f in
                                      ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "pvx", "pvy", "pvz", "m"]]
                + [("flags", u32p), ("id", u32p)] + [(f, f32p) for f in SPH3_F])


class _Domain3(C.Structure):
    _fields_ = [("w", C.c_float), ("h", C.c_float), ("d", C.c_float)]


class _Params3(C.Structure):
    _fields_ = [("G", C.c_float), ("theta", C.c_float), ("softening", C.c_float),
                ("dt_start", C.c_float), ("dt", C.c_float), ("naive", C.c_int32),
                ("direct_below", C.c_int32), ("sph_clamp", C.c_int32), ("sph_max_vel", C.c_float),
                ("periodic", C.c_int32), ("domain_w", C.c_float), ("domain_h", C.c_float),
                ("domain_d", C.c_float), ("leaf_cap", C.c_uint32), ("sph", C.c_int32),
                ("sph_h", C.c_float), ("sph_cell", C.c_float), ("sph_unit_mass", C.c_float),
                ("viscosity", C.c_float), ("cohesion_coef", C.c_float), ("delta", C.c_float),
                ("stiff_mult", C.c_float), ("sph_ground", C.c_int32), ("vertical_gravity", C.c_float),
                ("boundary_friction", C.c_float)]


_ready = False


def load():
    global _ready
    lib = _o.load()
    if _ready:
        return lib
    lib.geo3_bbox.argtypes = [C.c_int32, f32p, f32p, f32p, f32p]
    lib.geo3_scale_to_grid.restype = C.c_uint64
    lib.geo3_scale_to_grid.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.geo3_spread_bits.restype = C.c_uint64
    lib.geo3_spread_bits.argtypes = [C.c_uint64]
    lib.geo3_morton.restype = C.c_uint64
    lib.geo3_morton.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.geo3_morton_keys.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, u64p]
    lib.geo3_build_tree.restype = C.c_int32
    lib.geo3_build_tree.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, f32p, C.c_uint32, C.c_void_p, C.c_int32]
    lib.geo3_bh_accel.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, C.c_void_p, C.c_int32, C.c_float,
                                  C.c_float, C.c_float, C.POINTER(_Domain3), f32p, f32p, f32p, u64p]
    lib.geo3_direct_accel.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, C.c_float, C.c_float,
                                      C.POINTER(_Domain3), f32p, f32p, f32p]
    lib.geo3_direct_fp64.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, C.c_double, C.c_double, C.c_int32,
                                     u32p, f64p, f64p, f64p]
    lib.geo3_integrate_start.argtypes = [C.POINTER(_State3), C.c_float, C.c_int32, C.c_float,
                                         C.POINTER(_Domain3)]
    lib.geo3_integrate_end.argtypes = [C.POINTER(_State3), C.c_float]
    lib.geo3_step.restype = C.c_int32
    lib.geo3_step.argtypes = [C.POINTER(_State3), C.POINTER(_Params3), C.c_int32]
    lib.geo3_sph_constants.argtypes = [C.c_float, f32p]
    lib.geo3_sph_candidates.restype = C.c_int64
    lib.geo3_sph_candidates.argtypes = [C.c_int32, f32p, f32p, f32p, C.c_int32, f32p, f32p, f32p, C.c_float,
                                        C.POINTER(C.c_int64), u32p, C.c_int64]
    lib.geo3_sph_pcisph.restype = C.c_int32
    lib.geo3_sph_pcisph.argtypes = [C.POINTER(_State3), C.POINTER(_Params3)]
    lib.geo3_sph_ground.argtypes = [C.POINTER(_State3), C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float]
    _ready = True
    return lib


def _f(a):
    return a.ctypes.data_as(f32p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _dom(domain):
    return _Domain3(*domain) if domain is not None else _Domain3(0.0, 0.0, 0.0)


def bbox(x, y, z):
    x, y, z = f32(x), f32(y), f32(z)
    out = np.zeros(4, np.float32)
    load().geo3_bbox(len(x), _f(x), _f(y), _f(z), _f(out))
    return out


def scale_to_grid(p, lo, hi):
    return load().geo3_scale_to_grid(float(np.float32(p)), float(np.float32(lo)), float(np.float32(hi)))


def spread_bits(v):
    return load().geo3_spread_bits(int(v))


def morton3d(ix, iy, iz):
    return load().geo3_morton(int(ix), int(iy), int(iz))


def morton_keys(x, y, z, bb):
    x, y, z, bb = f32(x), f32(y), f32(z), f32(bb)
    k = np.zeros(len(x), np.uint64)
    load().geo3_morton_keys(len(x), _f(x), _f(y), _f(z), _f(bb), k.ctypes.data_as(u64p))
    return k


def sort_perm(keys):
    return _o.sort_perm(keys)


def build_tree(x, y, z, m, bb, leaf_cap=0):
    x, y, z, m, bb = f32(x), f32(y), f32(z), f32(m), f32(bb)
    lib = load()
    cap = 8 * len(x) + 16
    nodes = np.zeros(cap, NODE3_DTYPE)
    c = lib.geo3_build_tree(len(x), _f(x), _f(y), _f(z), _f(m), _f(bb), int(leaf_cap),
                            C.c_void_p(nodes.ctypes.data), cap)
    if c < 0:
        raise RuntimeError("octree build failed")
    if c > cap:
        nodes = np.zeros(c, NODE3_DTYPE)
        lib.geo3_build_tree(len(x), _f(x), _f(y), _f(z), _f(m), _f(bb), int(leaf_cap),
                            C.c_void_p(nodes.ctypes.data), c)
    return nodes[:c].copy()


def bh_accel(x, y, z, m, nodes, G, theta, eps, domain=None):
    x, y, z, m = f32(x), f32(y), f32(z), f32(m)
    nodes = np.ascontiguousarray(nodes, dtype=NODE3_DTYPE)
    a = [np.zeros(len(x), np.float32) for _ in range(3)]
    cnt = C.c_uint64()
    d = _dom(domain)
    load().geo3_bh_accel(len(x), _f(x), _f(y), _f(z), _f(m), C.c_void_p(nodes.ctypes.data), len(nodes),
                         G, theta, eps, C.byref(d), _f(a[0]), _f(a[1]), _f(a[2]), C.byref(cnt))
    return a[0], a[1], a[2], cnt.value


def direct_accel(x, y, z, m, G, eps, domain=None):
    x, y, z, m = f32(x), f32(y), f32(z), f32(m)
    a = [np.zeros(len(x), np.float32) for _ in range(3)]
    d = _dom(domain)
    load().geo3_direct_accel(len(x), _f(x), _f(y), _f(z), _f(m), G, eps, C.byref(d), _f(a[0]), _f(a[1]),
                             _f(a[2]))
    return a[0], a[1], a[2]


def direct_fp64(x, y, z, m, G, eps, targets=None):
    x, y, z, m = f32(x), f32(y), f32(z), f32(m)
    if targets is None:
        targets = np.arange(len(x), dtype=np.uint32)
    targets = np.ascontiguousarray(targets, dtype=np.uint32)
    a = [np.zeros(len(targets), np.float64) for _ in range(3)]
    load().geo3_direct_fp64(len(x), _f(x), _f(y), _f(z), _f(m), G, eps, len(targets),
                            targets.ctypes.data_as(u32p), *[v.ctypes.data_as(f64p) for v in a])
    return a[0], a[1], a[2]


def sph_constants(h):
    out = np.zeros(3, np.float32)
    load().geo3_sph_constants(float(np.float32(h)), _f(out))
    return out


def sph_candidates(x, y, z, qx, qy, qz, cell):
    """queryNeighborsAVX2 candidate sets: CSR (off, idx), idx ascending per query."""
    x, y, z, qx, qy, qz = (f32(v) for v in (x, y, z, qx, qy, qz))
    lib = load()
    off = np.zeros(len(qx) + 1, np.int64)
    offp = off.ctypes.data_as(C.POINTER(C.c_int64))
    total = lib.geo3_sph_candidates(len(x), _f(x), _f(y), _f(z), len(qx), _f(qx), _f(qy), _f(qz),
                                    float(np.float32(cell)), offp, None, 0)
    total = int(off[-1])
    idx = np.zeros(max(total, 1), np.uint32)
    lib.geo3_sph_candidates(len(x), _f(x), _f(y), _f(z), len(qx), _f(qx), _f(qy), _f(qz),
                            float(np.float32(cell)), offp, idx.ctypes.data_as(u32p), total)
    return off, idx[:total]


DEFAULTS3 = dict(G=1.0, theta=0.8, softening=1.0, dt=0.01, naive=0, direct_below=5000, sph_clamp=0,
                 sph_max_vel=250.0, periodic=0, domain_w=1300.0, domain_h=1300.0, domain_d=1300.0,
                 leaf_cap=0, sph=0, sph_h=3.0, sph_cell=3.0, sph_unit_mass=0.03, viscosity=0.3,
                 cohesion_coef=1.0, delta=19000.0, stiff_mult=1.0, sph_ground=0, vertical_gravity=0.0,
                 boundary_friction=0.0)


def _params3(prm):
    p = _Params3()
    d = dict(DEFAULTS3)
    d.update({k: v for k, v in prm.items() if k in d or k == "dt_start"})
    d.setdefault("dt_start", d["dt"])
    for k, v in d.items():
        setattr(p, k, v)
    return p


class State3:
    """Host 3D particle state for the oracle step; arrays are permuted in place."""

    def __init__(self, st):
        n = len(st["x"])
        self.n = n
        self.a = {}
        for f in STATE3_F:
            self.a[f] = f32(st[f]).copy() if f in st and st[f] is not None else np.zeros(n, np.float32)
        self.a["flags"] = np.ascontiguousarray(st.get("flags", np.zeros(n)), dtype=np.uint32).copy()
        self.a["id"] = np.ascontiguousarray(st.get("id", np.arange(n)), dtype=np.uint32).copy()
        self._s = _State3()
        self._s.n = n
        for f in STATE3_F:
            setattr(self._s, "m" if f == "mass" else f, _f(self.a[f]))
        self._s.flags = self.a["flags"].ctypes.data_as(u32p)
        self._s.id = self.a["id"].ctypes.data_as(u32p)
        # SPH fields: uploaded values or the device defaults (sph_mass 1, else 0)
        for f in SPH3_F:
            v = st.get(f)
            self.a[f] = f32(v).copy() if v is not None else (np.ones(n, np.float32) if f == "sph_mass"
                                                             else np.zeros(n, np.float32))
            setattr(self._s, f, _f(self.a[f]))

    def step(self, prm, nsteps=1):
        p = _params3(prm)
        rc = load().geo3_step(C.byref(self._s), C.byref(p), int(nsteps))
        if rc != 0:
            raise RuntimeError(f"oracle 3D step failed {rc}")

    def sph_pcisph(self, prm):
        p = _params3(prm)
        load().geo3_sph_pcisph(C.byref(self._s), C.byref(p))

    def sph_ground(self, radius, w, h, d, vertical_gravity, friction):
        load().geo3_sph_ground(C.byref(self._s), radius, w, h, d, vertical_gravity, friction)

    def integrate_start(self, dt, sph_clamp=0, sph_max_vel=250.0, domain=None):
        d = _dom(domain)
        load().geo3_integrate_start(C.byref(self._s), dt, sph_clamp, sph_max_vel, C.byref(d))

    def integrate_end(self, dt):
        load().geo3_integrate_end(C.byref(self._s), dt)

    def __getitem__(self, k):
        return self.a[k]
