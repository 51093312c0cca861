"""ctypes loader for the CPU restatement (libge_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and the bench's
cpu_baseline leg, never by the product library or its Python binding.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libge_oracle.so")

GEO_PINNED, GEO_SPH, GEO_DRAWN, GEO_DARK = 1, 2, 4, 8

NODE_DTYPE = np.dtype([("posx", "<f4"), ("posy", "<f4"), ("comx", "<f4"), ("comy", "<f4"),
                       ("size", "<f4"), ("mass", "<f4"), ("start", "<u4"), ("end", "<u4"),
                       ("next", "<u4"), ("child", "<u4", (4,))])

STATE_F = ["x", "y", "vx", "vy", "ax", "ay", "pvx", "pvy", "mass"]
STATE_SPH = ["sph_mass", "rest_dens", "stiff", "visc", "cohesion", "dens", "press", "press_tmp",
             "pred_dens", "ppx", "ppy", "pfx", "pfy"]


STATE_TEMP = ["temp", "label", "ke", "prev_ke"]
GEO_HOT_POINT, GEO_SOLIDIFIED = 64, 128


class _State(C.Structure):
    _fields_ = ([("n", C.c_int32)] + [(f, C.POINTER(C.c_float)) for f in
                                      ["x", "y", "vx", "vy", "ax", "ay", "pvx", "pvy", "m"]]
                + [("flags", C.POINTER(C.c_uint32)), ("id", C.POINTER(C.c_uint32))]
                + [(f, C.POINTER(C.c_float)) for f in STATE_SPH]
                + [("temp", C.POINTER(C.c_float)), ("label", C.POINTER(C.c_uint32)),
                   ("ke", C.POINTER(C.c_float)), ("prev_ke", C.POINTER(C.c_float))])


class Material(C.Structure):
    """SPHMaterial (GE/include/Physics/materialsSPH.h:3-47), physics fields."""
    _fields_ = ([("id", C.c_uint32)]
                + [(f, C.c_float) for f in (
                    "mass_mult", "rest_dens", "stiff", "visc", "cohesion",
                    "hot_point", "hot_rest_dens", "hot_mass_mult", "hot_stiff", "hot_visc", "hot_cohesion",
                    "cold_point", "cold_rest_dens", "cold_mass_mult", "cold_stiff", "cold_visc", "cold_cohesion",
                    "heat_conductivity", "constraint_resistance", "constraint_plastic_point",
                    "constraint_plastic_point_mult", "constraint_stiffness")]
                + [("is_plastic", C.c_int32)])


def _mat(id_, mass, rest, stiff, visc, coh, hot, hrest, hmass, hstiff, hvisc, hcoh, cold, crest, cmass,
         cstiff, cvisc, ccoh, k, res, pp_frac, ppm, cst, plastic):
    m = Material()
    vals = dict(id=id_, mass_mult=mass, rest_dens=rest, stiff=stiff, visc=visc, cohesion=coh, hot_point=hot,
                hot_rest_dens=hrest, hot_mass_mult=hmass, hot_stiff=hstiff, hot_visc=hvisc, hot_cohesion=hcoh,
                cold_point=cold, cold_rest_dens=crest, cold_mass_mult=cmass, cold_stiff=cstiff, cold_visc=cvisc,
                cold_cohesion=ccoh, heat_conductivity=k, constraint_resistance=res,
                constraint_plastic_point=float(np.float32(res) * np.float32(pp_frac)),
                constraint_plastic_point_mult=ppm, constraint_stiffness=cst, is_plastic=plastic)
    for key, v in vals.items():
        setattr(m, key, v)
    return m


def reference_materials():
    """The seven materials SPHMaterials::Init registers (materialsSPH.h:49-300); a
    material whose cold point is 0 takes its base values as cold values."""
    return [
        _mat(1, 0.6, 0.095, 1.0, 0.075, 0.05, 373.2, 0.065, 0.25, 1.0, 0.055, 0.0, 273.2, 0.046, 0.5, 0.6, 2.2,
             2500.0, 0.15, 0.06, 0.5, 0.0, 55.0, 0),
        _mat(2, 3.3, 0.008, 1.4, 3.0, 1750.0, 1370.0, 0.005, 2.8, 1.0, 0.6, 300.0, 0.0, 0.008, 3.3, 1.4, 3.0,
             1750.0, 0.02, 0.211, 0.5, 0.0, 60.0, 0),
        _mat(3, 4.0, 0.008, 1.4, 3.0, 1750.0, 1740.0, 0.005, 2.8, 1.0, 0.6, 300.0, 0.0, 0.008, 4.0, 1.4, 3.0,
             1750.0, 0.02, 0.24, 0.45, 1.8, 60.0, 1),
        _mat(4, 2.1, 0.008, 1.255, 0.74, 1.0, 1200.0, 0.011, 1.9, 1.12, 0.6, 1.0, 0.0, 0.008, 2.1, 1.255, 0.74,
             1.0, 0.01, 0.08, 0.5, 0.0, 55.0, 0),
        _mat(5, 1.9, 0.008, 1.0, 2.23, 3000.0, 950.0, 0.013, 1.8, 0.9, 1.8, 600.0, 0.0, 0.008, 1.9, 1.0, 2.23,
             3000.0, 0.02, 0.09, 0.5, 2.0, 30.0, 1),
        _mat(6, 2.3, 0.0095, 1.0, 0.6, 100.0, 1000.0, 0.011, 2.1, 1.0, 0.5, 40.0, 0.0, 0.0095, 2.3, 1.0, 0.6,
             100.0, 0.1, 0.03, 0.5, 2.0, 20.0, 1),
        _mat(7, 1.7, 0.0095, 1.0, 0.6, 100.0, 453.0, 0.011, 1.6, 1.0, 0.5, 40.0, 0.0, 0.0095, 1.7, 1.0, 0.6,
             100.0, 1.1, 5.5, 0.5, 2.0, 6.0, 1),
    ]


class _Params(C.Structure):
    _fields_ = [("G", C.c_float), ("theta", C.c_float), ("softening", C.c_float),
                ("dt_start", C.c_float), ("dt", C.c_float), ("naive", C.c_int32),
                ("direct_below", C.c_int32), ("sph", C.c_int32), ("sph_h", C.c_float),
                ("sph_cell", C.c_float), ("sph_unit_mass", C.c_float), ("viscosity", C.c_float),
                ("cohesion_coef", C.c_float), ("delta", C.c_float), ("stiff_mult", C.c_float),
                ("sph_max_vel", C.c_float), ("sph_search", C.c_int32), ("sph_search_radius", C.c_float),
                ("periodic", C.c_int32), ("domain_w", C.c_float), ("domain_h", C.c_float),
                ("sph_ground", C.c_int32), ("vertical_gravity", C.c_float), ("boundary_friction", C.c_float),
                ("temperature", C.c_int32), ("heat_conductivity", C.c_float), ("ambient_temp", C.c_float),
                ("ambient_heat_rate", C.c_float), ("materials", C.c_void_p), ("nmaterials", C.c_int32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    f32p = C.POINTER(C.c_float)
    u32p = C.POINTER(C.c_uint32)
    u64p = C.POINTER(C.c_uint64)
    f64p = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_int64)
    lib.geo_bbox.argtypes = [C.c_int32, f32p, f32p, f32p]
    lib.geo_scale_to_grid.restype = C.c_uint64
    lib.geo_scale_to_grid.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.geo_spread_bits.restype = C.c_uint64
    lib.geo_spread_bits.argtypes = [C.c_uint64]
    lib.geo_morton2d.restype = C.c_uint64
    lib.geo_morton2d.argtypes = [C.c_uint64, C.c_uint64]
    lib.geo_morton_keys.argtypes = [C.c_int32, f32p, f32p, f32p, u64p]
    lib.geo_sort_perm.argtypes = [C.c_int32, u64p, u32p]
    lib.geo_build_tree.restype = C.c_int32
    lib.geo_build_tree.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, C.c_void_p, C.c_int32]
    lib.geo_sph_candidates_mode.restype = C.c_int64
    lib.geo_sph_candidates_mode.argtypes = [C.c_int32, f32p, f32p, C.c_int32, f32p, f32p, C.c_float, C.c_int32,
                                            C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_uint32), C.c_int64]
    lib.geo_neighbor_amount.restype = None
    lib.geo_neighbor_amount.argtypes = [C.c_int32, f32p, f32p, C.POINTER(C.c_uint32), C.c_float, C.c_float,
                                        C.POINTER(C.c_uint32)]
    lib.geo_build_tree_cap.restype = C.c_int32
    lib.geo_build_tree_cap.argtypes = [C.c_int32, f32p, f32p, f32p, f32p, C.c_uint32, C.c_void_p, C.c_int32]
    lib.geo_bh_accel.argtypes = [C.c_int32, f32p, f32p, f32p, C.c_void_p, C.c_int32, C.c_float,
                                 C.c_float, C.c_float, f32p, f32p, u64p]
    lib.geo_direct_accel.argtypes = [C.c_int32, f32p, f32p, f32p, C.c_float, C.c_float, f32p, f32p]
    lib.geo_direct_fp64.argtypes = [C.c_int32, f32p, f32p, f32p, C.c_double, C.c_double, C.c_int32,
                                    u32p, f64p, f64p]
    lib.geo_sph_candidates.restype = C.c_int64
    lib.geo_sph_candidates.argtypes = [C.c_int32, f32p, f32p, C.c_int32, f32p, f32p, C.c_float,
                                       i64p, u32p, C.c_int64]
    lib.geo_sph_constants.argtypes = [C.c_float, f32p]
    lib.geo_build_info.restype = C.c_char_p
    lib.geo_build_info.argtypes = []
    lib.geo_step.restype = C.c_int32
    lib.geo_step.argtypes = [C.POINTER(_State), C.POINTER(_Params), C.c_int32]
    lib.geo_sph_pcisph.restype = C.c_int32
    lib.geo_sph_pcisph.argtypes = [C.POINTER(_State), C.POINTER(_Params)]
    lib.geo_integrate_start.restype = None
    lib.geo_integrate_start.argtypes = [C.POINTER(_State), C.c_float, C.c_int32, C.c_float]
    lib.geo_node_temps.restype = None
    lib.geo_node_temps.argtypes = [C.c_void_p, C.c_int32, f32p, f32p]
    lib.geo_walk_heat.restype = None
    lib.geo_walk_heat.argtypes = [C.c_int32, f32p, f32p, C.c_void_p, C.c_int32, f32p, C.c_float, C.c_float,
                                  C.c_float, C.c_float, C.c_void_p, f32p]
    lib.geo_constraints.restype = C.c_int32
    lib.geo_constraints.argtypes = [C.POINTER(_State), C.c_void_p, C.POINTER(C.c_uint32), C.c_int32,
                                    C.POINTER(_CParams), C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]
    lib.geo_create_constraints.restype = C.c_int32
    lib.geo_create_constraints.argtypes = [C.POINTER(_State), C.c_void_p, C.c_int32, C.c_int32, C.c_float,
                                           C.c_float, C.c_void_p, C.c_int32]
    lib.geo_merge.restype = C.c_int32
    lib.geo_merge.argtypes = [C.POINTER(_State), f32p, f32p, C.POINTER(C.c_int32), C.c_void_p]
    lib.geo_temperature.restype = None
    lib.geo_temperature.argtypes = [C.POINTER(_State), C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float]
    _lib = lib
    return lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def bbox(x, y):
    x, y = f32(x), f32(y)
    out = np.zeros(3, np.float32)
    load().geo_bbox(len(x), _f(x), _f(y), _f(out))
    return out


def scale_to_grid(p, lo, hi):
    return load().geo_scale_to_grid(float(np.float32(p)), float(np.float32(lo)), float(np.float32(hi)))


def morton2d(ix, iy):
    return load().geo_morton2d(int(ix), int(iy))


def morton_keys(x, y, bb):
    x, y, bb = f32(x), f32(y), f32(bb)
    k = np.zeros(len(x), np.uint64)
    load().geo_morton_keys(len(x), _f(x), _f(y), _f(bb), k.ctypes.data_as(C.POINTER(C.c_uint64)))
    return k


def sort_perm(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    p = np.zeros(len(keys), np.uint32)
    load().geo_sort_perm(len(keys), keys.ctypes.data_as(C.POINTER(C.c_uint64)), _u(p))
    return p


def build_tree(x, y, m, bb, leaf_cap=0):
    x, y, m, bb = f32(x), f32(y), f32(m), f32(bb)
    lib = load()
    cap = 4 * len(x) + 16
    nodes = np.zeros(cap, NODE_DTYPE)
    c = lib.geo_build_tree_cap(len(x), _f(x), _f(y), _f(m), _f(bb), int(leaf_cap),
                               C.c_void_p(nodes.ctypes.data), cap)
    if c < 0:
        raise RuntimeError("tree build failed")
    if c > cap:
        nodes = np.zeros(c, NODE_DTYPE)
        lib.geo_build_tree_cap(len(x), _f(x), _f(y), _f(m), _f(bb), int(leaf_cap),
                               C.c_void_p(nodes.ctypes.data), c)
    return nodes[:c].copy()


def bh_accel(x, y, m, nodes, G, theta, eps):
    x, y, m = f32(x), f32(y), f32(m)
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    ax = np.zeros(len(x), np.float32)
    ay = np.zeros(len(x), np.float32)
    cnt = C.c_uint64()
    load().geo_bh_accel(len(x), _f(x), _f(y), _f(m), C.c_void_p(nodes.ctypes.data), len(nodes),
                        G, theta, eps, _f(ax), _f(ay), C.byref(cnt))
    return ax, ay, cnt.value


def direct_accel(x, y, m, G, eps):
    x, y, m = f32(x), f32(y), f32(m)
    ax = np.zeros(len(x), np.float32)
    ay = np.zeros(len(x), np.float32)
    load().geo_direct_accel(len(x), _f(x), _f(y), _f(m), G, eps, _f(ax), _f(ay))
    return ax, ay


def direct_fp64(x, y, m, G, eps, targets=None):
    x, y, m = f32(x), f32(y), f32(m)
    if targets is None:
        targets = np.arange(len(x), dtype=np.uint32)
    targets = np.ascontiguousarray(targets, dtype=np.uint32)
    ax = np.zeros(len(targets), np.float64)
    ay = np.zeros(len(targets), np.float64)
    load().geo_direct_fp64(len(x), _f(x), _f(y), _f(m), G, eps, len(targets), _u(targets),
                           ax.ctypes.data_as(C.POINTER(C.c_double)),
                           ay.ctypes.data_as(C.POINTER(C.c_double)))
    return ax, ay


def sph_candidates(x, y, qx, qy, cell, mode=0, radius=3.5):
    """Candidate sets (CSR, ascending per query): mode 0 NeighborSearchV2AVX2,
    mode 1 NeighborSearchV2 (truncated cells, radius filter)."""
    x, y, qx, qy = f32(x), f32(y), f32(qx), f32(qy)
    off = np.zeros(len(qx) + 1, np.int64)
    i64p = C.POINTER(C.c_int64)
    lib = load()
    tot = lib.geo_sph_candidates_mode(len(x), _f(x), _f(y), len(qx), _f(qx), _f(qy), cell, mode, radius,
                                      off.ctypes.data_as(i64p), None, 0)
    if tot < 0:
        tot = int(off[-1])
    idx = np.zeros(max(tot, 1), np.uint32)
    lib.geo_sph_candidates_mode(len(x), _f(x), _f(y), len(qx), _f(qx), _f(qy), cell, mode, radius,
                                off.ctypes.data_as(i64p), _u(idx), tot)
    return off, idx[:tot]


def neighbor_amount(x, y, flags, cell=3.0, radius=3.5):
    x, y = f32(x), f32(y)
    fl = np.ascontiguousarray(flags, dtype=np.uint32)
    out = np.zeros(len(x), np.uint32)
    load().geo_neighbor_amount(len(x), _f(x), _f(y), _u(fl), cell, radius, _u(out))
    return out


CONSTRAINT_DTYPE = np.dtype([("id1", "<u4"), ("id2", "<u4"), ("rest_length", "<f4"), ("original_length", "<f4"),
                             ("stiffness", "<f4"), ("resistance", "<f4"), ("displacement", "<f4"),
                             ("plasticity_point", "<f4"), ("is_broken", "u1"), ("is_plastic", "u1"),
                             ("pad", "u1", (2,))])


class _CParams(C.Structure):
    _fields_ = [("stiffness_mult", C.c_float), ("resistance_mult", C.c_float), ("unbreakable", C.c_int32),
                ("forces", C.c_int32), ("damping", C.c_float), ("correction_ratio", C.c_float),
                ("periodic", C.c_int32), ("domain_w", C.c_float), ("domain_h", C.c_float)]


def constraints(st, cons, colors, materials=(), frame_count=0, **prm):
    """Physics::constraints once, serially in (colour, index) order.  st: dict of
    arrays (x, y, vx, vy, ax, ay, mass, flags, id, optional label); returns
    (state dict, remaining constraints, their colours, frame_count)."""
    temperature = st.get("label") is not None
    s = State(st, sph=False, temperature=temperature)
    cs = np.ascontiguousarray(cons, dtype=CONSTRAINT_DTYPE).copy()
    col = np.ascontiguousarray(colors, dtype=np.uint32).copy()
    p = _CParams()
    d = dict(stiffness_mult=1.0, resistance_mult=1.0, unbreakable=0, forces=1, damping=0.001,
             correction_ratio=0.013333, periodic=0, domain_w=0.0, domain_h=0.0)
    d.update(prm)
    for k, v in d.items():
        setattr(p, k, v)
    mats = list(materials)
    arr = (Material * max(len(mats), 1))(*mats)
    fc = C.c_uint32(frame_count)
    m = load().geo_constraints(C.byref(s._s), cs.ctypes.data, _u(col), len(cs), C.byref(p),
                               C.addressof(arr) if mats else None, len(mats), C.byref(fc))
    return s.a, cs[:m].copy(), col[:m].copy(), fc.value


def create_constraints(st, materials=(), mode=0, cell=3.0, radius=3.5):
    """Physics::createConstraints' automatic branch (solidified particles) on
    the state dict st (x, y, flags, id, optional label).  Returns (flags
    after the call, the appended constraints in creation order)."""
    temperature = st.get("label") is not None
    s = State(st, sph=False, temperature=temperature)
    mats = list(materials)
    arr = (Material * max(len(mats), 1))(*mats)
    cap = 64
    while True:
        out = np.zeros(cap, CONSTRAINT_DTYPE)
        flags = s.a["flags"].copy()
        m = load().geo_create_constraints(C.byref(s._s), C.addressof(arr) if mats else None, len(mats), int(mode),
                                          float(cell), float(radius), out.ctypes.data, cap)
        if m >= 0:
            return s.a["flags"].copy(), out[:m].copy()
        s.a["flags"][:] = flags  # restore and retry with room for every constraint
        cap *= 4


class _MParams(C.Structure):
    _fields_ = [("density_radius", C.c_float), ("cell_size", C.c_float), ("texture_half_size", C.c_float),
                ("size_multiplier", C.c_float)]


def merge(st, size, prev_size, neighbors, density_radius=10.0, cell_size=3.0, texture_half_size=16.0,
          size_multiplier=0.6):
    """UpdateNeighbors + mergerSolver serially on a copy of st (dict of arrays
    with x, y, vx, vy, mass, flags, id, ...); returns (state dict, size,
    prev_size, neighbors), all cut to the particles left."""
    s = State(st, sph=st.get("sph_mass") is not None, temperature=st.get("label") is not None)
    sz = f32(size).copy()
    ps = f32(prev_size).copy()
    nb = np.ascontiguousarray(neighbors, dtype=np.int32).copy()
    p = _MParams(density_radius, cell_size, texture_half_size, size_multiplier)
    m = load().geo_merge(C.byref(s._s), _f(sz), _f(ps), nb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(p))
    out = {k: v[:m].copy() for k, v in s.a.items()}
    return out, sz[:m], ps[:m], nb[:m]


def node_temps(nodes, temp):
    """Node::gridTemp of every node (NODE_DTYPE array) from the particles' temp."""
    nd = np.ascontiguousarray(nodes)
    t = f32(temp)
    out = np.zeros(len(nd), np.float32)
    load().geo_node_temps(nd.ctypes.data, len(nd), _f(t), _f(out))
    return out


def walk_heat(x, y, nodes, gtemp, theta, eps, conductivity, dt, temp):
    x, y, gt = f32(x), f32(y), f32(gtemp)
    nd = np.ascontiguousarray(nodes)
    t = f32(temp).copy()
    dom = _Domain(0.0, 0.0)
    load().geo_walk_heat(len(x), _f(x), _f(y), nd.ctypes.data, len(nd), _f(gt), theta, eps, conductivity, dt,
                         C.byref(dom), _f(t))
    return t


def temperature(st, materials, ambient=274.0, rate=1.0, dt=0.045):
    """Physics::temperatureCalculation on a copy of st (dict of arrays); returns the new dict."""
    s = State(st, sph=True, temperature=True)
    arr = (Material * max(len(materials), 1))(*materials)
    load().geo_temperature(C.byref(s._s), C.addressof(arr) if materials else None, len(materials), ambient, rate,
                           dt)
    return s.a


def build_info():
    """Compiler and flags of libge_oracle.so."""
    return load().geo_build_info().decode()


def set_threads(k):
    """OpenMP threads of the oracle's parallel loops (omp_set_num_threads in
    the libgomp the library loaded)."""
    load()
    C.CDLL("libgomp.so.1").omp_set_num_threads(int(k))


def sph_constants(h):
    out = np.zeros(3, np.float32)
    load().geo_sph_constants(h, _f(out))
    return out


class State:
    """Host particle state for the oracle step; arrays are permuted in place."""

    def __init__(self, st, sph=False, temperature=False):
        n = len(st["x"])
        self.n = n
        self.a = {}
        for f in STATE_F:
            self.a[f] = f32(st[f]).copy() if f in st and st[f] is not None else np.zeros(n, np.float32)
        self.a["flags"] = np.ascontiguousarray(st.get("flags", np.zeros(n)), dtype=np.uint32).copy()
        self.a["id"] = np.ascontiguousarray(st.get("id", np.arange(n)), dtype=np.uint32).copy()
        self.sph = sph
        if sph:
            for f in STATE_SPH:
                dflt = np.ones(n, np.float32) if f == "sph_mass" else np.zeros(n, np.float32)
                self.a[f] = f32(st[f]).copy() if f in st and st[f] is not None else dflt
        self._s = _State()
        self._s.n = n
        for f in STATE_F:
            setattr(self._s, "m" if f == "mass" else f, _f(self.a[f]))
        self._s.flags = _u(self.a["flags"])
        self._s.id = _u(self.a["id"])
        if sph:
            for f in STATE_SPH:
                setattr(self._s, f, _f(self.a[f]))
        self.temperature = temperature
        if temperature:  # ParticlePhysics' temp starts at 288 (particle.h:61)
            self.a["temp"] = f32(st["temp"]).copy() if st.get("temp") is not None else np.full(n, 288.0, np.float32)
            self.a["label"] = np.ascontiguousarray(st.get("label", np.zeros(n)), dtype=np.uint32).copy()
            self.a["ke"] = f32(st["ke"]).copy() if st.get("ke") is not None else np.zeros(n, np.float32)
            self.a["prev_ke"] = f32(st["prev_ke"]).copy() if st.get("prev_ke") is not None else np.zeros(n, np.float32)
            self._s.temp = _f(self.a["temp"])
            self._s.label = _u(self.a["label"])
            self._s.ke = _f(self.a["ke"])
            self._s.prev_ke = _f(self.a["prev_ke"])

    def step(self, prm, nsteps=1):
        p, keep = _make_params(prm)
        rc = load().geo_step(C.byref(self._s), C.byref(p), int(nsteps))
        del keep
        if rc != 0:
            raise RuntimeError(f"oracle step failed {rc}")

    def pcisph(self, prm):
        """geo_sph_pcisph alone (SPH::PCISPH, SPH.cpp:152-281) on the current
        state, in the current order: no integration, no sort."""
        p, keep = _make_params(dict(prm, sph=1))
        rc = load().geo_sph_pcisph(C.byref(self._s), C.byref(p))
        del keep
        if rc != 0:
            raise RuntimeError(f"oracle pcisph failed {rc}")

    def integrate_start(self, dt, sph, sph_max_vel):
        """geo_integrate_start alone (Physics::integrateStart,
        physics.cpp:1069-1112)."""
        load().geo_integrate_start(C.byref(self._s), float(dt), int(sph), float(sph_max_vel))

    def __getitem__(self, k):
        return self.a[k]


def _make_params(prm):
    """geo_params from a dict of step parameters (defaults: the reference's
    UpdateVariables values); returns (params, objects to keep alive)."""
    p = _Params()
    d = dict(G=1.0, theta=0.8, softening=1.0, dt=0.01, naive=0, direct_below=5000, sph=0,
             sph_h=3.0, sph_cell=3.0, sph_unit_mass=0.03, viscosity=0.3, cohesion_coef=1.0,
             delta=19000.0, stiff_mult=1.0, sph_max_vel=250.0, sph_search=0, sph_search_radius=3.5,
             periodic=0, domain_w=3840.0, domain_h=2160.0, sph_ground=0, vertical_gravity=3.0,
             boundary_friction=0.0, temperature=0, heat_conductivity=0.045, ambient_temp=274.0,
             ambient_heat_rate=1.0)
    d.update({k: v for k, v in prm.items() if k in d or k == "dt_start"})
    d.setdefault("dt_start", d["dt"])
    for k, v in d.items():
        setattr(p, k, v)
    arr = None
    mats = prm.get("materials")
    if mats:
        arr = (Material * len(mats))(*mats)
        p.materials = C.addressof(arr)
        p.nmaterials = len(mats)
    return p, arr


class _Domain(C.Structure):
    _fields_ = [("w", C.c_float), ("h", C.c_float)]


def direct_accel_periodic(x, y, m, G, eps, w, h):
    """naiveGravity with the periodic wrap of each separation (physics.cpp:417-420)."""
    x, y, m = f32(x), f32(y), f32(m)
    ax = np.zeros(len(x), np.float32)
    ay = np.zeros(len(x), np.float32)
    lib = load()
    lib.geo_direct_accel_dom.argtypes = [C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.c_float, C.c_float, C.POINTER(_Domain), C.POINTER(C.c_float),
                                         C.POINTER(C.c_float)]
    d = _Domain(w, h)
    lib.geo_direct_accel_dom(len(x), _f(x), _f(y), _f(m), G, eps, C.byref(d), _f(ax), _f(ay))
    return ax, ay


def bh_accel_periodic(x, y, m, nodes, G, theta, eps, w, h):
    x, y, m = f32(x), f32(y), f32(m)
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    ax = np.zeros(len(x), np.float32)
    ay = np.zeros(len(x), np.float32)
    cnt = C.c_uint64()
    lib = load()
    lib.geo_bh_accel_dom.argtypes = [C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float, C.POINTER(_Domain),
                                     C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint64)]
    d = _Domain(w, h)
    lib.geo_bh_accel_dom(len(x), _f(x), _f(y), _f(m), C.c_void_p(nodes.ctypes.data), len(nodes), G, theta, eps,
                         C.byref(d), _f(ax), _f(ay), C.byref(cnt))
    return ax, ay, cnt.value
