"""GPU parity of the particle constraints (include/ge_constraints.h) against
the oracle's serial loop in the device's (colour, index) order: positions,
accelerations and every constraint field bit-exact over frames that remove
broken constraints and apply the triangle rule; the colouring is checked to
be proper; a ge_step frame with constraints conserves momentum and replays
identically from a hipGraph."""
import numpy as np
import pytest

import oracle as O
from galaxy_step import constraints as GC
from galaxy_step import default_params
from galaxy_step import thermal as T
from galaxy_step.lib import GE_FLAG_HOT_POINT, GE_FLAG_SOLIDIFIED, Context

pytestmark = pytest.mark.gpu

F = np.float32


def _lattice(side, seed, spacing=1.0):
    """A triangulated lattice: right, up and up-right springs; random ids."""
    rng = np.random.default_rng(seed)
    n = side * side
    gx, gy = np.meshgrid(np.arange(side), np.arange(side), indexing="xy")
    x = (gx.ravel() * spacing + rng.normal(0, 0.08, n)).astype(F)
    y = (gy.ravel() * spacing + rng.normal(0, 0.08, n)).astype(F)
    ids = rng.permutation(np.arange(100, 100 + 3 * n, 3)).astype(np.uint32)
    rows = []
    for j in range(side):
        for i in range(side):
            a = j * side + i
            for di, dj in ((1, 0), (0, 1), (1, 1)):
                if i + di < side and j + dj < side:
                    rows.append((a, (j + dj) * side + i + di))
    rows = np.array(rows)
    rng.shuffle(rows)
    cons = np.zeros(len(rows), GC.CONSTRAINT_DTYPE)
    cons["id1"], cons["id2"] = ids[rows[:, 0]], ids[rows[:, 1]]
    d = np.hypot(x[rows[:, 1]] - x[rows[:, 0]], y[rows[:, 1]] - y[rows[:, 0]]).astype(F)
    rest = (d * rng.uniform(0.9, 1.1, len(rows))).astype(F)
    cons["rest_length"] = rest
    cons["original_length"] = d
    cons["stiffness"] = rng.uniform(20, 60, len(rows)).astype(F)
    cons["resistance"] = rng.uniform(0.05, 0.3, len(rows)).astype(F)
    cons["plasticity_point"] = (cons["resistance"] * F(0.5)).astype(F)
    st = {"x": x, "y": y, "vx": rng.normal(0, 0.5, n).astype(F), "vy": rng.normal(0, 0.5, n).astype(F),
          "ax": rng.normal(0, 0.1, n).astype(F), "ay": rng.normal(0, 0.1, n).astype(F),
          "mass": rng.uniform(0.5, 2.0, n).astype(F), "id": ids,
          "flags": np.where(rng.random(n) < 0.02, GE_FLAG_HOT_POINT, 0).astype(np.uint32),
          "label": rng.choice([0, 1, 3, 5], n).astype(np.uint32)}  # none, water, iron, soil
    return st, cons


def _params(**kw):
    p = default_params(G=0.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0, constraints=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _context(st, cons, n):
    c = Context(n, sph=True, temperature=True, max_constraints=len(cons) + 16)
    c.upload(st)
    T.thermal_upload(c, st)
    GC.upload(c, cons)
    return c


def _proper(cons, colors, st):
    idx = {int(v): k for k, v in enumerate(st["id"])}
    for q in np.unique(colors):
        sel = cons[colors == q]
        ends = [idx[int(a)] for a in sel["id1"]] + [idx[int(b)] for b in sel["id2"]]
        assert len(ends) == len(set(ends)), q


def _cmp(got, want, fields):
    for f in fields:
        np.testing.assert_array_equal(got[f], want[f], f)


def test_solve_frames_bitexact():
    st, cons = _lattice(48, 1)
    n = len(st["x"])
    mats = O.reference_materials()
    p = _params()
    with _context(st, cons, n) as c:
        colors, ncol = GC.colors(c, len(cons))
        assert 3 <= ncol <= 20 and np.all(colors < ncol)
        _proper(cons, colors, st)
        ref_st, ref_c, ref_col, fc = dict(st), cons.copy(), colors.copy(), 0
        removed = 0
        for frame in range(7):
            GC.solve(c, p)
            got_c = GC.download(c)
            got = c.download(["x", "y", "ax", "ay"])
            ref_st, ref_c, ref_col, fc = O.constraints(ref_st, ref_c, ref_col, materials=mats, frame_count=fc)
            _cmp(got_c, ref_c, GC.CONSTRAINT_DTYPE.names)
            _cmp(got, ref_st, ("x", "y", "ax", "ay"))
            removed = len(cons) - len(got_c)
        assert removed > 0                 # breaks happened and were removed
        assert ref_c["is_plastic"].any()


def test_step_with_constraints_conserves_momentum_and_replays():
    st, cons = _lattice(40, 2)
    st["label"][:] = 0                      # no materials: nothing breaks
    st["flags"][:] = 0
    st["ax"][:] = 0                         # the first kick then adds no momentum
    st["ay"][:] = 0
    n = len(st["x"])
    out = []
    for graph in (0, 1):
        with _context(st, cons, n) as c:
            p = _params(use_graph=graph)
            m = st["mass"].astype(np.float64)
            p0 = np.array([(m * st["vx"]).sum(), (m * st["vy"]).sum()])
            for _ in range(2):
                c.step(p, 5)                # the first frame eager, then graph pairs
                c.sync()
            d = c.download(["x", "y", "vx", "vy", "mass", "id"])
            o = np.argsort(d["id"])
            out.append({k: v[o] for k, v in d.items()})
            p1 = np.array([(d["mass"] * d["vx"]).astype(np.float64).sum(), (d["mass"] * d["vy"]).astype(np.float64).sum()])
            scale = np.abs(m * st["vx"]).sum()
            assert np.all(np.abs(p1 - p0) < 1e-4 * scale)
            assert len(GC.download(c)) == len(cons)
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_array_equal(out[0][k], out[1][k], k)


def test_constraint_rejections():
    st, cons = _lattice(8, 3)
    n = len(st["x"])
    with Context(n, max_constraints=10) as c:
        c.upload(st)
        with pytest.raises(Exception):      # above max_constraints
            GC.upload(c, cons)
        GC.upload(c, cons[:10])
        GC.clear(c)
        assert len(GC.download(c)) == 0


# ---- createConstraints' automatic branch (physics.cpp:654-766) -------------

def _solid_lattice(side, seed, frac=0.35):
    rng = np.random.default_rng(seed)
    n = side * side
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel() * 1.6 + (gy.ravel() % 2) * 0.8 + rng.normal(0, 0.12, n) - 30).astype(F)
    y = (gy.ravel() * 1.39 + rng.normal(0, 0.12, n) - 20).astype(F)
    ids = rng.permutation(n).astype(np.uint32) * 2 + 11
    flags = (np.uint32(2) | np.where(rng.random(n) < frac, np.uint32(GE_FLAG_SOLIDIFIED), np.uint32(0))).astype(np.uint32)
    st = {"x": x, "y": y, "vx": rng.normal(0, 0.2, n).astype(F), "vy": rng.normal(0, 0.2, n).astype(F),
          "mass": np.ones(n, F), "id": ids, "flags": flags, "sph_mass": np.ones(n, F),
          "rest_dens": np.full(n, 0.008, F), "stiff": np.ones(n, F), "visc": np.ones(n, F),
          "cohesion": np.ones(n, F), "label": rng.integers(0, 8, n).astype(np.uint32),
          "temp": np.full(n, 250.0, F)}
    return st


CREATE_FIELDS = ("id1", "id2", "rest_length", "original_length", "stiffness", "resistance", "displacement",
                 "plasticity_point", "is_broken", "is_plastic")


@pytest.mark.parametrize("search", [0, 1])
def test_create_constraints_stage_matches_oracle(search):
    st = _solid_lattice(56, 5)
    n = len(st["x"])
    _, pre = _lattice(6, 9)  # constraints already present: creation appends after them
    mats = O.reference_materials()
    p = _params(sph_search=search, sph_cell=3.0, sph_search_radius=3.5)
    with Context(n, sph=True, temperature=True, max_constraints=len(pre) + 20 * n) as c:
        c.upload(st)
        T.thermal_upload(c, st)
        GC.upload(c, pre)
        created = GC.create(c, p)
        got = GC.download(c)
        flags = c.download(["flags"])["flags"]
    want_flags, want = O.create_constraints(st, mats, mode=search, cell=3.0, radius=3.5)
    assert created == len(want) > 1000
    assert len(got) == len(pre) + created
    _cmp(got[:len(pre)], pre, CREATE_FIELDS)
    _cmp(got[len(pre):], want, CREATE_FIELDS)
    np.testing.assert_array_equal(flags, want_flags)


def test_create_constraints_capacity_error():
    st = _solid_lattice(20, 6, frac=1.0)
    n = len(st["x"])
    with Context(n, sph=True, temperature=True, max_constraints=50) as c:
        c.upload(st)
        T.thermal_upload(c, st)
        with pytest.raises(Exception):
            GC.create(c, _params())
        assert len(GC.download(c)) == 50  # what fitted stays


def test_frame_creates_constraints_at_the_reference_point():
    """One ge_step frame with temperature and constraints: the constraints it
    creates are the oracle's creation on the state at createConstraints'
    point of updateScene (after integrateStart and the Morton sort), found
    with the stage calls on a second context."""
    st = _solid_lattice(48, 7)
    n = len(st["x"])
    mats = O.reference_materials()
    p = _params(G=1.0, sph=1, sph_h=3.0, sph_cell=3.0, temperature=1, dt=0.004, dt_start=0.004)
    with Context(n, sph=True, temperature=True, max_constraints=20 * n) as c:
        c.upload(st)
        T.thermal_upload(c, st)
        c.step(p, 1)
        c.sync()
        got = GC.download(c)
    with Context(n, sph=True, temperature=True, max_constraints=8) as c:
        c.upload(st)
        T.thermal_upload(c, st)
        c.integrate_start(p)
        bb = c.bounding_box()
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        at = c.download(["x", "y", "id", "flags"])
        at.update(T.thermal_download(c))
    _, want = O.create_constraints({k: at[k] for k in ("x", "y", "id", "flags", "label")}, mats)
    assert len(want) > 500 and len(got) == len(want)
    # the frame's solve moved no end of a fresh constraint past breaking
    _cmp(got, want, ("id1", "id2", "original_length", "stiffness", "resistance", "plasticity_point"))
