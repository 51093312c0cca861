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
from galaxy_step.lib import GE_FLAG_HOT_POINT, Context

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
