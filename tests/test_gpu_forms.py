"""GPU tests of the execution forms built on the single-system step:
hipGraph replay of frames, autotunable leaf capacity and walk block size,
and continuous batching (systems leaving and joining a resident batch).

Graph replays, block sizes and batch membership change how work is launched,
not what any particle computes, so each is held bit-identical to the plain
eager single-system path.  Leaf-capacity trees are held bit-exact to the
oracle's build with the same leaf rule.
"""
import numpy as np
import pytest

import oracle as O
from galaxy_step import Context, configs, default_params
from test_gpu_gravity import _assert_acc_close

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "vx", "vy", "ax", "ay", "id"]


def _params(**kw):
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _eq(a, b, msg=""):
    for f in a:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f"{msg} field {f}")


@pytest.mark.parametrize("accuracy", [0, 1, 2])
def test_graph_replay_bitidentical_to_eager(accuracy):
    st, _ = configs.config0(seed=21, n=20000)
    n = len(st["x"])
    pe = _params(accuracy=accuracy)
    pg = _params(accuracy=accuracy, use_graph=1)
    with Context(n) as e, Context(n) as g:
        e.upload(st)
        g.upload(st)
        # 7 frames: one eager, three replays of the captured pair; then 5 and
        # 4 more (the same graph replayed, with eager frames between)
        for k, steps in enumerate([7, 5, 4, 3]):
            e.step(pe, steps)
            g.step(pg, steps)
            e.sync()
            g.sync()
            _eq(g.download(FIELDS), e.download(FIELDS), f"call {k}")


def test_graph_recaptures_when_params_change():
    st, _ = configs.config0(seed=22, n=6000)
    n = len(st["x"])
    with Context(n) as e, Context(n) as g:
        e.upload(st)
        g.upload(st)
        for dt in (0.01, 0.02, 0.01):
            e.step(_params(dt=dt, dt_start=dt), 5)
            g.step(_params(dt=dt, dt_start=dt, use_graph=1), 5)
        e.sync()
        g.sync()
        _eq(g.download(FIELDS), e.download(FIELDS))


def test_graph_replay_sph_bitidentical():
    st, _ = configs.config2(n_target=3000, n_impactor=800, r_target=40.0, r_impactor=16.0)
    n = len(st["x"])
    kw = dict(G=1.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, sph=1, sph_h=4.0, sph_cell=4.0)
    with Context(n, sph=True) as e, Context(n, sph=True) as g:
        e.upload(st)
        g.upload(st)
        e.step(_params(**kw), 7)
        g.step(_params(use_graph=1, **kw), 7)
        e.sync()
        g.sync()
        f = FIELDS + ["pred_dens", "press"]
        _eq(g.download(f), e.download(f))


def test_graph_replay_batch_bitidentical():
    sizes = [3000, 17, 900, 5000]
    systems = [configs.config3_system(k, m, 60.0 + 20 * k, seed=4)[0] for k, m in enumerate(sizes)]
    batch = {k: np.concatenate([s[k] for s in systems]) for k in ("x", "y", "vx", "vy", "mass")}
    n = len(batch["x"])
    with Context(n, max_systems=4) as e, Context(n, max_systems=4) as g:
        e.upload(batch, counts=sizes)
        g.upload(batch, counts=sizes)
        e.step(_params(), 6)
        g.step(_params(use_graph=1), 6)
        e.sync()
        g.sync()
        _eq(g.download(FIELDS), e.download(FIELDS))


@pytest.mark.parametrize("cap", [8, 16, 32])
def test_leaf_capacity_tree_matches_oracle(cap):
    st, prm = configs.config0(seed=23, n=30000)
    p = _params(theta=prm["theta"], leaf_capacity=cap)
    with Context(30000) as c:
        c.upload(st)
        for _ in range(2):
            c.step(p, 1)
            c.sync()
            d = c.download(["x", "y", "mass"])
            bb = O.bbox(d["x"], d["y"])
            ref = O.build_tree(d["x"], d["y"], d["mass"], bb, leaf_cap=cap)
            got = c.nodes()
            assert len(got) == len(ref)
            for f in ("start", "end", "next", "mass", "comx", "comy"):
                np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
            np.testing.assert_array_equal(got["child"], ref["child"])


@pytest.mark.parametrize("cap", [8, 32])
def test_leaf_capacity_forces_match_oracle_walk(cap):
    # the walk over a leaf-capacity tree against the oracle's walk of the same
    # tree: same accepted nodes, leaf sums over up to cap partners
    st, _ = configs.config0(seed=24, n=12000)
    with Context(12000) as c:
        c.upload(st)
        p = _params(theta=0.7, leaf_capacity=cap, compensated=0)
        c.step(p, 1)
        c.sync()
        d = c.download(["x", "y", "mass"])
        c.compute_gravity(p)
        g = c.download(["ax", "ay"])
    bb = O.bbox(d["x"], d["y"])
    nodes = O.build_tree(d["x"], d["y"], d["mass"], bb, leaf_cap=cap)
    ax, ay, _ = O.bh_accel(d["x"], d["y"], d["mass"], nodes, 1.0, 0.7, 1.0)
    _assert_acc_close(g["ax"], g["ay"], ax, ay)


@pytest.mark.parametrize("block", [64, 128])
def test_walk_block_size_bitidentical(block):
    st, _ = configs.config0(seed=25, n=15000)
    with Context(15000) as a, Context(15000) as b:
        a.upload(st)
        b.upload(st)
        a.step(_params(accuracy=1), 3)
        b.step(_params(accuracy=1, walk_block=block), 3)
        a.sync()
        b.sync()
        _eq(b.download(FIELDS), a.download(FIELDS))


def _alone(st, p, steps):
    n = len(st["x"])
    with Context(n) as a:
        a.upload(st)
        a.step(p, steps)
        a.sync()
        return a.download(FIELDS)


def test_systems_update_keeps_results_bitidentical():
    sizes = [4000, 700, 2500, 9000]
    systems = []
    for k, m in enumerate(sizes):
        s, _ = configs.config3_system(k, m, 50.0 + 30 * k, seed=6)
        s["id"] = np.arange(m, dtype=np.uint32)
        systems.append(s)
    keys = ("x", "y", "vx", "vy", "mass", "id")
    cat = lambda ss: {k: np.concatenate([s[k] for s in ss]) for k in keys}
    p = _params(accuracy=1)
    with Context(20000, max_systems=8) as c:
        c.upload(cat(systems[:3]), counts=sizes[:3])
        c.step(p, 3)
        # system 1 leaves, system 3 joins; 0 and 2 stay resident (2 moves)
        c.systems_update([0, 2], cat(systems[3:]), [sizes[3]])
        np.testing.assert_array_equal(c.system_offsets(), [0, 4000, 6500, 15500])
        c.step(p, 4)
        c.sync()
        got = [c.download_system(t, FIELDS) for t in range(3)]
        # a system joins while the others are mid-run; then the first leaves
        c.systems_update([1, 2], cat([systems[1]]), [sizes[1]])
        c.step(p, 2)
        c.sync()
        late = c.download_system(2, FIELDS)
    _eq(got[0], _alone(systems[0], p, 7), "kept system 0")
    _eq(got[1], _alone(systems[2], p, 7), "moved system 2")
    _eq(got[2], _alone(systems[3], p, 4), "joined system 3")
    _eq(late, _alone(systems[1], p, 2), "re-joined system 1")


def test_systems_update_rejects_bad_requests():
    with Context(1000, max_systems=2) as c:
        c.upload({"x": np.arange(10.0), "y": np.zeros(10), "mass": np.ones(10)}, counts=[4, 6])
        with pytest.raises(Exception):
            c.systems_update([1, 0])          # not ascending
        with pytest.raises(Exception):
            c.systems_update([2])             # no such system
        with pytest.raises(Exception):        # above max_systems
            c.systems_update([0, 1], {"x": np.zeros(3), "y": np.zeros(3), "mass": np.ones(3)}, [3])
        c.systems_update([1])
        assert c.count() == 6
        c.systems_update([])
        assert c.count() == 0


def test_direct_fp64_matches_oracle():
    st, _ = configs.config0(seed=26, n=8000)
    p = _params()
    rng = np.random.default_rng(3)
    with Context(8000) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        d = c.download(["x", "y", "mass"])
        t = rng.choice(8000, 300, replace=False).astype(np.uint32)
        gx, gy = c.direct_fp64(t, p)
    rx, ry = O.direct_fp64(d["x"], d["y"], d["mass"], 1.0, 1.0, t)
    mag = np.hypot(rx, ry)
    assert np.all(np.hypot(gx - rx, gy - ry) <= 1e-11 * mag + 1e-14 * mag.max())


def test_diagnostics_match_host_fp64():
    st, _ = configs.config0(seed=27, n=3000)
    p = _params()
    with Context(3000) as c:
        c.upload(st)
        c.step(p, 2)
        c.sync()
        d = c.download(["x", "y", "vx", "vy", "mass"])
        g = c.diagnostics(0, p, potential=True)
    x, y, vx, vy, m = (d[k].astype(np.float64) for k in ("x", "y", "vx", "vy", "mass"))
    want = {"mass": m.sum(), "px": (m * vx).sum(), "py": (m * vy).sum(), "lz": (m * (x * vy - y * vx)).sum(),
            "kinetic": 0.5 * (m * (vx * vx + vy * vy)).sum(), "comx": (m * x).sum() / m.sum(),
            "comy": (m * y).sum() / m.sum()}
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    inv = 1.0 / np.sqrt(dx * dx + dy * dy + 1.0)
    np.fill_diagonal(inv, 0.0)
    want["potential"] = -0.5 * float(m @ inv @ m)
    for k, v in want.items():
        assert g[k] == pytest.approx(v, rel=1e-10, abs=1e-9 * abs(want["kinetic"]) if k in ("px", "py", "lz") else 0), k


# a record shaped like the reference's ParticlePhysics (particle.h:5-73);
# field names that match ge_particles are transferred, the rest are not
PARTICLE_PHYSICS = np.dtype([("x", "f4"), ("y", "f4"), ("vx", "f4"), ("vy", "f4"), ("ax", "f4"), ("ay", "f4"),
                             ("ppx", "f4"), ("ppy", "f4"), ("pvx", "f4"), ("pvy", "f4"), ("predvx", "f4"),
                             ("predvy", "f4"), ("pfx", "f4"), ("pfy", "f4"), ("morton", "u8"), ("mass", "f4"),
                             ("press", "f4"), ("press_tmp", "f4"), ("dens", "f4"), ("pred_dens", "f4"),
                             ("sph_mass", "f4"), ("rest_dens", "f4"), ("stiff", "f4"), ("visc", "f4"),
                             ("cohesion", "f4"), ("temp", "f4"), ("ke", "f4"), ("prev_ke", "f4"), ("id", "u4"),
                             ("neighbor_offset", "u4"), ("hot", "u1"), ("solid", "u1")], align=True)


def test_aos_transfer_matches_soa():
    st, _ = configs.config0(seed=28, n=7000)
    n = len(st["x"])
    rec = np.zeros(n, PARTICLE_PHYSICS)
    for f in ("x", "y", "vx", "vy", "mass", "id"):
        rec[f] = st[f]
    rec["temp"] = 300.0
    p = _params()
    with Context(n) as a, Context(n) as b:
        a.upload(st)
        b.upload_aos(rec, flags=st["flags"])
        a.step(p, 3)
        b.step(p, 3)
        a.sync()
        b.sync()
        da = a.download(FIELDS)
        db, fl = b.download_aos(PARTICLE_PHYSICS, with_flags=True)
    for f in FIELDS:
        np.testing.assert_array_equal(db[f], da[f], err_msg=f)
    assert np.all(db["temp"] == 0.0) and np.all(db["morton"] == 0)  # fields outside the path untouched
    np.testing.assert_array_equal(fl, np.zeros(n, np.uint32))
