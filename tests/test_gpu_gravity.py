"""GPU parity of the gravity path through the C ABI against the oracle.

Integer and index work is bit-exact: bounding box, Morton keys, sort
permutation, and the whole quadtree (ranges, pre-order skip counts, node
geometry, and masses / centres of mass, which follow the reference's rounding
order).  Accelerations use v_rsq_f32 and fused multiply-adds, so they are
compared with a tolerance: per particle |a_gpu - a_ref| <= 2e-5 |a_ref| +
1e-6 max|a_ref|, and the accepted-interaction counts must match exactly
(same tree, same unfused acceptance test).
"""
import math

import numpy as np
import pytest

import oracle as O
from galaxy_step import Context, default_params, configs

pytestmark = pytest.mark.gpu

ACC_RTOL = 2e-5
ACC_ATOL_FRAC = 1e-6


def _params(**kw):
    p = default_params()
    p.G, p.theta, p.softening, p.dt, p.dt_start = 1.0, 0.8, 1.0, 0.01, 0.01
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_sorted(st):
    x, y = st["x"], st["y"]
    bb = O.bbox(x, y)
    keys = O.morton_keys(x, y, bb)
    perm = O.sort_perm(keys)
    return bb, keys, perm


def _assert_acc_close(ax, ay, rx, ry):
    mag = np.hypot(rx, ry)
    tol = ACC_RTOL * mag + ACC_ATOL_FRAC * mag.max()
    err = np.hypot(ax - rx, ay - ry)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (bad[:10], err[bad[:10]], mag[bad[:10]])


@pytest.fixture(scope="module")
def cfg0():
    st, prm = configs.config0()
    return st, prm


def test_bbox_keys_perm_bitexact(cfg0):
    st, _ = cfg0
    n = len(st["x"])
    bb_ref, keys_ref, perm_ref = _oracle_sorted(st)
    with Context(n) as c:
        c.upload(st)
        bb = c.bounding_box()
        np.testing.assert_array_equal(bb, bb_ref)
        c.compute_morton_keys(bb)
        np.testing.assert_array_equal(c.keys(), keys_ref)
        c.sort_by_morton_key()
        np.testing.assert_array_equal(c.perm(), perm_ref)
        d = c.download()
        np.testing.assert_array_equal(d["x"], st["x"][perm_ref])
        np.testing.assert_array_equal(d["id"], st["id"][perm_ref])
        np.testing.assert_array_equal(d["mass"], st["mass"][perm_ref])


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 257, 4999, 5000, 20000])
def test_tree_bitexact_sizes(n):
    st, _ = configs.config0(seed=11, n=max(n, 2))
    st = {k: v[:n] for k, v in st.items()}
    bb, keys, perm = _oracle_sorted(st)
    xs, ys, ms = st["x"][perm], st["y"][perm], st["mass"][perm]
    ref = O.build_tree(xs, ys, ms, bb)
    with Context(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        got = c.nodes()
    assert len(got) == len(ref)
    _assert_nodes_equal(got, ref)


def test_tree_mass_fold_order_known_answer():
    """computeInternalMass's child order q = 0, 2, 1, 3 (quadtree.h:70-94) on
    the device: with 2^24 in quadrant 2 and unit masses elsewhere only that
    order gives the root mass 2^24 (0, 1, 2, 3 gives 2^24 + 4)."""
    st = {"x": np.array([1.0, 7.0, 1.0, 7.0], np.float32), "y": np.array([1.0, 1.0, 7.0, 7.0], np.float32),
          "mass": np.array([1.0, 1.0, 2.0 ** 24, 1.0], np.float32)}
    with Context(4) as c:
        c.upload(st)
        bb = c.bounding_box()
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        got = c.nodes()
    assert len(got) == 5 and got[0]["mass"] == np.float32(2.0 ** 24)
    _assert_nodes_equal(got, O.build_tree(st["x"], st["y"], st["mass"], bb))


def _assert_nodes_equal(got, ref, fields=("start", "end", "next", "child", "posx", "posy", "size",
                                          "mass", "comx", "comy")):
    for f in fields:
        a, b = got[f], ref[f]
        eq = np.all(a == b, axis=tuple(range(1, a.ndim))) if a.ndim > 1 else (a == b)
        if not np.all(eq):
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"field {f}: first mismatch at node {i} of {len(ref)}:\n"
                                 f" got {got[i]}\n want {ref[i]}")


def test_tree_rebuilt_every_frame_matches_oracle():
    # several frames in one context: each frame's tree is rebuilt from scratch
    st, prm = configs.config0(seed=13, n=30000)
    p = _params(theta=prm["theta"], dt=prm["dt"], dt_start=prm["dt"])
    with Context(30000) as c:
        c.upload(st)
        for _ in range(4):
            c.step(p, 1)
            c.sync()
            d = c.download(["x", "y", "mass"])
            bb = O.bbox(d["x"], d["y"])
            ref = O.build_tree(d["x"], d["y"], d["mass"], bb)
            got = c.nodes()
            assert len(got) == len(ref)
            _assert_nodes_equal(got, ref, ("start", "end", "next", "child", "mass", "comx", "comy"))


def _run_gravity(st, prm_kw, theta):
    n = len(st["x"])
    bb, keys, perm = _oracle_sorted(st)
    xs, ys, ms = st["x"][perm], st["y"][perm], st["mass"][perm]
    with Context(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        p = _params(theta=theta, **prm_kw)
        c.compute_gravity(p)
        d = c.download(["ax", "ay"])
        inter = c.interactions()
    return xs, ys, ms, bb, d, inter


@pytest.mark.parametrize("theta", [0.5, 0.8, 1.0])
@pytest.mark.parametrize("compensated", [0, 1])
def test_bh_accel_matches_oracle(cfg0, theta, compensated):
    st, _ = cfg0
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(compensated=compensated), theta)
    nodes = O.build_tree(xs, ys, ms, bb)
    rx, ry, cnt = O.bh_accel(xs, ys, ms, nodes, 1.0, theta, 1.0)
    assert inter == cnt
    _assert_acc_close(d["ax"], d["ay"], rx, ry)


@pytest.mark.parametrize("theta", [0.5, 0.7, 1.0])
def test_bh_accel_million_two_disks(theta):
    """The 1M two-disk collision at each theta of the sweep: accelerations of
    the reference walk against the oracle on 4096 sampled particles, with the
    same total interaction count."""
    st, prm = configs.config1(n=1_000_000)
    xs, ys, ms, bb, d, inter = _run_gravity(st, {}, theta)
    nodes = O.build_tree(xs, ys, ms, bb)
    rng = np.random.default_rng(0)
    t = np.sort(rng.choice(len(xs), 4096, replace=False)).astype(np.uint32)
    rx, ry, cnt = O.bh_accel(xs, ys, ms, nodes, 1.0, theta, 1.0)
    assert inter == cnt
    _assert_acc_close(d["ax"][t], d["ay"][t], rx[t], ry[t])
    # FP64 direct-sum error of the reference algorithm
    fx, fy = O.direct_fp64(xs, ys, ms, 1.0, 1.0, t)
    e = np.hypot(d["ax"][t] - fx, d["ay"][t] - fy) / np.hypot(fx, fy)
    rms = math.sqrt(np.mean(e ** 2))
    print(f"BH theta={theta} RMS rel force error vs FP64 direct: {rms:.3e}, interactions/particle "
          f"{inter / len(xs):.1f}")
    assert rms < 0.1 * theta ** 2


@pytest.mark.parametrize("n", [3, 100, 1000, 4999])
def test_direct_sum_small_n(n):
    st, _ = configs.config0(seed=5, n=max(n, 3))
    st = {k: v[:n] for k, v in st.items()}
    bb, keys, perm = _oracle_sorted(st)
    xs, ys, ms = st["x"][perm], st["y"][perm], st["mass"][perm]
    with Context(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        c.compute_gravity(_params())  # N < 5000 -> direct sum, softening unsquared
        d = c.download(["ax", "ay"])
    rx, ry = O.direct_accel(xs, ys, ms, 1.0, 1.0)
    _assert_acc_close(d["ax"], d["ay"], rx, ry)


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


@pytest.mark.parametrize("n,steps", [(20000, 10), (3000, 10)])
def test_step_trajectory_matches_oracle(n, steps):
    st, prm = configs.config0(seed=7, n=n)
    p = _params(theta=prm["theta"], dt=prm["dt"], dt_start=prm["dt"])
    s = O.State(st)
    s.step(dict(prm, dt_start=prm["dt"]), nsteps=steps)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "vx", "vy", "ax", "ay", "id")})
    with Context(n) as c:
        c.upload(st)
        c.step(p, steps)
        c.sync()
        got = _by_id(c.download(["x", "y", "vx", "vy", "ax", "ay", "id"]))
    ext = float(np.ptp(ref["x"]))
    assert np.abs(got["x"] - ref["x"]).max() < 1e-4 * ext
    assert np.abs(got["y"] - ref["y"]).max() < 1e-4 * ext
    vs = float(np.abs(ref["vx"]).max())
    assert np.abs(got["vx"] - ref["vx"]).max() < 1e-3 * vs


def test_upload_edge_cases():
    with Context(10) as c:
        c.upload({"x": np.zeros(0), "y": np.zeros(0), "mass": np.zeros(0)})
        c.step(_params(), 3)
        assert c.count() == 0
        # one particle: no force, drifts with its velocity
        c.upload({"x": np.array([1.0]), "y": np.array([2.0]), "vx": np.array([1.0]),
                  "mass": np.array([1.0])})
        c.step(_params(dt=0.5, dt_start=0.5), 2)
        d = c.download(["x", "y", "ax"])
        assert d["x"][0] == pytest.approx(2.0) and d["ax"][0] == 0.0
        with pytest.raises(Exception):
            c.upload({"x": np.zeros(11), "y": np.zeros(11), "mass": np.ones(11)})


def test_coincident_particles():
    # all particles at one point: the root is a leaf (size 0 <= 1e-6)
    n = 6000
    st = {"x": np.full(n, 3.0), "y": np.full(n, -2.0), "mass": np.ones(n)}
    with Context(n) as c:
        c.upload(st)
        bb = c.bounding_box()
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        assert c.node_count() == 1
        c.compute_gravity(_params())
        d = c.download(["ax", "ay"])
    assert np.all(d["ax"] == 0) and np.all(d["ay"] == 0)


def test_unsupported_flags_rejected():
    with Context(100) as c:
        c.upload({"x": np.arange(100.0), "y": np.zeros(100), "mass": np.ones(100)})
        with pytest.raises(Exception):
            c.step(_params(temperature=1), 1)
        with pytest.raises(Exception):  # a periodic domain needs a size
            c.step(_params(periodic=1, domain_w=0.0), 1)


@pytest.mark.parametrize("n,direct_below", [(20000, 5000), (3000, 5000)])
def test_periodic_trajectory_matches_oracle(n, direct_below):
    # a disk near the corner of a periodic domain: particles and forces wrap
    # across both seams (tree walk for 20000, the direct sum below 5000)
    W, H = 900.0, 700.0
    st, prm = configs.config0(seed=33, n=n, scale=60.0)
    st = dict(st)
    st["x"] = np.mod(st["x"], W).astype(np.float32)
    st["y"] = np.mod(st["y"], H).astype(np.float32)
    p = _params(theta=prm["theta"], dt=prm["dt"], dt_start=prm["dt"], periodic=1, domain_w=W, domain_h=H,
                direct_below=direct_below)
    s = O.State(st)
    s.step(dict(prm, dt_start=prm["dt"], periodic=1, domain_w=W, domain_h=H, direct_below=direct_below), nsteps=5)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "vx", "vy", "id")})
    with Context(n) as c:
        c.upload(st)
        c.step(p, 5)
        c.sync()
        got = _by_id(c.download(["x", "y", "vx", "vy", "id"]))
    assert np.all((got["x"] >= 0) & (got["x"] < W) & (got["y"] >= 0) & (got["y"] < H))
    dx = np.abs(got["x"] - ref["x"])
    dx = np.minimum(dx, W - dx)  # a particle on the seam may sit on either side
    assert dx.max() < 1e-3, dx.max()
    assert np.abs(got["vx"] - ref["vx"]).max() < 1e-3 * float(np.abs(ref["vx"]).max())


def _fp64_errors(d, xs, ys, ms, t):
    fx, fy = O.direct_fp64(xs, ys, ms, 1.0, 1.0, t)
    err = np.hypot(d["ax"][t] - fx, d["ay"][t] - fy)
    mag = np.hypot(fx, fy)
    rel = err / mag
    return math.sqrt(np.mean(rel ** 2)), math.sqrt(np.mean(err ** 2) / np.mean(mag ** 2))


def test_accuracy_mode_two_disks_theta07():
    # quadrupole moments + offset-corrected opening test, checked against the
    # FP64 direct sum on 4096 sampled targets
    st, _ = configs.config1(n=1_000_000)
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=1), 0.7)
    t = np.sort(np.random.default_rng(0).choice(len(xs), 4096, replace=False)).astype(np.uint32)
    rms_rel, rms_norm = _fp64_errors(d, xs, ys, ms, t)
    _, _, _, _, d0, _ = _run_gravity(st, dict(accuracy=0), 0.7)
    rms_rel0, rms_norm0 = _fp64_errors(d0, xs, ys, ms, t)
    print(f"theta=0.7 accuracy mode: per-particle RMS {rms_rel:.3e}, force-normalised RMS {rms_norm:.3e}; "
          f"reference walk: {rms_rel0:.3e}, {rms_norm0:.3e}; interactions/particle {inter / len(xs):.1f}")
    assert rms_norm < 1e-3
    assert rms_rel < 3e-3
    assert rms_rel < 0.25 * rms_rel0


def test_hexadecapole_mode_two_disks_theta07():
    # accuracy 2: moments to the hexadecapole; the per-particle RMS error vs
    # the FP64 direct sum at theta = 0.7 must stay below 1e-3
    st, _ = configs.config1(n=1_000_000)
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=2), 0.7)
    t = np.sort(np.random.default_rng(1).choice(len(xs), 4096, replace=False)).astype(np.uint32)
    rms_rel, rms_norm = _fp64_errors(d, xs, ys, ms, t)
    print(f"theta=0.7 hexadecapole mode: per-particle RMS {rms_rel:.3e}, force-normalised RMS {rms_norm:.3e}")
    assert rms_rel < 1e-3
    assert rms_norm < 2e-4


@pytest.mark.parametrize("accuracy", [1, 2, 3])
def test_multipole_modes_converge_to_direct(accuracy):
    # theta -> 0 opens every node: the walk becomes the direct sum
    st, _ = configs.config0(seed=31, n=5000)
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=accuracy), 0.02)
    t = np.arange(len(xs), dtype=np.uint32)
    rms_rel, _ = _fp64_errors(d, xs, ys, ms, t)
    assert rms_rel < 2e-5


def test_hexadecapole_beats_quadrupole_per_node_order():
    # same tree and opening test: the higher order must cut the error
    st, _ = configs.config0(seed=32, n=40000)
    t = np.arange(0, 40000, 7, dtype=np.uint32)
    errs = {}
    for acc in (0, 1, 3, 2):  # monopole (reference), quadrupole, octupole, hexadecapole
        xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=acc), 0.7)
        errs[acc] = _fp64_errors(d, xs, ys, ms, t)[0]
    assert errs[2] < 0.7 * errs[3] < 0.7 * errs[1] < 0.5 * errs[0], errs


def test_accuracy_mode_disk_per_particle():
    st, _ = configs.config0()
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=1), 0.7)
    t = np.arange(0, len(xs), 5, dtype=np.uint32)
    rms_rel, rms_norm = _fp64_errors(d, xs, ys, ms, t)
    assert rms_rel < 1e-3 and rms_norm < 1e-4


def test_accuracy_mode_small_theta_is_direct():
    st, _ = configs.config0(seed=21, n=6000)
    xs, ys, ms, bb, d, inter = _run_gravity(st, dict(accuracy=1), 0.05)
    t = np.arange(len(xs), dtype=np.uint32)
    rms_rel, _ = _fp64_errors(d, xs, ys, ms, t)
    assert rms_rel < 2e-5


def test_tree_bitexact_at_full_size_after_evolution():
    # the 1M two-disk collision after 12 frames: every node field of the
    # device tree equals the oracle's build on the device's particle order
    st, _ = configs.config1(n=1_000_000)
    p = _params(theta=0.7, dt=0.005, dt_start=0.005, direct_below=0)
    with Context(1_000_000) as c:
        c.upload(st)
        c.step(p, 12)
        c.sync()
        d = c.download(["x", "y", "mass"])
        got = c.nodes()
    bb = O.bbox(d["x"], d["y"])
    ref = O.build_tree(d["x"], d["y"], d["mass"], bb)
    assert len(got) == len(ref)
    _assert_nodes_equal(got, ref)


def _midline_set(seed=5, ncl=40000, nmid=24000, levels=9):
    """Particles within three float spacings of the midlines of the node
    squares of the first `levels` levels (subGridMaker's float arithmetic),
    plus a dense cluster that keeps the top levels big and two corners that
    fix the bounding box.  Their float quadrant tests disagree with their
    Morton digits at every one of those levels (hundreds per level)."""
    f32 = np.float32
    rng = np.random.default_rng(seed)
    corners = np.array([[-1000.3, -800.7], [1500.9, 1200.1]], np.float32)
    cl = (rng.normal(0, 30, (ncl, 2)) + np.array([300.0, 250.0])).astype(np.float32)
    pts = np.vstack([corners, cl])
    bb = O.bbox(pts[:, 0], pts[:, 1])
    x0, y0, size = f32(bb[0]), f32(bb[1]), f32(bb[2])
    mids = ([], [])
    pos, s = [(x0, y0)], size
    for _ in range(levels):
        half = f32(s * f32(0.5))
        mids[0].extend(f32(px + half) for px, _ in pos)
        mids[1].extend(f32(py + half) for _, py in pos)
        pos = [(f32(px + (half if q & 1 else f32(0))), f32(py + (half if q & 2 else f32(0))))
               for px, py in pos for q in range(4)]
        s = half
    mx, my = np.unique(np.array(mids[0], np.float32)), np.unique(np.array(mids[1], np.float32))
    out = np.zeros((nmid, 2), np.float32)
    for k in range(nmid):
        ax = int(rng.integers(2))
        line = mx if ax == 0 else my
        v = line[rng.integers(len(line))]
        off = int(rng.integers(-3, 4))
        for _ in range(abs(off)):
            v = np.nextafter(v, f32(np.inf) if off > 0 else f32(-np.inf), dtype=np.float32)
        other = f32(rng.uniform(-900, 1400) if ax == 0 else rng.uniform(-700, 1100))
        out[k] = (v, other) if ax == 0 else (other, v)
    pts = np.vstack([pts, out])
    n = len(pts)
    z = np.zeros(n, np.float32)
    return {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "vx": z, "vy": z, "mass": np.ones(n, np.float32),
            "id": np.arange(n, dtype=np.uint32)}


@pytest.mark.parametrize("leaf_capacity", [0, 32])
def test_tree_bitexact_particles_on_midlines(leaf_capacity):
    # the frame's tree (whose top levels split Morton-prefix runs, checked
    # against the particles' float quadrant tests) equals the oracle's on
    # particles sitting on and next to the node midlines of nine levels
    st = _midline_set()
    n = len(st["x"])
    p = _params(theta=0.7, dt=0.001, dt_start=0.001, direct_below=0, leaf_capacity=leaf_capacity)
    with Context(n) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        d = c.download(["x", "y", "mass"])
        got = c.nodes()
    np.testing.assert_array_equal(np.sort(d["x"]), np.sort(st["x"]))  # zero velocities: nothing moved
    bb = O.bbox(d["x"], d["y"])
    ref = O.build_tree(d["x"], d["y"], d["mass"], bb, leaf_capacity)
    assert len(got) == len(ref)
    _assert_nodes_equal(got, ref)
