"""GPU parity of the 3D gravity path (include/ge_physics3d.h) against the 3D
oracle.

Integer and index work is bit-exact: bounding box, 63-bit Morton keys, the
sort permutation and every octree node field (ranges, pre-order skip counts,
octant children, corners and sizes, masses and centres of mass in the
reference's 0,4,2,6,1,5,3,7 rounding order).  Accelerations use v_rsq_f32 and
fused multiply-adds, so per particle |a_gpu - a_ref| <= 2e-5 |a_ref| +
1e-6 max|a_ref|, and the accepted-interaction counts match exactly (same
tree, same unfused opening test with the softening added unsquared).
"""
import math

import numpy as np
import pytest

import oracle3d as O3
from galaxy_step import configs
from galaxy_step.lib3d import Context3D, default_params3

pytestmark = pytest.mark.gpu

ACC_RTOL = 2e-5
ACC_ATOL_FRAC = 1e-6


def _params(**kw):
    p = default_params3()
    p.G, p.theta, p.softening, p.dt, p.dt_start = 1.0, 0.8, 1.0, 0.01, 0.01
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_sorted(st):
    bb = O3.bbox(st["x"], st["y"], st["z"])
    keys = O3.morton_keys(st["x"], st["y"], st["z"], bb)
    perm = O3.sort_perm(keys)
    return bb, keys, perm


def _assert_acc_close(a, r):
    mag = np.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2)
    tol = ACC_RTOL * mag + ACC_ATOL_FRAC * mag.max()
    err = np.sqrt((a[0] - r[0]) ** 2 + (a[1] - r[1]) ** 2 + (a[2] - r[2]) ** 2)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (bad[:10], err[bad[:10]], mag[bad[:10]])


NODE_FIELDS = ("start", "end", "next", "child", "posx", "posy", "posz", "size", "mass", "comx", "comy", "comz")


def _assert_nodes_equal(got, ref, fields=NODE_FIELDS):
    assert len(got) == len(ref), (len(got), len(ref))
    for f in fields:
        a, b = got[f], ref[f]
        eq = np.all(a == b, axis=tuple(range(1, a.ndim))) if a.ndim > 1 else (a == b)
        if not np.all(eq):
            i = int(np.nonzero(~eq)[0][0])
            raise AssertionError(f"field {f}: first mismatch at node {i} of {len(ref)}:\n"
                                 f" got {got[i]}\n want {ref[i]}")


@pytest.fixture(scope="module")
def gal():
    return configs.galaxy3d_small(n=20000)


def test_bbox_keys_perm_bitexact(gal):
    st, _ = gal
    n = len(st["x"])
    bb_ref, keys_ref, perm_ref = _oracle_sorted(st)
    with Context3D(n) as c:
        c.upload(st)
        bb = c.bounding_box()
        np.testing.assert_array_equal(bb, bb_ref)
        c.compute_morton_keys(bb)
        np.testing.assert_array_equal(c.keys(), keys_ref)
        c.sort_by_morton_key()
        np.testing.assert_array_equal(c.perm(), perm_ref)
        d = c.download()
        for f in ("x", "y", "z", "vz", "mass", "id"):
            np.testing.assert_array_equal(d[f], st[f][perm_ref])


def test_equal_keys_keep_index_order():
    # many particles on one point and on a line: equal 63-bit keys sort stably
    st, _ = configs.galaxy3d_small(seed=8, n=6000)
    st = dict(st)
    for f in ("x", "y", "z"):
        st[f] = st[f].copy()
        st[f][100:900] = st[f][50]
    bb, keys, perm = _oracle_sorted(st)
    with Context3D(6000) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        np.testing.assert_array_equal(c.keys(), keys)
        c.sort_by_morton_key()
        np.testing.assert_array_equal(c.perm(), perm)


@pytest.mark.parametrize("n", [1, 2, 8, 63, 64, 65, 257, 4999, 20000, 150000])
def test_octree_bitexact_sizes(n):
    st, _ = configs.galaxy3d_small(seed=11, n=max(n, 2))
    st = {k: v[:n] for k, v in st.items()}
    bb, keys, perm = _oracle_sorted(st)
    ref = O3.build_tree(st["x"][perm], st["y"][perm], st["z"][perm], st["mass"][perm], bb)
    with Context3D(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_octree(bb)
        got = c.nodes()
    _assert_nodes_equal(got, ref)


def test_octree_coincident_and_dense_clumps():
    # a clump of 5000 particles on one point and 3000 in a tiny cube: deep
    # chains down to the 1e-6 size floor in the level and workgroup phases
    st, _ = configs.galaxy3d_small(seed=12, n=30000)
    st = dict(st)
    rng = np.random.default_rng(1)
    for f in ("x", "y", "z"):
        st[f] = st[f].copy()
        st[f][:5000] = st[f][0]
        st[f][5000:8000] = st[f][1] + rng.uniform(0, 1e-3, 3000).astype(np.float32)
    bb, keys, perm = _oracle_sorted(st)
    ref = O3.build_tree(st["x"][perm], st["y"][perm], st["z"][perm], st["mass"][perm], bb)
    with Context3D(30000) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_octree(bb)
        got = c.nodes()
    _assert_nodes_equal(got, ref)


def _run_gravity(st, theta, **kw):
    n = len(st["x"])
    bb, keys, perm = _oracle_sorted(st)
    s = {f: st[f][perm] for f in ("x", "y", "z", "mass")}
    with Context3D(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_octree(bb)
        c.compute_gravity(_params(theta=theta, **kw))
        d = c.download(["ax", "ay", "az"])
        inter = c.interaction_counts()[0]
    return s, bb, d, inter


@pytest.mark.parametrize("theta", [0.5, 0.8, 1.0])
@pytest.mark.parametrize("compensated", [0, 1])
def test_bh_accel_matches_oracle(gal, theta, compensated):
    st, _ = gal
    s, bb, d, inter = _run_gravity(st, theta, compensated=compensated)
    nodes = O3.build_tree(s["x"], s["y"], s["z"], s["mass"], bb)
    rx, ry, rz, cnt = O3.bh_accel(s["x"], s["y"], s["z"], s["mass"], nodes, 1.0, theta, 1.0)
    assert inter == cnt
    _assert_acc_close((d["ax"], d["ay"], d["az"]), (rx, ry, rz))


@pytest.mark.parametrize("n", [3, 100, 1000, 4999])
def test_direct_sum_small_n(n):
    st, _ = configs.galaxy3d_small(seed=5, n=max(n, 3))
    st = {k: v[:n] for k, v in st.items()}
    bb, keys, perm = _oracle_sorted(st)
    s = {f: st[f][perm] for f in ("x", "y", "z", "mass")}
    with Context3D(n) as c:
        c.upload(st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_octree(bb)
        c.compute_gravity(_params())  # N < 5000 -> naiveGravity3D
        d = c.download(["ax", "ay", "az"])
    r = O3.direct_accel(s["x"], s["y"], s["z"], s["mass"], 1.0, 1.0)
    _assert_acc_close((d["ax"], d["ay"], d["az"]), r)


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


@pytest.mark.parametrize("n,steps,periodic", [(20000, 8, 0), (3000, 8, 0), (20000, 5, 1), (3000, 5, 1)])
def test_step_trajectory_matches_oracle(n, steps, periodic):
    st, prm = configs.galaxy3d_small(seed=7, n=n)
    st = dict(st)
    dom = dict(periodic=periodic, domain_w=900.0, domain_h=800.0, domain_d=300.0)
    if periodic:  # the galaxy overlaps the cube's faces: wrap it in first
        for f, w in (("x", 900.0), ("y", 800.0), ("z", 300.0)):
            st[f] = (np.mod(st[f] + 0.5 * w + 200.0, w) - 0.5 * w).astype(np.float32)
    p = _params(theta=prm["theta"], dt=prm["dt"], dt_start=prm["dt"], **dom)
    s = O3.State3(st)
    s.step(dict(prm, dt_start=prm["dt"], **dom), nsteps=steps)
    keys = ("x", "y", "z", "vx", "vy", "vz", "id")
    ref = _by_id({k: s[k].copy() for k in keys})
    with Context3D(n) as c:
        c.upload(st)
        c.step(p, steps)
        c.sync()
        got = _by_id(c.download(list(keys)))
    np.testing.assert_array_equal(got["id"], ref["id"])
    for f in ("x", "y", "z"):
        ext = float(np.ptp(ref[f]))
        diff = np.abs(got[f] - ref[f])
        if periodic:  # a particle on a seam may sit on either side of it
            w = {"x": 900.0, "y": 800.0, "z": 300.0}[f]
            diff = np.minimum(diff, np.abs(diff - w))
        assert diff.max() < 1e-4 * ext, f
    vs = float(np.abs(ref["vx"]).max())
    assert np.abs(got["vx"] - ref["vx"]).max() < 1e-3 * vs
    if periodic:
        assert np.all(got["x"] >= -450.0) and np.all(got["x"] < 450.0)
        assert np.all(got["z"] >= -150.0) and np.all(got["z"] < 150.0)


def test_bh_accel_million_two_galaxies():
    st, prm = configs.config3d(n=1_000_000)
    s, bb, d, inter = _run_gravity(st, 0.7, direct_below=0)
    nodes = O3.build_tree(s["x"], s["y"], s["z"], s["mass"], bb)
    rng = np.random.default_rng(0)
    t = np.sort(rng.choice(len(s["x"]), 2048, replace=False)).astype(np.uint32)
    rx, ry, rz, _ = O3.bh_accel(s["x"], s["y"], s["z"], s["mass"], nodes, 1.0, 0.7, 1.0)
    _assert_acc_close((d["ax"][t], d["ay"][t], d["az"][t]), (rx[t], ry[t], rz[t]))
    fx, fy, fz = O3.direct_fp64(s["x"], s["y"], s["z"], s["mass"], 1.0, 1.0, t)
    e = np.sqrt((d["ax"][t] - fx) ** 2 + (d["ay"][t] - fy) ** 2 + (d["az"][t] - fz) ** 2)
    rms = math.sqrt(np.mean(e ** 2) / np.mean(fx ** 2 + fy ** 2 + fz ** 2))
    print(f"3D BH theta=0.7 RMS force error vs FP64 direct: {rms:.3e}, interactions/particle "
          f"{inter / len(s['x']):.1f}")
    assert rms < 5e-2


def test_direct_fp64_matches_oracle(gal):
    st, _ = gal
    n = len(st["x"])
    t = np.arange(0, n, 97, dtype=np.uint32)
    with Context3D(n) as c:
        c.upload(st)
        got = c.direct_fp64(t, 1.0, 1.0)
    ref = O3.direct_fp64(st["x"], st["y"], st["z"], st["mass"], 1.0, 1.0, t)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(b).max())


def test_tree_bitexact_at_full_size_after_evolution():
    st, _ = configs.config3d(n=1_000_000)
    p = _params(theta=0.7, dt=0.005, dt_start=0.005, direct_below=0)
    with Context3D(1_000_000) as c:
        c.upload(st)
        c.step(p, 6)
        c.sync()
        d = c.download(["x", "y", "z", "mass"])
        got = c.nodes()
    bb = O3.bbox(d["x"], d["y"], d["z"])
    ref = O3.build_tree(d["x"], d["y"], d["z"], d["mass"], bb)
    _assert_nodes_equal(got, ref)


def test_leaf_capacity_and_block_sizes_agree():
    st, prm = configs.galaxy3d_small(seed=9, n=30000)
    out = []
    for leaf, blk in ((0, 256), (16, 128), (32, 64)):
        with Context3D(30000) as c:
            c.upload(st)
            c.step(_params(theta=0.7, leaf_capacity=leaf, walk_block=blk), 1)
            c.sync()
            out.append(_by_id(c.download(["ax", "ay", "az", "id"])))
    for o in out[1:]:
        a, r = (o["ax"], o["ay"], o["az"]), (out[0]["ax"], out[0]["ay"], out[0]["az"])
        mag = np.sqrt(sum(v ** 2 for v in r))
        err = np.sqrt(sum((x - y) ** 2 for x, y in zip(a, r)))
        assert np.sqrt(np.mean(err ** 2) / np.mean(mag ** 2)) < 2e-2


def test_edge_cases_and_rejections():
    with Context3D(10) as c:
        c.upload({"x": np.zeros(0), "y": np.zeros(0), "z": np.zeros(0), "mass": np.zeros(0)})
        c.step(_params(), 3)
        assert c.count() == 0
        c.upload({"x": np.array([1.0]), "y": np.array([2.0]), "z": np.array([3.0]),
                  "vz": np.array([2.0]), "mass": np.array([1.0])})
        c.step(_params(dt=0.5, dt_start=0.5), 2)
        d = c.download(["z", "az"])
        assert d["z"][0] == pytest.approx(5.0) and d["az"][0] == 0.0
        with pytest.raises(Exception):
            c.upload({"x": np.zeros(11), "y": np.zeros(11), "z": np.zeros(11), "mass": np.ones(11)})
        with pytest.raises(Exception):
            c.step(_params(sph=1), 1)
        with pytest.raises(Exception):
            c.step(_params(temperature=1), 1)
        with pytest.raises(Exception):
            c.step(_params(periodic=1, domain_d=0.0), 1)


def _step_nodes(st, leaf_capacity=0):
    # one frame with zero velocities: the frame's own sort and octree (whose
    # top levels split Morton-prefix runs) on positions that do not move
    n = len(st["x"])
    z = np.zeros(n, np.float32)
    st = dict(st, vx=z, vy=z, vz=z)
    with Context3D(n) as c:
        c.upload(st)
        c.step(_params(theta=0.7, dt=0.001, dt_start=0.001, direct_below=0, leaf_capacity=leaf_capacity), 1)
        c.sync()
        d = c.download(["x", "y", "z", "mass"])
        got = c.nodes()
    np.testing.assert_array_equal(np.sort(d["x"]), np.sort(st["x"]))
    bb = O3.bbox(d["x"], d["y"], d["z"])
    ref = O3.build_tree(d["x"], d["y"], d["z"], d["mass"], bb, leaf_capacity)
    return got, ref


def _midline_set3(seed=6, ncl=60000, nmid=30000, levels=5):
    """Particles within three float spacings of the midlines of the node cubes
    of the first `levels` levels (subGridMaker3D's float arithmetic) on one
    axis, plus a dense cluster and two corners that fix the bounding cube."""
    f32 = np.float32
    rng = np.random.default_rng(seed)
    corners = np.array([[-1000.3, -800.7, -600.1], [1500.9, 1200.1, 1300.3]], np.float32)
    cl = (rng.normal(0, 40, (ncl, 3)) + np.array([300.0, 250.0, 100.0])).astype(np.float32)
    pts = np.vstack([corners, cl])
    bb = O3.bbox(pts[:, 0], pts[:, 1], pts[:, 2])
    lo, size = [f32(v) for v in bb[:3]], f32(bb[3])
    mids = ([], [], [])
    pos, s = [tuple(lo)], size
    for _ in range(levels):
        half = f32(s * f32(0.5))
        for a in range(3):
            mids[a].extend(f32(p[a] + half) for p in pos)
        pos = [tuple(f32(p[a] + (half if (q >> a) & 1 else f32(0))) for a in range(3)) for p in pos for q in range(8)]
        s = half
    lines = [np.unique(np.array(m, np.float32)) for m in mids]
    out = np.zeros((nmid, 3), np.float32)
    for k in range(nmid):
        ax = int(rng.integers(3))
        v = lines[ax][rng.integers(len(lines[ax]))]
        off = int(rng.integers(-3, 4))
        for _ in range(abs(off)):
            v = np.nextafter(v, f32(np.inf) if off > 0 else f32(-np.inf), dtype=np.float32)
        row = rng.uniform([-900, -700, -500], [1400, 1100, 1200]).astype(np.float32)
        row[ax] = v
        out[k] = row
    pts = np.vstack([pts, out])
    n = len(pts)
    return {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "z": pts[:, 2].copy(), "mass": np.ones(n, np.float32),
            "id": np.arange(n, dtype=np.uint32)}


@pytest.mark.parametrize("leaf_capacity", [0, 32])
def test_frame_octree_bitexact_on_midlines(leaf_capacity):
    got, ref = _step_nodes(_midline_set3(), leaf_capacity)
    _assert_nodes_equal(got, ref)


def test_frame_octree_bitexact_dense_clumps():
    st, _ = configs.galaxy3d_small(seed=12, n=30000)
    st = dict(st)
    rng = np.random.default_rng(1)
    for f in ("x", "y", "z"):
        st[f] = st[f].copy()
        st[f][:5000] = st[f][0]
        st[f][5000:8000] = st[f][1] + rng.uniform(0, 1e-3, 3000).astype(np.float32)
    got, ref = _step_nodes(st)
    _assert_nodes_equal(got, ref)
