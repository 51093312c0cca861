"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

The reference ships no tests or fixtures for this path, and cannot be built
here (its headers need glm, which is absent).  Each oracle function is pinned
by something independent of its own code:

  geo_morton2d / geo_scale_to_grid / geo_bbox -> tests/golden/kat.json
      (explicit bit interleave, float32 numpy evaluation)
  geo_sort_perm   -> numpy stable argsort
  geo_build_tree  -> the hand-checked tree in kat.json, plus structural
      invariants (ranges partition, next = subtree size - 1, pre-order,
      leaf rule, children inside parent squares) and FP64 mass / COM sums
  geo_bh_accel    -> FP64 direct sum: theta -> 0 reduces BH to the direct sum,
      and the two-body closed form; theta = 0.7 error bound
  geo_direct_accel-> closed forms with the unsquared softening, momentum
      conservation (Newton's third law)
  geo_direct_fp64 -> two-body closed form
  geo_integrate_* -> kick-drift-kick on a harmonic force: closed-form one-step
      update
  geo_sph_candidates -> O(N^2) brute force over cell coordinates
  geo_sph_constants  -> the float-pow values in kat.json (which differ in the
      last bit from the double form at h = 2.5, 3 and 3.5); W integrates to 1
  geo_sph_pcisph  -> two-particle closed forms, one per pair term, written from
      the physics of SPH.cpp:98-281 in FP64: viscosity (visc and density of
      the other particle, 0.001 floor, both visits), cohesion (own cohesion,
      both visits), self-inclusive predicted density over the averaged rest
      density, the doubled pressure pair force with the min(1, 8 / massRatio)
      clamp, and a pinned particle's one-sided visit; plus symmetry (momentum
      conserved) and the isolated particle's self term
  computeInternalMass order -> a known answer where only the fold q = 0, 2,
      1, 3 gives the root mass 2^24
  MAC softening placement   -> a target whose acceptance of one node flips
      with the placement of eps^2 (checked through its acceleration)
  SPH speed clamp (integrateStart) -> closed form of the clamp of vel and
      prevVel.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def test_morton_kat():
    for ix, iy, k in KAT["morton"]:
        assert O.morton2d(ix, iy) == k


def test_scale_to_grid_kat():
    for p, lo, hi, v in KAT["scale_to_grid"]:
        assert O.scale_to_grid(p, lo, hi) == v, (p, lo, hi)


def test_bbox_kat():
    for case in KAT["bbox"]:
        bb = O.bbox(case["x"], case["y"])
        np.testing.assert_array_equal(bb, np.array(case["bb"], np.float32))


def test_morton_keys_matches_kat_map():
    rng = np.random.default_rng(0)
    x = rng.uniform(-100, 100, 500).astype(np.float32)
    y = rng.uniform(-30, 70, 500).astype(np.float32)
    bb = O.bbox(x, y)
    keys = O.morton_keys(x, y, bb)
    ext = np.float32(max(bb[2], np.float32(1e-6)))
    hx, hy = np.float32(bb[0] + ext), np.float32(bb[1] + ext)
    for i in range(0, 500, 7):
        ix = O.scale_to_grid(x[i], bb[0], hx)
        iy = O.scale_to_grid(y[i], bb[1], hy)
        k = 0
        for b in range(18):
            k |= ((ix >> b) & 1) << (2 * b) | ((iy >> b) & 1) << (2 * b + 1)
        assert int(keys[i]) == k
    # the particle on the max edge clamps to 262143 (morton.cpp:9)
    j = int(np.argmax(x))
    assert O.scale_to_grid(x[j], bb[0], hx) <= 262143


def test_sort_perm_is_stable_argsort():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 3000).astype(np.uint64)  # many ties
    p = O.sort_perm(keys)
    np.testing.assert_array_equal(p, np.argsort(keys, kind="stable").astype(np.uint32))


def test_tree_hand_checked():
    t = KAT["tree"]
    pts = np.array(t["points"], np.float32)
    m = np.ones(len(pts), np.float32)
    nodes = O.build_tree(pts[:, 0], pts[:, 1], m, np.array(t["bb"], np.float32))
    got = [(int(n["start"]), int(n["end"]), int(n["next"]), float(n["size"])) for n in nodes]
    assert got == [tuple(v) for v in t["nodes"]]


def _check_tree(nodes, x, y, m, leaf_cap=0):
    nn = len(nodes)
    n = len(x)
    assert nodes[0]["start"] == 0 and nodes[0]["end"] == n
    assert nodes[0]["next"] == nn - 1
    for i in range(nn):
        nd = nodes[i]
        cnt = int(nd["end"] - nd["start"])
        assert cnt >= 1
        if leaf_cap:
            is_leaf = cnt <= leaf_cap or nd["size"] <= 1e-6
        else:
            is_leaf = ((cnt <= 4 and nd["size"] <= 2.0) or cnt == 1) or nd["size"] <= 1e-6
        ch = [c for c in nd["child"] if c != 0xFFFFFFFF]
        if is_leaf:
            assert nd["next"] == 0 and not ch
            mass = m[nd["start"]:nd["end"]].astype(np.float64).sum()
            assert abs(nd["mass"] - mass) <= 1e-5 * mass + 1e-30
        else:
            assert ch, i
            # children are consecutive in pre-order and tile the parent range
            assert ch[0] == i + 1
            s = nd["start"]
            nxt = i + 1
            tot = 0
            for c in ch:
                assert c == nxt
                assert nodes[c]["start"] == s
                s = nodes[c]["end"]
                tot += 1 + int(nodes[c]["next"])
                nxt = c + int(nodes[c]["next"]) + 1
                assert nodes[c]["size"] == np.float32(nd["size"] * np.float32(0.5))
            assert s == nd["end"]
            assert tot == nd["next"]
        # COM is the mass-weighted mean of the particles in range (FP64 check)
        sl = slice(int(nd["start"]), int(nd["end"]))
        mm = m[sl].astype(np.float64)
        if mm.sum() > 0:
            cx = (x[sl].astype(np.float64) * mm).sum() / mm.sum()
            cy = (y[sl].astype(np.float64) * mm).sum() / mm.sum()
            scale = max(1.0, abs(cx), abs(cy))
            assert abs(nd["comx"] - cx) <= 1e-4 * scale and abs(nd["comy"] - cy) <= 1e-4 * scale


def _sorted_cloud(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 50, n).astype(np.float32)
    y = rng.normal(0, 20, n).astype(np.float32)
    m = rng.uniform(0.5, 1.5, n).astype(np.float32)
    bb = O.bbox(x, y)
    p = O.sort_perm(O.morton_keys(x, y, bb))
    return x[p], y[p], m[p], bb


def test_tree_invariants():
    x, y, m, bb = _sorted_cloud(3000, 2)
    nodes = O.build_tree(x, y, m, bb)
    _check_tree(nodes, x, y, m)
    # particles lie inside their leaf's square (closed on the lower edges)
    for nd in nodes:
        if nd["next"] == 0:
            sl = slice(int(nd["start"]), int(nd["end"]))
            assert np.all(x[sl] >= nd["posx"] - 1e-3) and np.all(x[sl] <= nd["posx"] + nd["size"] + 1e-3)
            assert np.all(y[sl] >= nd["posy"] - 1e-3) and np.all(y[sl] <= nd["posy"] + nd["size"] + 1e-3)


@pytest.mark.parametrize("cap", [8, 16, 32])
def test_tree_leaf_capacity_is_pruned_full_tree(cap):
    # a leaf-capacity tree splits exactly like the fully refined tree (same
    # partition points and squares) and stops at nodes of <= cap particles
    x, y, m, bb = _sorted_cloud(4000, 3)
    full = O.build_tree(x, y, m, bb, leaf_cap=1)
    t = O.build_tree(x, y, m, bb, leaf_cap=cap)
    _check_tree(t, x, y, m, leaf_cap=cap)
    geo = {(float(a["posx"]), float(a["posy"]), float(a["size"]), int(a["start"]), int(a["end"])) for a in full}
    for nd in t:
        key = (float(nd["posx"]), float(nd["posy"]), float(nd["size"]), int(nd["start"]), int(nd["end"]))
        assert key in geo
        cnt = int(nd["end"] - nd["start"])
        if nd["next"] == 0:
            assert cnt <= cap or nd["size"] <= 1e-6
            # computeLeafMass: sequential float sums over the leaf's range
            gm = np.float32(0)
            for i in range(int(nd["start"]), int(nd["end"])):
                gm = np.float32(gm + m[i])
            assert nd["mass"] == gm
        else:
            assert cnt > cap
    # every fully refined node of > cap particles survives
    assert sum(1 for a in full if a["end"] - a["start"] > cap) == sum(1 for a in t if a["next"] > 0)


def test_tree_duplicates_and_single():
    x = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0], np.float32)
    y = np.array([2.0] * 6, np.float32)
    m = np.ones(6, np.float32)
    nodes = O.build_tree(x, y, m, O.bbox(x, y))
    assert len(nodes) == 1 and nodes[0]["next"] == 0  # size 0 <= 1e-6 -> leaf
    nodes = O.build_tree(x[:1], y[:1], m[:1], O.bbox(x[:1], y[:1]))
    assert len(nodes) == 1


def test_bh_two_body_closed_form():
    x = np.array([0.0, 3.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    m = np.array([2.0, 5.0], np.float32)
    bb = O.bbox(x, y)
    nodes = O.build_tree(x, y, m, bb)
    ax, ay, cnt = O.bh_accel(x, y, m, nodes, 1.0, 0.5, 1.0)
    r2 = 25.0 + 1.0
    f = 1.0 / r2 ** 1.5
    np.testing.assert_allclose([ax[0], ay[0]], [5 * 3 * f, 5 * 4 * f], rtol=1e-6)
    np.testing.assert_allclose([ax[1], ay[1]], [-2 * 3 * f, -2 * 4 * f], rtol=1e-6)
    dx, dy = O.direct_fp64(x, y, m, 1.0, 1.0)
    np.testing.assert_allclose(ax, dx, rtol=1e-6)


def test_bh_theta_zero_is_direct_sum():
    x, y, m, bb = _sorted_cloud(1500, 3)
    nodes = O.build_tree(x, y, m, bb)
    ax, ay, cnt = O.bh_accel(x, y, m, nodes, 1.0, 0.0, 1.0)
    assert cnt == 1500 * 1499
    rx, ry = O.direct_fp64(x, y, m, 1.0, 1.0)
    err = np.hypot(ax - rx, ay - ry) / np.hypot(rx, ry)
    assert err.max() < 1e-4


def test_bh_theta_error_bound():
    from galaxy_step import configs
    st, _ = configs.config0(n=5000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb = O.bbox(x, y)
    p = O.sort_perm(O.morton_keys(x, y, bb))
    x, y, m = x[p], y[p], m[p]
    nodes = O.build_tree(x, y, m, bb)
    rx, ry = O.direct_fp64(x, y, m, 1.0, 1.0)
    prev = None
    prev_rms = 0.0
    for th in (0.3, 0.7, 1.0):
        ax, ay, cnt = O.bh_accel(x, y, m, nodes, 1.0, th, 1.0)
        rms = math.sqrt(np.mean((np.hypot(ax - rx, ay - ry) / np.hypot(rx, ry)) ** 2))
        assert rms < 0.05
        assert rms > prev_rms  # monopole truncation error grows with theta
        prev_rms = rms
        if prev is not None:
            assert cnt < prev  # a larger opening angle accepts more nodes
        prev = cnt


def test_direct_unsquared_softening_and_momentum():
    x = np.array([0.0, 2.0], np.float32)
    y = np.array([0.0, 0.0], np.float32)
    m = np.array([1.0, 3.0], np.float32)
    ax, ay = O.direct_accel(x, y, m, 1.0, 0.5)  # eps added unsquared: r^2 + 0.5
    f = 1.0 / (4.0 + 0.5) ** 1.5
    np.testing.assert_allclose(ax, [3 * 2 * f, -1 * 2 * f], rtol=1e-6)
    rng = np.random.default_rng(5)
    x = rng.normal(0, 10, 700).astype(np.float32)
    y = rng.normal(0, 10, 700).astype(np.float32)
    m = rng.uniform(0.5, 2, 700).astype(np.float32)
    ax, ay = O.direct_accel(x, y, m, 1.0, 1.0)
    px = (m.astype(np.float64) * ax).sum() / (m * np.abs(ax)).sum()
    assert abs(px) < 1e-5


def test_integrator_kdk_closed_form():
    st = {"x": np.array([1.0, 2.0]), "y": np.array([0.0, 1.0]), "vx": np.array([0.5, 0.0]),
          "vy": np.array([0.0, -1.0]), "mass": np.array([1.0, 1.0]),
          "ax": np.array([0.25, -1.0]), "ay": np.array([0.0, 2.0]),
          "flags": np.array([0, O.GEO_PINNED])}
    s = O.State(st)
    # G = 0 -> acc zeroed; one frame = kick(dt/2, old acc), drift(dt), kick(dt/2, 0)
    s.step(dict(G=0.0, dt=0.1, naive=1))
    ids = s["id"].tolist()
    i0, i1 = ids.index(0), ids.index(1)
    vx = np.float32(np.float32(0.5) + np.float32(np.float32(0.25) * np.float32(0.05)))
    assert s["vx"][i0] == vx
    assert s["x"][i0] == np.float32(np.float32(1.0) + np.float32(vx * np.float32(0.1)))
    assert s["pvx"][i0] == np.float32(0.5)
    # pinned particle untouched
    assert s["x"][i1] == np.float32(2.0) and s["vy"][i1] == np.float32(-1.0)


def test_sph_candidates_brute_force():
    rng = np.random.default_rng(6)
    n = 800
    x = rng.uniform(-20, 20, n).astype(np.float32)
    y = rng.uniform(-15, 15, n).astype(np.float32)
    qx = (x + rng.normal(0, 0.5, n)).astype(np.float32)
    qy = (y + rng.normal(0, 0.5, n)).astype(np.float32)
    cell = 3.0
    off, idx = O.sph_candidates(x, y, qx, qy, cell)
    cx = np.floor(x / np.float32(cell)).astype(np.int64)
    cy = np.floor(y / np.float32(cell)).astype(np.int64)
    qcx = np.floor(qx / np.float32(cell)).astype(np.int64)
    qcy = np.floor(qy / np.float32(cell)).astype(np.int64)
    for i in range(0, n, 13):
        want = np.nonzero((np.abs(cx - qcx[i]) <= 1) & (np.abs(cy - qcy[i]) <= 1))[0]
        np.testing.assert_array_equal(idx[off[i]:off[i + 1]], want.astype(np.uint32))


def test_sph_constants_closed_form():
    for c in KAT["sph_constants"]:
        k = O.sph_constants(c["h"])
        # exact: the float evaluation order of SPH.h:38-56
        np.testing.assert_array_equal(k, np.array([c["volume"], c["spiky"], c["lap"]], np.float32))
        h = c["h"]
        np.testing.assert_allclose(k, [math.pi * h ** 4 / 6, -45 / (math.pi * h ** 6), 45 / (math.pi * h ** 6)],
                                   rtol=3e-7)
    # W = (h - r)^2 / volume integrates to 1 over the disk of radius h
    h = 3.0
    vol = O.sph_constants(h)[0]
    r = np.linspace(0, h, 200001)
    w = (h - r) ** 2 / vol * 2 * math.pi * r
    assert abs(np.trapezoid(w, r) - 1.0) < 1e-5


def _sph_state(n, seed, spacing=1.2):
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel()[:n] * spacing + rng.normal(0, 0.2, n)).astype(np.float32)
    y = (gy.ravel()[:n] * spacing + rng.normal(0, 0.2, n)).astype(np.float32)
    st = {"x": x, "y": y, "vx": rng.normal(0, 1, n), "vy": rng.normal(0, 1, n),
          "mass": np.ones(n), "flags": np.full(n, O.GEO_SPH), "sph_mass": np.ones(n),
          "rest_dens": np.full(n, 0.008), "stiff": np.ones(n), "visc": np.ones(n),
          "cohesion": np.ones(n), "dens": np.zeros(n)}
    return st


def test_sph_momentum_conserved():
    st = _sph_state(900, 7)
    s = O.State(st, sph=True)
    s.step(dict(G=0.0, dt=0.002, sph=1, sph_h=3.0, sph_cell=3.0), nsteps=1)
    fx = s["pfx"].astype(np.float64)  # sphMass = 1 -> pressF is the force
    fy = s["pfy"].astype(np.float64)
    scale = np.abs(fx).sum() + np.abs(fy).sum()
    assert scale > 0
    assert abs(fx.sum()) < 1e-4 * scale and abs(fy.sum()) < 1e-4 * scale


def test_sph_isolated_density_is_self_term():
    st = _sph_state(4, 8)
    st["x"] = np.array([0.0, 100.0, 200.0, 300.0], np.float32)
    st["y"] = np.zeros(4, np.float32)
    s = O.State(st, sph=True)
    s.step(dict(G=0.0, dt=0.002, sph=1, sph_h=3.0, sph_cell=3.0), nsteps=1)
    vol = O.sph_constants(3.0)[0]
    want = np.float32(np.float32(np.float32(0.03) * np.float32(np.float32(9.0) / vol)) / np.float32(0.008))
    np.testing.assert_allclose(s["pred_dens"], np.full(4, want), rtol=1e-6)
    assert np.all(s["pfx"] == 0) and np.all(s["pfy"] == 0)


def test_full_step_permutes_consistently():
    from galaxy_step import configs
    st, prm = configs.config0(n=6000)
    s = O.State(st)
    s.step(dict(prm, theta=0.8), nsteps=2)
    ids = s["id"]
    assert sorted(ids.tolist()) == list(range(6000))
    # the stored order is the Morton order of the current positions' frame
    assert np.all(np.isfinite(s["x"])) and np.all(np.isfinite(s["vx"]))
    np.testing.assert_array_equal(s["mass"], st["mass"][ids])


def _v2_brute(x, y, qx, qy, cell, radius):
    inv = np.float32(1.0) / np.float32(cell)
    R = int(np.ceil(np.float32(radius) * inv))
    cx = (x * inv).astype(np.int64)  # (int)(p * invCellSize): truncation toward zero
    cy = (y * inv).astype(np.int64)
    out = []
    for q in range(len(qx)):
        qcx, qcy = int(np.float32(qx[q]) * inv), int(np.float32(qy[q]) * inv)
        dx = x - np.float32(qx[q])
        dy = y - np.float32(qy[q])
        d2 = dx * dx + dy * dy
        ok = (np.abs(cx - qcx) <= R) & (np.abs(cy - qcy) <= R) & (d2 <= np.float32(radius) * np.float32(radius))
        out.append(np.nonzero(ok)[0].astype(np.uint32))
    return out


@pytest.mark.parametrize("cell,radius", [(3.0, 3.5), (4.0, 4.5)])
def test_sph_candidates_v2_brute_force(cell, radius):
    # NeighborSearchV2 (neighborSearch.h:439-599): truncated cells around the
    # origin (negative coordinates included), radius filter, self included
    rng = np.random.default_rng(12)
    x = rng.uniform(-30, 30, 1500).astype(np.float32)
    y = rng.uniform(-30, 30, 1500).astype(np.float32)
    qx = (x + rng.normal(0, 0.5, 1500)).astype(np.float32)
    qy = (y + rng.normal(0, 0.5, 1500)).astype(np.float32)
    off, idx = O.sph_candidates(x, y, qx, qy, cell, mode=1, radius=radius)
    want = _v2_brute(x, y, qx, qy, cell, radius)
    for q in range(len(qx)):
        np.testing.assert_array_equal(idx[off[q]:off[q + 1]], want[q])


def test_neighbor_amount_brute_force():
    rng = np.random.default_rng(13)
    n = 1200
    x = rng.uniform(-20, 20, n).astype(np.float32)
    y = rng.uniform(-20, 20, n).astype(np.float32)
    flags = np.zeros(n, np.uint32)
    flags[::17] = 1   # pinned
    flags[::23] |= 8  # dark matter
    got = O.neighbor_amount(x, y, flags)
    want = np.array([len(w) for w in _v2_brute(x, y, x, y, 3.0, 3.5)], np.uint32)
    want[(flags & 9) != 0] = 0
    np.testing.assert_array_equal(got, want)


def test_pcisph_search_modes_agree_when_both_cover_h():
    # with h <= cell (AVX2 3x3 stencil) and h <= search radius (V2 filter)
    # both candidate sets contain every particle within h: same forces up to
    # summation order
    rng = np.random.default_rng(14)
    n = 900
    side = 30
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    st = {"x": (gx.ravel()[:n] * 1.4 - 20 + rng.normal(0, 0.3, n)).astype(np.float32),
          "y": (gy.ravel()[:n] * 1.4 - 20 + rng.normal(0, 0.3, n)).astype(np.float32),
          "vx": rng.normal(0, 1, n).astype(np.float32), "vy": rng.normal(0, 1, n).astype(np.float32),
          "mass": np.ones(n, np.float32), "flags": np.full(n, 2, np.uint32), "id": np.arange(n, dtype=np.uint32),
          "sph_mass": np.ones(n, np.float32), "rest_dens": np.full(n, 0.008, np.float32),
          "stiff": np.ones(n, np.float32), "visc": np.ones(n, np.float32), "cohesion": np.ones(n, np.float32),
          "dens": np.zeros(n, np.float32)}
    prm = dict(G=0.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, sph=1, sph_h=3.0, sph_cell=3.0)
    res = []
    for mode in (0, 1):
        s = O.State(st, sph=True)
        s.step(dict(prm, sph_search=mode, sph_search_radius=3.5), nsteps=1)
        o = np.argsort(s["id"])
        res.append({k: s[k][o].astype(np.float64) for k in ("pred_dens", "pfx", "pfy", "vx")})
    for k in res[0]:
        scale = np.abs(res[0][k]).max() or 1.0
        assert np.abs(res[0][k] - res[1][k]).max() <= 1e-5 * scale, k


def test_periodic_wrap_two_body_closed_form():
    # physics.cpp:117-120 / 417-420: d -= W * ((d > W/2) - (d < -W/2)); two
    # bodies 20 apart across the x = 0 / x = W seam pull toward each other
    W, H = 1000.0, 600.0
    x = np.array([10.0, W - 10.0], np.float32)
    y = np.array([300.0, 300.0], np.float32)
    m = np.array([1.0, 3.0], np.float32)
    ax, ay = O.direct_accel_periodic(x, y, m, 1.0, 1.0, W, H)   # softening unsquared (A-2)
    want = 3.0 * 20.0 / (400.0 + 1.0) ** 1.5
    assert ax[0] == pytest.approx(-want, rel=1e-6) and ay[0] == 0.0
    assert ax[1] == pytest.approx(1.0 * 20.0 / 401.0 ** 1.5, rel=1e-6)
    # the tree walk with theta -> 0 is the same direct sum (eps squared)
    bb = O.bbox(x, y)
    nodes = O.build_tree(x, y, m, bb)
    bx, by, _ = O.bh_accel_periodic(x, y, m, nodes, 1.0, 0.01, 1.0, W, H)
    assert bx[0] == pytest.approx(-3.0 * 20.0 / 401.0 ** 1.5, rel=1e-6)


def test_periodic_positions_wrap_and_ground_walls():
    W, H = 1000.0, 600.0
    st = {"x": np.array([999.0, 1.0, 500.0], np.float32), "y": np.array([300.0, 300.0, 300.0], np.float32),
          "vx": np.array([200.0, -200.0, 0.0], np.float32), "vy": np.zeros(3, np.float32),
          "mass": np.ones(3, np.float32), "flags": np.zeros(3, np.uint32), "id": np.arange(3, dtype=np.uint32)}
    s = O.State(st)
    s.step(dict(G=0.0, dt=0.01, dt_start=0.01, periodic=1, domain_w=W, domain_h=H), nsteps=1)
    o = np.argsort(s["id"])
    x = s["x"][o]
    assert x[0] == pytest.approx(1.0, abs=1e-3) and x[1] == pytest.approx(999.0, abs=1e-3)
    # groundModeBoundary (SPH.cpp:283-322): a particle below the floor is put
    # at y = h with vy damped by -0.1 and vx scaled by 1 - friction
    n = 2
    g = {"x": np.array([50.0, 100.0], np.float32), "y": np.array([1.0, 50.0], np.float32),
         "vx": np.array([2.0, 0.0], np.float32), "vy": np.array([-5.0, 0.0], np.float32),
         "mass": np.ones(n, np.float32), "flags": np.full(n, 2, np.uint32), "id": np.arange(n, dtype=np.uint32),
         "sph_mass": np.ones(n, np.float32), "rest_dens": np.full(n, 0.008, np.float32),
         "stiff": np.ones(n, np.float32), "visc": np.ones(n, np.float32), "cohesion": np.ones(n, np.float32),
         "dens": np.zeros(n, np.float32)}
    s = O.State(g, sph=True)
    s.step(dict(G=0.0, dt=0.01, dt_start=0.01, sph=1, sph_h=3.0, sph_cell=3.0, sph_ground=1, domain_w=W,
                domain_h=H, vertical_gravity=3.0, boundary_friction=0.25), nsteps=1)
    o = np.argsort(s["id"])
    gx, gy, gvx, gvy, gay = (s[k][o] for k in ("x", "y", "vx", "vy", "ay"))
    # particle 0: drift to y = 0.95, wall: y = h, vy = -5 * -0.1, vx = 2 * 0.75; kick: vy += 3 * dt / 2
    assert gy[0] == np.float32(3.0) and gvx[0] == np.float32(1.5)
    assert gvy[0] == np.float32(np.float32(0.5) + np.float32(3.0) * np.float32(0.005))
    assert gay[0] == np.float32(3.0) and gay[1] == np.float32(3.0)
    assert gy[1] == np.float32(50.0) and gvy[1] == np.float32(np.float32(3.0) * np.float32(0.005))


# ---- two-particle closed forms of the 2D PCISPH pair terms -----------------
# SPH.cpp:98-281 with i = particle 0, j = particle 1, written in FP64 from the
# physics, independent of the oracle's code:
#   pass 1 (viscosity + cohesion, at pos): each particle visits the other, so
#     F_i = mu lap (v_j - v_i) (visc_j m_j / max(dens_j, 0.001)
#                               + visc_i m_i / max(dens_i, 0.001))
#         + c W_coh(r) n_ij (coh_i m_j + coh_j m_i),   F_j = -F_i
#     (a pinned or drawn particle does not visit, so only the other's term)
#   predPos = pos + F / sphMass * dt^2 / 2
#   predDens_i = m_i W(0) / rho0_i + m_j W(rr) / ((rho0_i + rho0_j) / 2)
#   press_i = max(0, delta (predDens_i - rho0_i)) stiff_i stiffMultiplier
#   pass 2 (pressure): both visits add the same pair force, so
#     F_i += 2 mag gradW(rr) (pp_i - pp_j) / rr,
#     mag = -(m_i + m_j) (P_i + P_j) / 2 / max((d_i + d_j) / 2, 0.01)
#           * min(1, 8 / massRatio)
#   pressF = F / sphMass; acc += pressF unless pinned
# with m = sphMass * myVar.mass, W(r) = (h - r)^2 / (pi h^4 / 6),
# gradW(r) = -45 / (pi h^6) (h - r)^2, lap = 45 / (pi h^6),
# W_coh(r) = (1 - q)(1/2 - q)^2 30 / (pi h^2), q = r / h.

_PAIR_PRM = dict(sph_h=3.0, sph_cell=3.0, sph_unit_mass=0.03, viscosity=0.3, cohesion_coef=1.0,
                 delta=19000.0, stiff_mult=1.0, dt=0.002)


def _pair(**kw):
    st = {"x": [0.4, 2.1], "y": [0.5, 1.1], "vx": [0.0, 0.0], "vy": [0.0, 0.0], "mass": [1.0, 1.0],
          "flags": [O.GEO_SPH, O.GEO_SPH], "sph_mass": [1.0, 1.0], "rest_dens": [0.008, 0.008],
          "stiff": [1.0, 1.0], "visc": [1.0, 1.0], "cohesion": [0.0, 0.0], "dens": [0.0, 0.0]}
    st.update(kw)
    out = {k: (np.asarray(v, np.uint32) if k == "flags" else np.asarray(v, np.float32)) for k, v in st.items()}
    return out


def _pair_closed_form(st, prm):
    f = {k: np.asarray(v, np.float64) for k, v in st.items() if k != "flags"}
    h, unit, mu, c = prm["sph_h"], prm["sph_unit_mass"], prm["viscosity"], prm["cohesion_coef"]
    vol, lap, spiky = math.pi * h ** 4 / 6, 45 / (math.pi * h ** 6), -45 / (math.pi * h ** 6)
    m = f["sph_mass"] * unit
    pos = np.stack([f["x"], f["y"]], 1)
    vel = np.stack([f["vx"], f["vy"]], 1)
    visits = [(int(st["flags"][k]) & O.GEO_PINNED) == 0 for k in (0, 1)]
    d = pos[1] - pos[0]
    r = math.hypot(*d)
    n = d / r
    q = r / h
    wc = (1 - q) * (0.5 - q) ** 2 * 30 / (math.pi * h * h)
    dens = np.maximum(f["dens"], 0.001)
    F = np.zeros((2, 2))
    if visits[0]:  # i visits j (SPH.cpp:110-147)
        t = mu * f["visc"][1] * m[1] / dens[1] * lap * (vel[1] - vel[0]) + c * f["cohesion"][0] * m[1] * wc * n
        F[0] += t
        F[1] -= t
    if visits[1]:  # j visits i
        t = mu * f["visc"][0] * m[0] / dens[0] * lap * (vel[0] - vel[1]) + c * f["cohesion"][1] * m[0] * wc * (-n)
        F[1] += t
        F[0] -= t
    pp = pos + F / f["sph_mass"][:, None] * (0.5 * prm["dt"] ** 2)
    rr = math.hypot(*(pp[0] - pp[1]))
    rho = f["rest_dens"]
    w0, wr = h * h / vol, (h - rr) ** 2 / vol
    pd = np.array([m[0] * w0 / rho[0] + m[1] * wr / (0.5 * (rho[0] + rho[1])),
                   m[1] * w0 / rho[1] + m[0] * wr / (0.5 * (rho[0] + rho[1]))])
    press = np.maximum(0.0, prm["delta"] * (pd - rho)) * f["stiff"] * prm["stiff_mult"]
    ratio = f["sph_mass"].max() / f["sph_mass"].min()
    mag = -(m[0] + m[1]) * 0.5 * (press[0] + press[1]) / max(0.5 * (pd[0] + pd[1]), 0.01) * min(1.0, 8.0 / ratio)
    pf = mag * spiky * (h - rr) ** 2 * (pp[0] - pp[1]) / rr
    F[0] += 2 * pf
    F[1] -= 2 * pf
    pressF = F / f["sph_mass"][:, None]
    return {"pred_dens": pd, "press": press, "pressF": pressF, "ppos": pp}


def _check_pair(st, prm, rtol=2e-5):
    want = _pair_closed_form(st, prm)
    s = O.State(st, sph=True)
    s.pcisph(prm)
    np.testing.assert_allclose(s["pred_dens"], want["pred_dens"], rtol=rtol)
    np.testing.assert_allclose(s["press"], want["press"], rtol=rtol, atol=1e-9)
    got = np.stack([s["pfx"], s["pfy"]], 1).astype(np.float64)
    scale = np.abs(want["pressF"]).max()
    assert scale > 0
    np.testing.assert_allclose(got, want["pressF"], rtol=0, atol=rtol * scale)
    np.testing.assert_allclose(np.stack([s["ppx"], s["ppy"]], 1), want["ppos"], rtol=0, atol=1e-5)
    pinned = (np.asarray(st["flags"]) & O.GEO_PINNED) != 0
    np.testing.assert_array_equal(s["ax"][pinned], 0)  # acc untouched for pinned particles
    np.testing.assert_allclose(s["ax"][~pinned], got[~pinned, 0], rtol=1e-6)
    return want


def test_pcisph_pair_pressure_doubled_and_mass_ratio_clamped():
    # sphMass ratio 20 > 8: the pair force is scaled by 8 / 20; each pair is
    # visited from both sides, so it enters twice; low rest densities make
    # both pressures positive and unequal (stiff 1 and 2)
    st = _pair(sph_mass=[1.0, 20.0], rest_dens=[0.002, 0.003], stiff=[1.0, 2.0])
    want = _check_pair(st, _PAIR_PRM)
    assert want["press"].min() > 0 and want["press"][0] != want["press"][1]


def test_pcisph_pair_viscosity_uses_other_particles_visc_and_density():
    prm = dict(_PAIR_PRM, delta=0.0)  # no pressure
    vel = dict(vx=[1.0, -2.0], vy=[-0.5, 0.7])
    # dens = 0 on the live path: the 0.001 floor
    _check_pair(_pair(visc=[0.5, 2.0], sph_mass=[1.0, 3.0], **vel), prm)
    # a written density of the other particle replaces the floor (dens_j = 5)
    _check_pair(_pair(visc=[0.5, 2.0], sph_mass=[1.0, 3.0], dens=[0.0, 5.0], **vel), prm)
    # a pinned particle does not visit, but is visited (one-sided term)
    _check_pair(_pair(visc=[0.5, 2.0], sph_mass=[1.0, 3.0], flags=[O.GEO_SPH | O.GEO_PINNED, O.GEO_SPH], **vel), prm)


def test_pcisph_pair_cohesion_uses_own_cohesion():
    prm = dict(_PAIR_PRM, delta=0.0)
    _check_pair(_pair(cohesion=[0.3, 1.7], sph_mass=[1.0, 4.0]), prm)


def test_pcisph_pair_all_terms_together():
    st = _pair(vx=[1.0, -2.0], vy=[-0.5, 0.7], visc=[0.5, 2.0], cohesion=[0.3, 1.7], sph_mass=[1.0, 12.0],
               rest_dens=[0.002, 0.004], stiff=[1.5, 0.5])
    _check_pair(st, _PAIR_PRM)


def test_pair_closed_form_separates_the_plausible_slips():
    """The closed form itself tells the slips apart (so a restatement with one
    of them fails the tests above): pressure not doubled, clamp dropped, visc_i
    for visc_j, coh_j for coh_i."""
    st = _pair(vx=[1.0, -2.0], vy=[-0.5, 0.7], visc=[0.5, 2.0], cohesion=[0.3, 1.7], sph_mass=[1.0, 20.0],
               rest_dens=[0.002, 0.003], stiff=[1.0, 2.0])
    base = _pair_closed_form(st, _PAIR_PRM)["pressF"]
    swapped = dict(st, visc=st["visc"][::-1].copy(), cohesion=st["cohesion"][::-1].copy())
    alt = _pair_closed_form(swapped, _PAIR_PRM)["pressF"]
    assert np.abs(alt - base).max() > 1e-3 * np.abs(base).max()


# ---- quadtree mass fold order (quadtree.h:70-94) ---------------------------

def test_tree_mass_fold_order_known_answer():
    """computeInternalMass adds the children in loop order subGrids[0][0],
    [0][1], [1][0], [1][1], i.e. q = 0, 2, 1, 3 (q & 1 -> +x, q & 2 -> +y).
    One particle per quadrant, 2^24 in quadrant 2 and 1 in the others: in
    float, 1 + 2^24 rounds (to even) to 2^24, which absorbs the two later
    units, so the root mass is exactly 2^24.  The fold q = 0, 1, 2, 3 gives
    2 + 2^24 = 2^24 + 2, then 2^24 + 4."""
    x = np.array([1.0, 7.0, 1.0, 7.0], np.float32)  # Morton order: q = 0, 1, 2, 3
    y = np.array([1.0, 1.0, 7.0, 7.0], np.float32)
    m = np.array([1.0, 1.0, 2.0 ** 24, 1.0], np.float32)
    bb = O.bbox(x, y)
    np.testing.assert_array_equal(bb, [1.0, 1.0, 6.0])
    nodes = O.build_tree(x, y, m, bb)
    assert len(nodes) == 5 and nodes[0]["next"] == 4
    assert nodes[0]["mass"] == np.float32(2.0 ** 24)
    assert list(nodes[0]["child"]) == [1, 2, 3, 4]
    np.testing.assert_array_equal(nodes["mass"][1:], m)


# ---- MAC softening placement (physics.cpp:239-252) -------------------------

def _mac_scene():
    # A = (0, 0) and B = (8, 8) set the root square [0, 8]^2.  Quadrant 3
    # holds the pair P1 = (4.6, 5), P2 = (5.4, 5) in the leaf [4, 6]^2 (size
    # 2, COM (5, 5)) and B in the leaf [6, 8]^2; quadrant 1 holds the targets
    # T = (5, 1.3) and T2 = (5, 1.7) in the leaf [4, 6] x [0, 2].
    pts = {"A": (0.0, 0.0), "B": (8.0, 8.0), "P1": (4.6, 5.0), "P2": (5.4, 5.0), "T": (5.0, 1.3), "T2": (5.0, 1.7)}
    names = list(pts)
    x = np.array([pts[k][0] for k in names], np.float32)
    y = np.array([pts[k][1] for k in names], np.float32)
    bb = O.bbox(x, y)
    perm = O.sort_perm(O.morton_keys(x, y, bb))
    return [names[i] for i in perm], x[perm], y[perm], bb


def _soft(src, tgt, msrc, eps2):
    d = np.asarray(src, np.float64) - np.asarray(tgt, np.float64)
    return msrc * d / (d @ d + eps2) ** 1.5


def test_mac_softening_placement_flips_acceptance():
    """With theta = 0.5 and eps = 2 the size-2 pair leaf (s^2 = 4) is accepted
    when 4 < 0.25 (d^2 + 4), i.e. d^2 > 12.  From T, d^2 = 13.69: accepted
    (dropping eps^2 or using eps unsquared would open it).  From T2,
    d^2 = 10.89: opened (adding eps^2 outside theta^2 would accept it).
    Every other node's decision is the same for all these forms."""
    names, x, y, bb = _mac_scene()
    m = np.ones(len(x), np.float32)
    nodes = O.build_tree(x, y, m, bb)
    ax, ay, _ = O.bh_accel(x, y, m, nodes, 1.0, 0.5, 2.0)
    P = {k: (float(x[i]), float(y[i])) for i, k in enumerate(names)}
    eps2 = 4.0
    com = ((P["P1"][0] + P["P2"][0]) / 2, (P["P1"][1] + P["P2"][1]) / 2)
    want_T = (_soft(P["A"], P["T"], 1, eps2) + _soft(P["T2"], P["T"], 1, eps2) + _soft(com, P["T"], 2, eps2)
              + _soft(P["B"], P["T"], 1, eps2))
    want_T2 = (_soft(P["A"], P["T2"], 1, eps2) + _soft(P["T"], P["T2"], 1, eps2) + _soft(P["P1"], P["T2"], 1, eps2)
               + _soft(P["P2"], P["T2"], 1, eps2) + _soft(P["B"], P["T2"], 1, eps2))
    # the flipped decisions, for contrast
    open_T = want_T - _soft(com, P["T"], 2, eps2) + _soft(P["P1"], P["T"], 1, eps2) + _soft(P["P2"], P["T"], 1, eps2)
    acc_T2 = (want_T2 - _soft(P["P1"], P["T2"], 1, eps2) - _soft(P["P2"], P["T2"], 1, eps2)
              + _soft(com, P["T2"], 2, eps2))
    for k, want, flipped in (("T", want_T, open_T), ("T2", want_T2, acc_T2)):
        i = names.index(k)
        got = np.array([ax[i], ay[i]], np.float64)
        tol = 1e-5 * np.linalg.norm(want)
        assert np.linalg.norm(got - want) < tol, (k, got, want)
        assert np.linalg.norm(flipped - want) > 100 * tol, k


# ---- SPH speed clamp in integrateStart (physics.cpp:1086-1100) -------------

def test_sph_speed_clamp_closed_form():
    """With SPH on, a half-kicked speed above sphMaxVel is scaled back to it,
    and prevVel too (unless its square is at most 1e-5); pinned particles are
    skipped; with SPH off nothing is clamped."""
    vmax, dt = 250.0, 0.01
    st = {"x": np.zeros(4), "y": np.zeros(4), "vx": [300.0, 100.0, 0.0, 300.0], "vy": [400.0, 0.0, 0.001, 400.0],
          "ax": [0.0, 0.0, 60000.0, 0.0], "ay": [0.0, 0.0, 0.0, 0.0], "mass": np.ones(4),
          "flags": [O.GEO_SPH, O.GEO_SPH, O.GEO_SPH, O.GEO_SPH | O.GEO_PINNED]}
    s = O.State(st, sph=True)
    s.integrate_start(dt, 1, vmax)
    # particle 0: |v| = 500 -> (150, 200); prevVel (300, 400) -> (150, 200)
    np.testing.assert_allclose([s["vx"][0], s["vy"][0]], [150.0, 200.0], rtol=1e-6)
    np.testing.assert_allclose([s["pvx"][0], s["pvy"][0]], [150.0, 200.0], rtol=1e-6)
    np.testing.assert_allclose([s["x"][0], s["y"][0]], [1.5, 2.0], rtol=1e-6)
    # particle 1 below the limit: unchanged
    assert (s["vx"][1], s["pvx"][1]) == (100.0, 100.0)
    # particle 2: the half kick takes it to (300, 0.001) > 250; prevVel^2 = 1e-6 stays
    np.testing.assert_allclose([s["vx"][2], s["vy"][2]], [250.0, 0.001 * 250.0 / math.hypot(300.0, 0.001)],
                               rtol=1e-6)
    assert s["pvy"][2] == np.float32(0.001) and s["pvx"][2] == 0.0
    # particle 3 pinned: untouched
    assert (s["vx"][3], s["vy"][3], s["x"][3]) == (300.0, 400.0, 0.0)
    s2 = O.State(st, sph=True)
    s2.integrate_start(dt, 0, vmax)
    assert (s2["vx"][0], s2["vy"][0]) == (300.0, 400.0)
