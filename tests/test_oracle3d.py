"""Pins the 3D CPU restatement (oracle/ge_oracle3d.c) before it checks the GPU.

Each function is checked against something independent of its own code:

  geo3_spread_bits / geo3_morton -> a magic-mask spread written here and
      hand-worked known answers (single bits, all-ones = 2^63 - 1)
  geo3_scale_to_grid -> float32 numpy evaluation of the reference formula and
      the end points (clamp at 2097151, degenerate range -> 0)
  geo3_bbox        -> numpy min / max
  geo3_build_tree  -> hand-built trees (one particle per octant: eight
      children in octant order; the z -> y -> x partition of a two-particle
      case), a mass-order known answer (2^24 in octant 4 swallows the unit
      masses only in the reference order 0,4,2,6,1,5,3,7), and on random sets
      the structural invariants: child ranges tile the parent, pre-order
      `next`, leaf rule, every particle inside its leaf cube, FP64 masses
  geo3_bh_accel    -> theta = 0 reduces the walk to the direct sum over j != i
      (same j order); theta = 0.7 error bound against FP64
  geo3_direct_accel -> two-body closed form with the unsquared softening,
      momentum conservation, FP64 direct sum
  periodic wrap    -> a pair straddling the cube faces attracts through them
  geo3_integrate_* -> the closed-form kick-drift(-kick) update and the wrap
      into [-W/2, W/2)
"""
import math

import numpy as np
import pytest

import oracle3d as O3


def spread_magic(v):
    """Bits of a 21-bit value to every third bit (magic-mask form)."""
    v &= 0x1FFFFF
    v = (v | (v << 32)) & 0x1F00000000FFFF
    v = (v | (v << 16)) & 0x1F0000FF0000FF
    v = (v | (v << 8)) & 0x100F00F00F00F00F
    v = (v | (v << 4)) & 0x10C30C30C30C30C3
    v = (v | (v << 2)) & 0x1249249249249249
    return v


def test_spread_bits_matches_magic_masks():
    rng = np.random.default_rng(3)
    vals = [0, 1, 2, 3, 0x1FFFFF, 0x155555, 0xAAAAA] + [int(v) for v in rng.integers(0, 1 << 21, 200)]
    for v in vals:
        assert O3.spread_bits(v) == spread_magic(v), v


def test_morton3d_known_answers():
    assert O3.morton3d(1, 0, 0) == 1
    assert O3.morton3d(0, 1, 0) == 2
    assert O3.morton3d(0, 0, 1) == 4
    assert O3.morton3d(1, 1, 1) == 7
    assert O3.morton3d(2, 0, 0) == 8
    assert O3.morton3d(0, 0, 2) == 32
    assert O3.morton3d(0x1FFFFF, 0, 0) == 0x1249249249249249
    assert O3.morton3d(0x1FFFFF, 0x1FFFFF, 0x1FFFFF) == (1 << 63) - 1
    # a key is monotone in each coordinate with the others fixed
    assert O3.morton3d(5, 9, 3) < O3.morton3d(6, 9, 3)


def test_scale_to_grid_3d():
    f = np.float32
    assert O3.scale_to_grid(0.0, 0.0, 1.0) == 0
    assert O3.scale_to_grid(1.0, 0.0, 1.0) == 2097151
    assert O3.scale_to_grid(2.0, 0.0, 1.0) == 2097151        # clamped
    assert O3.scale_to_grid(-1.0, 0.0, 1.0) == 0
    assert O3.scale_to_grid(0.5, 1.0, 1.0) == 0              # maxVal <= minVal
    rng = np.random.default_rng(4)
    lo, hi = f(-123.25), f(877.5)
    for p in rng.uniform(-123.25, 877.5, 300).astype(np.float32):
        norm = (p - lo) / (hi - lo)
        want = min(int(np.float32(norm * f(2097151.0))), 2097151)
        assert O3.scale_to_grid(p, lo, hi) == want


def _cloud(n, seed, spread=100.0):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, spread, n).astype(np.float32)
    y = rng.normal(0, spread, n).astype(np.float32)
    z = rng.normal(0, spread * 0.3, n).astype(np.float32)
    m = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return x, y, z, m


def _sorted(x, y, z, m):
    bb = O3.bbox(x, y, z)
    k = O3.morton_keys(x, y, z, bb)
    p = O3.sort_perm(k)
    return x[p], y[p], z[p], m[p], bb, k[p]


def test_bbox_3d():
    x, y, z, m = _cloud(1000, 1)
    bb = O3.bbox(x, y, z)
    mn = np.array([x.min(), y.min(), z.min()], np.float32)
    mx = np.array([x.max(), y.max(), z.max()], np.float32)
    size = np.float32(max(mx - mn))
    assert bb[3] == size
    np.testing.assert_array_equal(bb[:3], ((mn + mx) * np.float32(0.5) - size * np.float32(0.5)).astype(np.float32))


def test_keys_sorted_and_stable():
    x, y, z, m = _cloud(3000, 2)
    x[10:20] = x[5]; y[10:20] = y[5]; z[10:20] = z[5]  # equal keys
    bb = O3.bbox(x, y, z)
    k = O3.morton_keys(x, y, z, bb)
    p = O3.sort_perm(k)
    np.testing.assert_array_equal(p, np.argsort(k, kind="stable"))


def test_octree_one_particle_per_octant():
    # corners of a cube of side 10: each octant holds one particle
    pts = [(x, y, z) for z in (0.5, 9.5) for y in (0.5, 9.5) for x in (0.5, 9.5)]
    x = np.array([p[0] for p in pts], np.float32)
    y = np.array([p[1] for p in pts], np.float32)
    z = np.array([p[2] for p in pts], np.float32)
    m = np.arange(1, 9, dtype=np.float32)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    assert len(nodes) == 9
    root = nodes[0]
    assert root["next"] == 8
    # children in octant order q = x + 2y + 4z at pre-order indices 1..8
    for q in range(8):
        c = nodes[root["child"][q]]
        assert root["child"][q] == q + 1
        assert c["end"] - c["start"] == 1
        assert (c["posx"] > bb[0]) == bool(q & 1)
        assert (c["posy"] > bb[1]) == bool(q & 2)
        assert (c["posz"] > bb[2]) == bool(q & 4)
    assert root["mass"] == np.float32(36.0)


def test_octree_partition_order_z_first():
    # two particles differing in x and z.  The cube is centred on y = 0, and
    # y < mid is false at y == mid, so both are in the high y half: octants
    # 1 + 2 = 3 (x high) and 4 + 2 = 6 (z high)
    x = np.array([0.0, 10.0], np.float32)
    y = np.array([0.0, 0.0], np.float32)
    z = np.array([10.0, 0.0], np.float32)
    m = np.ones(2, np.float32)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    # sorted order: the x-high particle (key bit 0) before the z-high one (bit 2)
    assert xs[0] == 10.0 and zs[1] == 10.0
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    root = nodes[0]
    assert root["child"][3] == 1 and root["child"][6] == 2
    assert sum(int(c != 0xFFFFFFFF) for c in root["child"]) == 2
    assert nodes[1]["start"] == 0 and nodes[2]["start"] == 1


def test_octree_mass_order_known_answer():
    # octant 0 and octants 1..3,5..7 hold unit masses, octant 4 holds 2^24.
    # computeInternalMass3D sums q = 0,4,2,6,1,5,3,7: 1 + 2^24 rounds to 2^24
    # and each later +1 is lost, giving exactly 2^24 (sequential q order would
    # give 2^24 + 4).
    pts = [(x, y, z) for z in (0.5, 9.5) for y in (0.5, 9.5) for x in (0.5, 9.5)]
    x = np.array([p[0] for p in pts], np.float32)
    y = np.array([p[1] for p in pts], np.float32)
    z = np.array([p[2] for p in pts], np.float32)
    m = np.ones(8, np.float32)
    m[4] = np.float32(2.0 ** 24)  # octant 4: x low, y low, z high
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    assert nodes[0]["mass"] == np.float32(2.0 ** 24)


def _check_tree(nodes, xs, ys, zs, ms, leaf_cap=0):
    n = len(xs)
    assert nodes[0]["start"] == 0 and nodes[0]["end"] == n
    assert nodes[0]["next"] == len(nodes) - 1
    leaf_cover = np.zeros(n, np.int32)
    for i, nd in enumerate(nodes):
        s, e = int(nd["start"]), int(nd["end"])
        kids = [int(c) for c in nd["child"] if c != 0xFFFFFFFF]
        cnt = e - s
        if nd["next"] == 0:
            assert not kids
            if leaf_cap:
                assert cnt <= leaf_cap or nd["size"] <= 1e-6
            else:
                assert (cnt <= 4 and nd["size"] <= 2.0) or cnt == 1 or nd["size"] <= 1e-6
            leaf_cover[s:e] += 1
            # every particle inside its leaf cube (closed, rounding slack)
            sl = slice(s, e)
            tol = 1e-4 * max(1.0, abs(float(nd["posx"])) + float(nd["size"]))
            for c, p in ((xs, "posx"), (ys, "posy"), (zs, "posz")):
                assert np.all(c[sl] >= nd[p] - tol) and np.all(c[sl] <= nd[p] + nd["size"] + tol)
            mass64 = float(np.sum(ms[sl].astype(np.float64)))
            assert abs(float(nd["mass"]) - mass64) <= 1e-5 * mass64
        else:
            # children in pre-order, ranges tiling [s, e) in octant order
            pos = i + 1
            cur = s
            for q in range(8):
                c = int(nd["child"][q])
                if c == 0xFFFFFFFF:
                    continue
                assert c == pos
                ch = nodes[c]
                assert ch["start"] == cur
                cur = int(ch["end"])
                assert ch["size"] == np.float32(nd["size"] * np.float32(0.5))
                half = np.float32(nd["size"] * np.float32(0.5))
                assert ch["posx"] == np.float32(nd["posx"] + (half if q & 1 else np.float32(0)))
                assert ch["posy"] == np.float32(nd["posy"] + (half if q & 2 else np.float32(0)))
                assert ch["posz"] == np.float32(nd["posz"] + (half if q & 4 else np.float32(0)))
                pos = c + int(ch["next"]) + 1
            assert cur == e
            assert pos == i + int(nd["next"]) + 1
    assert np.all(leaf_cover == 1)


@pytest.mark.parametrize("seed,leaf_cap", [(5, 0), (6, 0), (7, 8), (8, 32)])
def test_octree_invariants(seed, leaf_cap):
    x, y, z, m = _cloud(4000, seed)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb, leaf_cap)
    _check_tree(nodes, xs, ys, zs, ms, leaf_cap)


def test_octree_coincident_particles_end_at_size_floor():
    x = np.full(40, 3.0, np.float32)
    y = np.full(40, -2.0, np.float32)
    z = np.full(40, 1.0, np.float32)
    x[0] = 50.0
    m = np.ones(40, np.float32)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    _check_tree(nodes, xs, ys, zs, ms)
    leaves = nodes[nodes["next"] == 0]
    assert np.any((leaves["end"] - leaves["start"] == 39) & (leaves["size"] <= 1e-6))


def test_walk_theta0_equals_direct_sum():
    x, y, z, m = _cloud(1500, 9)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    ax, ay, az, cnt = O3.bh_accel(xs, ys, zs, ms, nodes, 1.0, 0.0, 1.0)
    dx, dy, dz = O3.direct_accel(xs, ys, zs, ms, 1.0, 1.0)
    assert cnt == len(xs) * (len(xs) - 1)
    for a, d in ((ax, dx), (ay, dy), (az, dz)):
        np.testing.assert_allclose(a, d, rtol=2e-5, atol=1e-6 * np.abs(d).max())


def test_walk_error_bound_theta07():
    x, y, z, m = _cloud(6000, 10)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    ax, ay, az, cnt = O3.bh_accel(xs, ys, zs, ms, nodes, 1.0, 0.7, 1.0)
    rx, ry, rz = O3.direct_fp64(xs, ys, zs, ms, 1.0, 1.0)
    err = np.sqrt((ax - rx) ** 2 + (ay - ry) ** 2 + (az - rz) ** 2)
    ref = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    rms = np.sqrt(np.mean(err ** 2)) / np.sqrt(np.mean(ref ** 2))
    assert rms < 2e-2
    assert cnt < len(xs) * (len(xs) - 1) // 2


def test_direct_two_body_closed_form():
    x = np.array([0.0, 3.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    z = np.array([0.0, 12.0], np.float32)
    m = np.array([2.0, 5.0], np.float32)
    eps = 0.5
    ax, ay, az = O3.direct_accel(x, y, z, m, 2.0, eps)
    r2 = 9.0 + 16.0 + 144.0 + eps  # softening added unsquared
    k = 2.0 / r2 ** 1.5
    np.testing.assert_allclose([ax[0], ay[0], az[0]], [k * 5 * 3, k * 5 * 4, k * 5 * 12], rtol=1e-6)
    np.testing.assert_allclose([ax[1], ay[1], az[1]], [-k * 2 * 3, -k * 2 * 4, -k * 2 * 12], rtol=1e-6)
    fx, fy, fz = O3.direct_fp64(x, y, z, m, 2.0, eps)
    np.testing.assert_allclose(fx, ax, rtol=1e-6)
    np.testing.assert_allclose(fz, az, rtol=1e-6)


def test_direct_momentum_conserved():
    x, y, z, m = _cloud(800, 11)
    ax, ay, az = O3.direct_accel(x, y, z, m, 1.0, 1.0)
    for a in (ax, ay, az):
        tot = np.sum(m.astype(np.float64) * a)
        assert abs(tot) < 1e-4 * np.sum(np.abs(m * a))


def test_periodic_pair_attracts_through_faces():
    W = 100.0
    eps = 0.01  # the direct sum includes j == i, so eps = 0 would give 0 * inf
    x = np.array([-48.0, 48.0], np.float32)
    y = np.zeros(2, np.float32)
    z = np.zeros(2, np.float32)
    m = np.ones(2, np.float32)
    ax, _, _ = O3.direct_accel(x, y, z, m, 1.0, eps, domain=(W, W, W))
    # separation wraps from +96 to -4: particle 0 is pulled in -x
    a = 4.0 / (16.0 + eps) ** 1.5
    np.testing.assert_allclose(ax[0], -a, rtol=1e-6)
    np.testing.assert_allclose(ax[1], a, rtol=1e-6)
    xs, ys, zs, ms, bb, _ = _sorted(x, y, z, m)
    nodes = O3.build_tree(xs, ys, zs, ms, bb)
    bx, _, _, _ = O3.bh_accel(xs, ys, zs, ms, nodes, 1.0, 0.7, eps, domain=(W, W, W))
    np.testing.assert_allclose(np.sort(bx), [-a, a], rtol=1e-6)


def test_integrator_closed_form_and_wrap():
    n = 4
    st = dict(x=np.array([0.0, 10.0, 49.9, -49.9], np.float32), y=np.zeros(n, np.float32),
              z=np.zeros(n, np.float32), vx=np.array([1.0, -2.0, 1.0, -1.0], np.float32),
              vy=np.zeros(n, np.float32), vz=np.array([0.5, 0.0, 0.0, 0.0], np.float32),
              ax=np.array([2.0, 0.0, 0.0, 0.0], np.float32), ay=np.zeros(n, np.float32),
              az=np.array([-1.0, 0.0, 0.0, 0.0], np.float32), mass=np.ones(n, np.float32),
              flags=np.array([0, 1, 0, 0], np.uint32))
    s = O3.State3(st)
    dt = np.float32(0.5)
    s.integrate_start(dt, domain=(100.0, 100.0, 100.0))
    # kick half, drift
    assert s["vx"][0] == np.float32(1.0 + 2.0 * 0.25)
    assert s["x"][0] == np.float32(0.0 + 1.5 * 0.5)
    assert s["vz"][0] == np.float32(0.5 - 0.25)
    assert s["pvx"][0] == np.float32(1.0)
    assert s["x"][1] == np.float32(10.0)        # pinned
    assert s["x"][2] == np.float32(np.float32(49.9 + 0.5) - 100.0)   # >= W/2 wraps down
    assert s["x"][3] == np.float32(np.float32(-49.9 - 0.5) + 100.0)  # < -W/2 wraps up
    s.integrate_end(dt)
    assert s["vx"][0] == np.float32(1.5 + 2.0 * 0.25)


def test_step_direct_conserves_momentum():
    x, y, z, m = _cloud(600, 12, spread=20.0)
    m[:] = 1.0
    s = O3.State3(dict(x=x, y=y, z=z, mass=m))
    s.step(dict(G=1.0, softening=1.0, dt=0.01), nsteps=5)
    for v in ("vx", "vy", "vz"):
        assert abs(float(np.sum(s[v].astype(np.float64)))) < 1e-3
    # ids stay a permutation
    assert sorted(s["id"].tolist()) == list(range(600))


# ---------------------------------------------------------------- SPH3D ---

def _sph_cloud(n, seed, spread=12.0, offset=-3.0):
    rng = np.random.default_rng(seed)
    # negative coordinates too, so floor and truncation would differ
    return [(rng.uniform(0.0, spread, n) + offset).astype(np.float32) for _ in range(3)]


def test_sph3_kernel_constants_closed_form_and_normalisation():
    from scipy.integrate import quad
    for h in (3.0, 4.0, 2.5):
        vol, spiky, lap = (float(v) for v in O3.sph_constants(h))
        # SPH3D.h:27-46: 2 pi h^5 / 15, -45 / (pi h^6), 45 / (pi h^6)
        assert vol == pytest.approx(2.0 * math.pi * h ** 5 / 15.0, rel=1e-6)
        assert spiky == pytest.approx(-45.0 / (math.pi * h ** 6), rel=1e-6)
        assert lap == pytest.approx(45.0 / (math.pi * h ** 6), rel=1e-6)
        # the density kernel (h - r)^2 / vol integrates to 1 over the ball
        integral, _ = quad(lambda r: 4.0 * math.pi * r * r * (h - r) ** 2 / vol, 0.0, h)
        assert integral == pytest.approx(1.0, rel=1e-6)


def test_sph3_candidates_match_brute_force_cells():
    x, y, z = _sph_cloud(1500, 40)
    rng = np.random.default_rng(41)
    qx, qy, qz = [(rng.uniform(-4.0, 10.0, 300)).astype(np.float32) for _ in range(3)]
    cell = 3.0
    off, idx = O3.sph_candidates(x, y, z, qx, qy, qz, cell)
    c = np.stack([np.floor(x / np.float32(cell)), np.floor(y / np.float32(cell)),
                  np.floor(z / np.float32(cell))], 1).astype(np.int64)
    for q in range(len(qx)):
        cq = np.floor(np.array([qx[q], qy[q], qz[q]]) / np.float32(cell)).astype(np.int64)
        want = np.nonzero(np.all(np.abs(c - cq) <= 1, axis=1))[0]
        np.testing.assert_array_equal(idx[off[q]:off[q + 1]], want)
    # a particle is its own candidate (no distance filter, self included)
    off2, idx2 = O3.sph_candidates(x, y, z, x, y, z, cell)
    for i in range(0, 1500, 97):
        assert i in idx2[off2[i]:off2[i + 1]]


def _sph3_state(n, seed, spread=10.0, moving=True, pinned=()):
    x, y, z = _sph_cloud(n, seed, spread, offset=-spread / 2)
    rng = np.random.default_rng(seed + 1)
    st = dict(x=x, y=y, z=z, mass=np.ones(n, np.float32),
              vx=(rng.normal(0, 1, n) if moving else np.zeros(n)).astype(np.float32),
              vy=(rng.normal(0, 1, n) if moving else np.zeros(n)).astype(np.float32),
              vz=(rng.normal(0, 1, n) if moving else np.zeros(n)).astype(np.float32),
              flags=np.full(n, 2, np.uint32), sph_mass=rng.uniform(0.8, 1.2, n).astype(np.float32),
              rest_dens=np.full(n, 0.2, np.float32), stiff=np.ones(n, np.float32),
              visc=np.ones(n, np.float32), cohesion=np.ones(n, np.float32))
    for i in pinned:
        st["flags"][i] |= 1
    return st


SPH3_PRM = dict(dt=0.002, sph=1, sph_h=3.0, sph_cell=3.0, sph_unit_mass=0.03, viscosity=0.3,
                cohesion_coef=1.0, delta=19000.0, stiff_mult=1.0)


def test_sph3_pcisph_conserves_momentum():
    # every pair force is applied to both particles with opposite signs, so
    # the sum of sphMass * pressF vanishes when nothing is pinned
    s = O3.State3(_sph3_state(800, 42))
    s.sph_pcisph(SPH3_PRM)
    f = [s["sph_mass"].astype(np.float64) * s[k].astype(np.float64) for k in ("pfx", "pfy", "pfz")]
    scale = sum(np.abs(v).sum() for v in f)
    assert scale > 0
    for v in f:
        assert abs(v.sum()) < 1e-5 * scale
    assert np.all(s["pred_dens"] > 0) and np.all(s["press"] >= 0)


def test_sph3_two_particles_closed_form():
    # two SPH particles at rest, r < h: viscosity vanishes, cohesion pulls
    # them together, pressure pushes them apart; the pressure pair force is
    # applied by both visits (SPH3D.cpp:238-291), so each feels it twice
    h, r = 3.0, 1.7
    st = dict(x=np.array([0.2, 0.2 + r], np.float32), y=np.full(2, 0.4, np.float32),
              z=np.full(2, 0.4, np.float32), mass=np.ones(2, np.float32), flags=np.full(2, 2, np.uint32),
              sph_mass=np.ones(2, np.float32), rest_dens=np.full(2, 0.05, np.float32),
              stiff=np.ones(2, np.float32), visc=np.ones(2, np.float32), cohesion=np.ones(2, np.float32))
    prm = dict(SPH3_PRM, delta=10.0)
    s = O3.State3(st)
    s.sph_pcisph(prm)
    vol = 2.0 * math.pi * h ** 5 / 15.0
    unit = prm["sph_unit_mass"]
    q = r / h
    wcoh = (1 - q) * (0.5 - q) ** 2 * 30.0 / (math.pi * h ** 3)
    coh = prm["cohesion_coef"] * 1.0 * unit * wcoh   # on particle 0, along +x, from both visits
    f_coh = 2.0 * coh
    dt = prm["dt"]
    # predicted positions move by the cohesion displacement
    disp = f_coh / 1.0 * 0.5 * dt * dt
    rp = r - 2.0 * disp
    m2 = 2.0 * unit
    dens = m2 * (h ** 2 / vol) / 0.05 + m2 * ((h - rp) ** 2 / vol) / 0.05
    press = max(prm["delta"] * (dens - 0.05), 0.0)
    grad = -45.0 / (math.pi * h ** 6) * (h - rp) ** 2
    mag = -(m2 + m2) * press / max(dens, 0.01)
    fp = mag * grad * (-1.0)          # along x for particle 0 (d = x0 - x1 < 0)
    want0 = f_coh + 2.0 * fp
    assert s["pred_dens"][0] == pytest.approx(dens, rel=1e-5)
    assert s["press"][0] == pytest.approx(press, rel=1e-4)
    assert s["pfx"][0] == pytest.approx(want0, rel=1e-4)
    assert s["pfx"][1] == pytest.approx(-want0, rel=1e-4)
    assert abs(s["pfy"][0]) < 1e-6 * abs(want0)


def test_sph3_ground_walls():
    st = dict(x=np.array([-49.5, 49.0, 0.0, 0.0], np.float32), y=np.array([0.0, -49.9, 0.0, 0.0], np.float32),
              z=np.array([0.0, 0.0, 49.8, 0.0], np.float32), vx=np.array([-2.0, 1.0, 1.0, 1.0], np.float32),
              vy=np.array([1.0, -3.0, 1.0, 1.0], np.float32), vz=np.array([1.0, 1.0, 4.0, 1.0], np.float32),
              mass=np.ones(4, np.float32), flags=np.array([2, 2, 2, 3], np.uint32))
    s = O3.State3(st)
    s.sph_ground(3.0, 100.0, 100.0, 100.0, 3.0, 0.25)
    # x wall: position clamped, vx damped by -0.1, friction on vy, vz
    assert s["x"][0] == np.float32(-50.0 + 3.0) and s["vx"][0] == np.float32(-2.0 * -0.1)
    assert s["vy"][0] == np.float32(0.75) and s["vz"][0] == np.float32(0.75)
    assert s["ay"][0] == np.float32(-3.0)
    # particle 1 hits the +x and the -y walls
    assert s["x"][1] == np.float32(47.0) and s["y"][1] == np.float32(-47.0)
    assert s["vx"][1] == np.float32(np.float32(1.0 * -0.1) * 0.75)
    assert s["vy"][1] == np.float32(np.float32(-3.0 * -0.1) * 0.75)
    assert s["z"][2] == np.float32(47.0) and s["vz"][2] == np.float32(np.float32(4.0 * -0.1))
    # pinned: untouched
    assert s["ay"][3] == 0.0 and s["vx"][3] == 1.0
