"""GPU parity of the SPH path (grid + PCISPH) through the C ABI against the oracle.

Neighbour candidate sets are compared bit-exactly (as sets per particle, the
reference's queryNeighborsAVX2 semantics).  PCISPH fields are float sums in a
different order than the reference's atomics, so they are compared with a
tolerance: |gpu - ref| <= 1e-4 * max|ref| + 1e-4 * |ref| per field.
"""
import numpy as np
import pytest

import oracle as O
from galaxy_step import GE_FLAG_DARK, GE_FLAG_DRAWN, GE_FLAG_PINNED, GE_FLAG_SPH, Context, configs, default_params

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _sph_block(n, seed, spacing=1.4, jitter=0.3, h=3.0):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel()[:n] * spacing + rng.normal(0, jitter, n)).astype(np.float32)
    y = (gy.ravel()[:n] * spacing + rng.normal(0, jitter, n)).astype(np.float32)
    st = {"x": x, "y": y, "vx": rng.normal(0, 1, n).astype(np.float32),
          "vy": rng.normal(0, 1, n).astype(np.float32), "mass": np.ones(n, np.float32),
          "flags": np.full(n, GE_FLAG_SPH, np.uint32), "id": np.arange(n, dtype=np.uint32),
          "sph_mass": rng.uniform(0.8, 1.2, n).astype(np.float32),
          "rest_dens": np.full(n, 0.008, np.float32), "stiff": np.ones(n, np.float32),
          "visc": np.ones(n, np.float32), "cohesion": np.ones(n, np.float32),
          "dens": np.zeros(n, np.float32)}
    return st


def _params(**kw):
    p = default_params(G=0.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, sph=1, sph_h=3.0,
                       sph_cell=3.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_prm(p):
    return {k: getattr(p, k) for k in ("G", "theta", "softening", "dt", "dt_start", "sph", "sph_h",
                                       "sph_cell", "sph_unit_mass", "viscosity", "cohesion_coef",
                                       "delta", "stiff_mult", "sph_max_vel")}


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


def _close(got, ref, name, rtol=RTOL):
    scale = float(np.abs(ref).max()) or 1.0
    err = np.abs(got.astype(np.float64) - ref)
    tol = rtol * scale + rtol * np.abs(ref)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (name, bad[:8], got[bad[:8]], ref[bad[:8]], scale)


def _mixed_flags(st, seed):
    rng = np.random.default_rng(seed)
    n = len(st["x"])
    f = np.full(n, GE_FLAG_SPH, np.uint32)
    r = rng.random(n)
    f[r < 0.05] = 0                                  # not SPH
    f[(r >= 0.05) & (r < 0.10)] |= GE_FLAG_PINNED    # pinned
    f[(r >= 0.10) & (r < 0.13)] |= GE_FLAG_DRAWN     # being drawn
    st["flags"] = f
    return st


@pytest.mark.parametrize("cell", [3.0, 4.0])
def test_neighbor_sets_bitexact_at_pos(cell):
    st = _sph_block(3000, 1)
    # a few far-away particles: sparse cells and large coordinates
    st["x"][:5] += np.float32(500.0)
    p = _params(sph_cell=cell, sph_h=cell)
    with Context(3000, sph=True) as c:
        c.upload(st)
        c.sph_new_grid(p)
        off, idx = c.neighbors(0)
        d = c.download(["x", "y"])
    ro, ri = O.sph_candidates(d["x"], d["y"], d["x"], d["y"], cell)
    np.testing.assert_array_equal(off, ro)
    np.testing.assert_array_equal(idx, ri)


def test_neighbor_sets_bitexact_at_predpos():
    st = _sph_block(2500, 2)
    p = _params()
    with Context(2500, sph=True) as c:
        c.upload(st)
        c.sph_new_grid(p)
        c.compute_gravity(p)  # G = 0: acc = 0
        c.sph_pcisph(p)
        off, idx = c.neighbors(1)
        d = c.download(["x", "y", "ppx", "ppy"])
    ro, ri = O.sph_candidates(d["x"], d["y"], d["ppx"], d["ppy"], 3.0)
    np.testing.assert_array_equal(off, ro)
    np.testing.assert_array_equal(idx, ri)


def _compress_corner(st, lim=20.0, f=0.3):
    """Squeezes the lattice's corner below (lim, lim) by f about its mean: its
    queries keep about 160 neighbours within h = 3, more than a per-particle
    list holds, so the SPH passes rescan their candidates there."""
    sel = (st["x"] < lim) & (st["y"] < lim)
    for k in ("x", "y"):
        v = st[k][sel]
        st[k][sel] = (v.mean() + (v - v.mean()) * f).astype(np.float32)
    return st


@pytest.mark.parametrize("mixed,dense", [(False, False), (True, False), (True, True)])
def test_pcisph_one_frame_matches_oracle(mixed, dense):
    n = 4000
    st = _sph_block(n, 3)
    if mixed:
        st = _mixed_flags(st, 4)
    if dense:
        st = _compress_corner(st)
    p = _params()
    s = O.State(st, sph=True)
    s.step(_oracle_prm(p), nsteps=1)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "vx", "vy", "ax", "ay", "pred_dens", "press",
                                             "press_tmp", "ppx", "ppy", "pfx", "pfy", "id")})
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        got = _by_id(c.download(["x", "y", "vx", "vy", "ax", "ay", "pred_dens", "press",
                                 "press_tmp", "ppx", "ppy", "pfx", "pfy", "id"]))
    np.testing.assert_array_equal(got["x"], ref["x"])  # drift happens before any SPH force
    for f in ("pred_dens", "press", "press_tmp", "ppx", "ppy", "pfx", "pfy", "ax", "ay", "vx", "vy"):
        _close(got[f], ref[f], f)


def test_speed_clamp_frame_matches_oracle():
    """integrateStart's SPH speed clamp (physics.cpp:1086-1100): particles
    whose half-kicked speed passes sphMaxVel (250) are scaled back to it, and
    their prevVel too; the drift that follows is bit-exact."""
    n = 4000
    st = _mixed_flags(_sph_block(n, 9), 10)
    rng = np.random.default_rng(11)
    fast = rng.random(n) < 0.08
    ang = rng.uniform(0, 2 * np.pi, n)
    st["vx"][fast] = (400.0 * np.cos(ang[fast])).astype(np.float32)
    st["vy"][fast] = (400.0 * np.sin(ang[fast])).astype(np.float32)
    kick = (~fast) & (rng.random(n) < 0.05)  # the half kick alone passes the limit
    st["ax"] = np.zeros(n, np.float32)
    st["ax"][kick] = np.float32(3.0e5)
    st["vx"][kick & (rng.random(n) < 0.5)] = 0.0  # prevVel below 1e-5: not rescaled
    st["vy"][kick] = np.where(st["vx"][kick] == 0.0, np.float32(0.001), st["vy"][kick])
    p = _params()
    s = O.State(st, sph=True)
    s.step(_oracle_prm(p), nsteps=1)
    fields = ("x", "y", "vx", "vy", "pvx", "pvy", "ax", "ay", "pfx", "pfy", "id")
    ref = _by_id({k: s[k].copy() for k in fields})
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        got = _by_id(c.download(list(fields)))
    clamped = np.hypot(ref["pvx"], ref["pvy"]) < np.hypot(st["vx"], st["vy"])[np.argsort(st["id"])] * 0.99
    assert clamped.sum() > 0.05 * n  # the clamp fired
    for f in ("x", "y", "pvx", "pvy"):
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    for f in ("vx", "vy", "pfx", "pfy"):
        _close(got[f], ref[f], f)


def test_pcisph_momentum_conserved_gpu():
    n = 5000
    st = _sph_block(n, 5)
    st["sph_mass"][:] = 1.0
    p = _params()
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        d = c.download(["pfx", "pfy"])
    fx, fy = d["pfx"].astype(np.float64), d["pfy"].astype(np.float64)
    scale = np.abs(fx).sum() + np.abs(fy).sum()
    assert scale > 0
    assert abs(fx.sum()) < 1e-4 * scale and abs(fy.sum()) < 1e-4 * scale


def test_impact_with_gravity_trajectory():
    st, prm = configs.config2(n_target=8000, n_impactor=2000, r_target=60.0, r_impactor=24.0)
    n = len(st["x"])
    p = _params(G=1.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, sph_h=4.0, sph_cell=4.0)
    s = O.State(st, sph=True)
    s.step(_oracle_prm(p), nsteps=5)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "vx", "vy", "pred_dens", "id")})
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 5)
        c.sync()
        got = _by_id(c.download(["x", "y", "vx", "vy", "pred_dens", "id"]))
    ext = float(np.ptp(ref["x"]))
    assert np.abs(got["x"] - ref["x"]).max() < 1e-4 * ext
    assert np.abs(got["y"] - ref["y"]).max() < 1e-4 * ext
    _close(got["vx"], ref["vx"], "vx", rtol=1e-3)
    _close(got["pred_dens"], ref["pred_dens"], "pred_dens", rtol=1e-3)


def test_sph_requires_context_flag():
    with Context(100) as c:
        c.upload({"x": np.arange(100.0), "y": np.zeros(100), "mass": np.ones(100)})
        with pytest.raises(Exception):
            c.step(_params(), 1)


# ---- NeighborSearchV2 semantics (the reference's path without AVX2)

@pytest.mark.parametrize("cell,radius", [(3.0, 3.5), (4.0, 4.5)])
def test_neighbor_sets_v2_bitexact(cell, radius):
    st = _sph_block(3000, 6)
    st["x"] -= np.float32(40.0)  # negative coordinates: truncation differs from floor
    st["x"][:4] -= np.float32(300.0)
    p = _params(sph_cell=cell, sph_h=cell, sph_search=1, sph_search_radius=radius)
    with Context(3000, sph=True) as c:
        c.upload(st)
        c.sph_new_grid(p)
        off, idx = c.neighbors(0)
        d = c.download(["x", "y"])
        c.compute_gravity(p)
        c.sph_pcisph(p)
        off1, idx1 = c.neighbors(1)
        d1 = c.download(["x", "y", "ppx", "ppy"])
    ro, ri = O.sph_candidates(d["x"], d["y"], d["x"], d["y"], cell, mode=1, radius=radius)
    np.testing.assert_array_equal(off, ro)
    np.testing.assert_array_equal(idx, ri)
    ro, ri = O.sph_candidates(d1["x"], d1["y"], d1["ppx"], d1["ppy"], cell, mode=1, radius=radius)
    np.testing.assert_array_equal(off1, ro)
    np.testing.assert_array_equal(idx1, ri)


@pytest.mark.parametrize("mixed,dense", [(False, False), (True, False), (True, True)])
def test_pcisph_v2_one_frame_matches_oracle(mixed, dense):
    n = 4000
    st = _sph_block(n, 7)
    if dense:
        st = _compress_corner(st)
    st["x"] -= np.float32(30.0)
    if mixed:
        st = _mixed_flags(st, 8)
    p = _params(sph_search=1, sph_search_radius=3.5)
    s = O.State(st, sph=True)
    s.step(dict(_oracle_prm(p), sph_search=1, sph_search_radius=3.5), nsteps=1)
    ref = _by_id({k: s[k].copy() for k in ("x", "vx", "vy", "pred_dens", "press", "pfx", "pfy", "id")})
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        got = _by_id(c.download(["x", "vx", "vy", "pred_dens", "press", "pfx", "pfy", "id"]))
    np.testing.assert_array_equal(got["x"], ref["x"])
    for f in ("pred_dens", "press", "pfx", "pfy", "vx", "vy"):
        _close(got[f], ref[f], f)


def test_neighbor_counts_match_oracle():
    n = 5000
    st = _sph_block(n, 9, spacing=1.1)
    st["x"] -= np.float32(35.0)
    st = _mixed_flags(st, 10)
    st["flags"][::29] |= GE_FLAG_DARK
    p = _params(sph_cell=3.0, sph_search_radius=3.5)
    with Context(n, sph=True) as c:
        c.upload(st)
        got = c.neighbor_counts(p)
        d = c.download(["x", "y", "flags"])
    want = O.neighbor_amount(d["x"], d["y"], d["flags"], 3.0, 3.5)
    np.testing.assert_array_equal(got, want)


def test_ground_mode_matches_oracle():
    # SPH::groundModeBoundary: vertical gravity and damped walls at h from the
    # domain edges, between PCISPH and integrateEnd
    n = 3000
    st = _sph_block(n, 11)
    st["y"] = (st["y"] + np.float32(1.0)).astype(np.float32)   # the bottom rows start inside the floor band
    st["vy"] = (st["vy"] - np.float32(5.0)).astype(np.float32)
    W, H = 200.0, 300.0
    p = _params(sph_ground=1, domain_w=W, domain_h=H, vertical_gravity=3.0, boundary_friction=0.2)
    s = O.State(st, sph=True)
    s.step(dict(_oracle_prm(p), sph_ground=1, domain_w=W, domain_h=H, vertical_gravity=3.0,
                boundary_friction=0.2), nsteps=2)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "vx", "vy", "ay", "id")})
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 2)
        c.sync()
        got = _by_id(c.download(["x", "y", "vx", "vy", "ay", "id"]))
    assert got["y"].min() >= np.float32(3.0)  # floor at radiusMultiplier = h
    for f in ("x", "y", "vx", "vy", "ay"):
        _close(got[f], ref[f], f, rtol=1e-4)
