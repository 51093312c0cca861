"""GPU parity of SPH3D (grid + PCISPH + ground walls) through the 3D C ABI
against the oracle's restatement of SPH3D.cpp.

Candidate sets (queryNeighborsAVX2: the 27 cells around the query cell, no
distance filter, self included) are compared bit-exactly as sets per
particle.  PCISPH fields are float sums in another order than the
reference's atomics, so they are compared with a tolerance:
|gpu - ref| <= 1e-4 * max|ref| + 1e-4 * |ref| per field.
"""
import numpy as np
import pytest

import oracle3d as O3
from galaxy_step import GE_FLAG_DARK, GE_FLAG_DRAWN, GE_FLAG_PINNED, GE_FLAG_SPH, configs
from galaxy_step.lib3d import Context3D, default_params3

pytestmark = pytest.mark.gpu

RTOL = 1e-4
ORACLE_KEYS = ("G", "theta", "softening", "dt", "dt_start", "sph", "sph_h", "sph_cell", "sph_unit_mass",
               "viscosity", "cohesion_coef", "delta", "stiff_mult", "sph_max_vel", "sph_ground",
               "vertical_gravity", "boundary_friction", "domain_w", "domain_h", "domain_d")


def _sph_block3(n, seed, spacing=1.3, jitter=0.25):
    """A jittered cubic lattice of SPH particles (about 30 neighbours within h = 3)."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"), -1)
    g = g.reshape(-1, 3)[:n].astype(np.float64)
    x, y, z = [(g[:, k] * spacing - side * spacing / 2 + rng.normal(0, jitter, n)).astype(np.float32)
               for k in range(3)]
    rest = np.float32(np.sqrt(2.0 * 0.03 / spacing ** 3))
    return {"x": x, "y": y, "z": z, "vx": rng.normal(0, 1, n).astype(np.float32),
            "vy": rng.normal(0, 1, n).astype(np.float32), "vz": rng.normal(0, 1, n).astype(np.float32),
            "mass": np.ones(n, np.float32), "flags": np.full(n, GE_FLAG_SPH, np.uint32),
            "id": np.arange(n, dtype=np.uint32), "sph_mass": rng.uniform(0.8, 1.2, n).astype(np.float32),
            "rest_dens": np.full(n, rest, np.float32), "stiff": np.ones(n, np.float32),
            "visc": np.ones(n, np.float32), "cohesion": np.ones(n, np.float32), "dens": np.zeros(n, np.float32)}


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


def _params(**kw):
    p = default_params3(G=0.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, sph=1, direct_below=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_prm(p):
    return {k: getattr(p, k) for k in ORACLE_KEYS}


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


def _close(got, ref, name, rtol=RTOL):
    scale = float(np.abs(ref).max()) or 1.0
    err = np.abs(got.astype(np.float64) - ref)
    tol = rtol * scale + rtol * np.abs(ref)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (name, bad[:8], got[bad[:8]], ref[bad[:8]], scale)


def test_neighbor_sets_bitexact_at_pos():
    n = 4000
    st = _sph_block3(n, 1)
    st["x"][:5] += np.float32(400.0)   # sparse far cells and large coordinates
    st["z"][5:9] -= np.float32(250.0)
    p = _params()
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.sph_new_grid(p)
        off, idx = c.neighbors(0)
        d = c.download(["x", "y", "z"])
    ro, ri = O3.sph_candidates(d["x"], d["y"], d["z"], d["x"], d["y"], d["z"], 3.0)
    np.testing.assert_array_equal(off, ro)
    np.testing.assert_array_equal(idx, ri)


def test_neighbor_sets_bitexact_at_predpos():
    n = 3000
    st = _sph_block3(n, 2)
    p = _params()
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.sph_new_grid(p)
        c.compute_gravity(p)  # G = 0: acc = 0
        c.sph_pcisph(p)
        off, idx = c.neighbors(1)
        d = c.download(["x", "y", "z", "ppx", "ppy", "ppz"])
    ro, ri = O3.sph_candidates(d["x"], d["y"], d["z"], d["ppx"], d["ppy"], d["ppz"], 3.0)
    np.testing.assert_array_equal(off, ro)
    np.testing.assert_array_equal(idx, ri)


@pytest.mark.parametrize("mixed,dense", [(False, False), (True, False), (True, True)])
def test_pcisph3_one_frame_matches_oracle(mixed, dense):
    n = 5000
    st = _sph_block3(n, 3)
    if mixed:
        st = _mixed_flags(st, 4)
    if dense:
        # a compressed corner: its queries keep more neighbours (about 250
        # within h) than a per-particle list holds, so the passes rescan
        # their candidates there and use the lists elsewhere
        sel = (st["x"] < -3.0) & (st["y"] < -3.0)
        for f in ("x", "y", "z"):
            v = st[f][sel]
            st[f][sel] = (v.mean() + (v - v.mean()) * 0.55).astype(np.float32)
    p = _params()
    s = O3.State3(st)
    s.step(_oracle_prm(p), nsteps=1)
    keys = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "pred_dens", "press", "press_tmp",
            "ppx", "ppy", "ppz", "pfx", "pfy", "pfz", "id")
    ref = _by_id({k: s[k].copy() for k in keys})
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        got = _by_id(c.download(list(keys)))
    np.testing.assert_array_equal(got["x"], ref["x"])  # the drift comes before any SPH force
    for f in keys[3:-1]:
        _close(got[f], ref[f], f)


def test_pcisph3_momentum_conserved_gpu():
    n = 6000
    st = _sph_block3(n, 5)
    st["sph_mass"][:] = 1.0
    p = _params()
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.step(p, 1)
        c.sync()
        d = c.download(["pfx", "pfy", "pfz"])
    f = [d[k].astype(np.float64) for k in ("pfx", "pfy", "pfz")]
    scale = sum(np.abs(v).sum() for v in f)
    assert scale > 0
    for v in f:
        assert abs(v.sum()) < 1e-4 * scale


def test_impact3_with_gravity_trajectory():
    st, prm = configs.config3d_sph(n_target=12000, n_impactor=3000)
    n = len(st["x"])
    p = _params(G=1.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002)
    s = O3.State3(st)
    s.step(_oracle_prm(p), nsteps=4)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "z", "vx", "vy", "vz", "pred_dens", "id")})
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.step(p, 4)
        c.sync()
        got = _by_id(c.download(["x", "y", "z", "vx", "vy", "vz", "pred_dens", "id"]))
    ext = float(np.ptp(ref["x"]))
    for f in ("x", "y", "z"):
        assert np.abs(got[f] - ref[f]).max() < 1e-4 * ext, f
    for f in ("vx", "vy", "vz"):
        _close(got[f], ref[f], f, rtol=1e-3)
    _close(got["pred_dens"], ref["pred_dens"], "pred_dens", rtol=1e-3)


def test_ground3_mode_matches_oracle():
    # SPH3D::groundModeBoundary: vertical gravity and damped walls at h inside
    # +-domain/2, between PCISPH and integrateEnd3D
    n = 3000
    st = _sph_block3(n, 11)
    W = 30.0   # the lattice (about 19 wide) plus its velocities reaches the walls
    st["vy"] = (st["vy"] - np.float32(30.0)).astype(np.float32)
    st["y"] = (st["y"] - np.float32(3.0)).astype(np.float32)
    p = _params(sph_ground=1, domain_w=W, domain_h=W, domain_d=W, vertical_gravity=3.0, boundary_friction=0.2)
    s = O3.State3(st)
    s.step(_oracle_prm(p), nsteps=2)
    ref = _by_id({k: s[k].copy() for k in ("x", "y", "z", "vx", "vy", "vz", "ay", "id")})
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.step(p, 2)
        c.sync()
        got = _by_id(c.download(["x", "y", "z", "vx", "vy", "vz", "ay", "id"]))
    assert got["y"].min() >= np.float32(-W / 2 + 3.0)  # floor at radiusMultiplier inside the wall
    for f in ("x", "y", "z", "vx", "vy", "vz", "ay"):
        _close(got[f], ref[f], f, rtol=1e-4)


def test_neighbor_counts3_match_oracle():
    n = 4000
    st = _mixed_flags(_sph_block3(n, 9, spacing=1.1), 10)
    st["flags"][::29] |= GE_FLAG_DARK
    p = _params()
    with Context3D(n, sph=True) as c:
        c.upload(st)
        got = c.neighbor_counts(p)
        d = c.download(["x", "y", "z", "flags"])
    off, _ = O3.sph_candidates(d["x"], d["y"], d["z"], d["x"], d["y"], d["z"], 3.0)
    want = np.diff(off).astype(np.uint32)
    want[(d["flags"] & (GE_FLAG_PINNED | GE_FLAG_DRAWN | GE_FLAG_DARK)) != 0] = 0
    np.testing.assert_array_equal(got, want)


def test_sph3_rejections():
    st = _sph_block3(500, 12)
    with Context3D(500) as c:  # no SPH state
        c.upload(st)
        with pytest.raises(Exception):
            c.step(_params(), 1)
    with Context3D(500, sph=True) as c:
        c.upload(st)
        with pytest.raises(Exception):   # PCISPH needs a grid over the current positions
            c.sph_pcisph(_params())
        with pytest.raises(Exception):
            c.step(_params(sph_h=0.0), 1)
