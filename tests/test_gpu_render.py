"""GPU parity of the steps either side of the frame (include/ge_render.h)
against the oracle restatement (oracle/render.py): render fields that travel
with the particles, neighbour counts and the size / colour consumers
(bit-exact: integer counts and colour bytes, float sizes from the same
float operations), PLY vertices (bit-exact) and text (byte-exact), and the
galaxy spawner (masses, dark-matter velocities, flags, ids and render fields
bit-exact from the same draws; libm-dependent positions and velocities within
2e-5 relative)."""
import numpy as np
import pytest

import oracle as O
import oracle3d as O3
import render as R
from galaxy_step import configs, default_params
from galaxy_step import render as GR
from galaxy_step.lib import Context
from galaxy_step.lib3d import Context3D, default_params3

pytestmark = pytest.mark.gpu

F = np.float32


def _block2(n, seed, spacing=1.2):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel()[:n] * spacing + rng.normal(0, 0.4, n)).astype(F)
    y = (gy.ravel()[:n] * spacing + rng.normal(0, 0.4, n)).astype(F)
    return {"x": x, "y": y, "vx": rng.normal(0, 1, n).astype(F), "vy": rng.normal(0, 1, n).astype(F),
            "ax": rng.normal(0, 25, n).astype(F), "ay": rng.normal(0, 25, n).astype(F),
            "mass": np.ones(n, F), "id": np.arange(n, dtype=np.uint32)}


def _flags(n, seed):
    rng = np.random.default_rng(seed)
    choices = [0, R.FLAG_SPH, R.FLAG_DARK, R.FLAG_SOLID, R.FLAG_UNIQUE, R.FLAG_PINNED, R.FLAG_DRAWN,
               R.FLAG_UNIQUE | R.FLAG_DARK]
    return rng.choice(choices, n, p=[0.4, 0.2, 0.08, 0.07, 0.1, 0.05, 0.05, 0.05]).astype(np.uint32)


def _render_fields(n, seed):
    rng = np.random.default_rng(seed)
    r32 = lambda: rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)  # noqa: E731
    return {"color": r32(), "pcolor": r32(), "scolor": r32(), "size": rng.uniform(0.1, 2.0, n).astype(F),
            "neighbors": rng.integers(0, 300, n).astype(np.int32),
            "previous_size": rng.uniform(0.1, 2.0, n).astype(F)}


ALL_MODES = dict(density_size=1, force_size=1, solid_color=1, density_color=1, force_color=1)


def test_render_fields_travel_with_the_particles():
    st, _ = configs.config0(n=20000)
    n = len(st["x"])
    rf = _render_fields(n, 1)
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0)
    with Context(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        c.step(p, 3)
        c.sync()
        ids = c.download(["id"])["id"]
        got = GR.render_download(c)
    assert not np.array_equal(ids, np.arange(n))  # the sort moved particles
    for k in GR.RENDER_FIELDS:
        np.testing.assert_array_equal(got[k], rf[k][ids], k)


def test_upload_gives_render_defaults():
    st = _block2(1000, 2)
    with Context(1000, render=True) as c:
        c.upload(st)
        got = GR.render_download(c)
    assert np.all(got["color"] == R.WHITE) and np.all(got["pcolor"] == R.WHITE) and np.all(got["scolor"] == R.WHITE)
    assert np.all(got["size"] == 1.0) and np.all(got["neighbors"] == 0) and np.all(got["previous_size"] == 1.0)


@pytest.mark.parametrize("search", [0, 1])
def test_visuals_2d_match_oracle(search):
    n = 6000
    st = _block2(n, 3)
    st["flags"] = _flags(n, 4)
    rf = _render_fields(n, 5)
    v = GR.default_visual_params(search=search, **ALL_MODES)
    with Context(n, sph=True, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        GR.visuals(c, v)
        got = GR.render_download(c)
        d = c.download(["x", "y", "ax", "ay", "flags"])
    off, _ = O.sph_candidates(d["x"], d["y"], d["x"], d["y"], 3.0, mode=search, radius=3.5)
    nb = R.neighbor_counts(off, d["flags"], rf["neighbors"])
    np.testing.assert_array_equal(got["neighbors"], nb)
    col, pc, sc, size = R.visuals(d["flags"], d["ax"], d["ay"], None, rf["color"], rf["pcolor"], rf["scolor"],
                                  rf["size"], nb, v.as_dict())
    np.testing.assert_array_equal(got["color"], col)
    np.testing.assert_array_equal(got["pcolor"], pc)
    np.testing.assert_array_equal(got["scolor"], sc)
    np.testing.assert_array_equal(got["size"], size)


def test_visuals_stored_counts_and_single_modes():
    n = 3000
    st = _block2(n, 6)
    st["flags"] = _flags(n, 7)
    rf = _render_fields(n, 8)
    with Context(n, render=True) as c:  # no SPH grid: counts come from the render fields
        c.upload(st)
        GR.render_upload(c, rf)
        for modes in (dict(density_color=1), dict(density_size=1, density_color=0),
                      dict(force_color=1, density_color=0), dict(solid_color=1, density_color=0)):
            v = GR.default_visual_params(count_neighbors=0, **modes)
            GR.render_upload(c, rf)
            GR.visuals(c, v)
            got = GR.render_download(c)
            col, pc, sc, size = R.visuals(st["flags"], st["ax"], st["ay"], None, rf["color"], rf["pcolor"],
                                          rf["scolor"], rf["size"], rf["neighbors"], v.as_dict())
            for k, want in (("color", col), ("pcolor", pc), ("scolor", sc), ("size", size)):
                np.testing.assert_array_equal(got[k], want, (modes, k))
        with pytest.raises(Exception):  # counting needs the SPH grid
            GR.visuals(c, GR.default_visual_params(count_neighbors=1))


def test_visuals_3d_match_oracle():
    n = 5000
    rng = np.random.default_rng(9)
    st = {k: rng.uniform(-14, 14, n).astype(F) for k in ("x", "y", "z")}
    st.update({k: rng.normal(0, 15, n).astype(F) for k in ("ax", "ay", "az")})
    st["mass"] = np.ones(n, F)
    st["flags"] = _flags(n, 10)
    rf = _render_fields(n, 11)
    v = GR.default_visual_params(**ALL_MODES)
    with Context3D(n, sph=True, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        GR.visuals(c, v)
        got = GR.render_download(c)
        d = c.download(["x", "y", "z", "ax", "ay", "az", "flags"])
    off, _ = O3.sph_candidates(d["x"], d["y"], d["z"], d["x"], d["y"], d["z"], 3.0)
    nb = R.neighbor_counts(off, d["flags"], rf["neighbors"])
    np.testing.assert_array_equal(got["neighbors"], nb)
    col, pc, sc, size = R.visuals(d["flags"], d["ax"], d["ay"], d["az"], rf["color"], rf["pcolor"], rf["scolor"],
                                  rf["size"], nb, v.as_dict())
    np.testing.assert_array_equal(got["color"], col)
    np.testing.assert_array_equal(got["size"], size)
    np.testing.assert_array_equal(got["pcolor"], pc)
    with Context3D(n, sph=True, render=True) as c:
        c.upload(st)
        with pytest.raises(Exception):  # 3D counts follow the AVX2 search only
            GR.visuals(c, GR.default_visual_params(search=1))


def test_ply_2d_vertices_and_file(tmp_path):
    n = 50000
    st, _ = configs.config0(n=n)
    n = len(st["x"])
    st["flags"] = _flags(n, 12)
    rf = _render_fields(n, 13)
    with Context(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        v = GR.ply_vertices(c, 3840.0, 2160.0)
        path = tmp_path / "frame.ply"
        GR.ply_write(c, path, 3840.0, 2160.0)
    want = R.ply_vertices(st["x"], st["y"], None, st["flags"], rf["color"], rf["size"], 3840.0, 2160.0)
    assert v.dtype == want.dtype or v.dtype.names == want.dtype.names
    for k in ("x", "y", "z", "r", "g", "b", "radius"):
        np.testing.assert_array_equal(v[k], want[k], k)
    assert path.read_bytes() == R.ply_text(want)


def test_ply_3d_vertices_and_file(tmp_path):
    n = 20000
    rng = np.random.default_rng(14)
    st = {k: rng.normal(0, 50, n).astype(F) for k in ("x", "y", "z")}
    st["mass"] = np.ones(n, F)
    st["flags"] = _flags(n, 15)
    rf = _render_fields(n, 16)
    with Context3D(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        v = GR.ply_vertices(c)
        path = tmp_path / "frame3d.ply"
        GR.ply_write(c, path)
    want = R.ply_vertices(st["x"], st["y"], st["z"], st["flags"], rf["color"], rf["size"])
    for k in ("x", "y", "z", "r", "g", "b", "radius"):
        np.testing.assert_array_equal(v[k], want[k], k)
    assert path.read_bytes() == R.ply_text(want)


def _close(got, want, scale, rtol=2e-5):
    err = np.abs(got.astype(np.float64) - want)
    assert np.all(err <= rtol * (np.abs(want) + scale)), (np.argmax(err), err.max())


def _check_galaxy(got, want, nv):
    for k in ("mass", "flags", "id", "sph_mass"):
        np.testing.assert_array_equal(got[k], want[k], k)
    np.testing.assert_array_equal(got["vx"][nv:], want["vx"][nv:])  # integer draws
    np.testing.assert_array_equal(got["vy"][nv:], want["vy"][nv:])
    _close(got["x"], want["x"], 1.0)
    _close(got["y"], want["y"], 1.0)
    _close(got["vx"][:nv], want["vx"][:nv], 1.0)
    _close(got["vy"][:nv], want["vy"][:nv], 1.0)


def test_spawn_galaxy_device_draws_match_oracle():
    g = GR.default_galaxy_params(visible=40000, dark=12000, center_x=120.0, center_y=-40.0, extra_vx=2.5,
                                 extra_vy=-1.5, first_id=1000)
    seed = 20250101
    with Context(60000, sph=True, render=True) as c:
        first = GR.spawn_galaxy(c, g, seed=seed)
        got = c.download(["x", "y", "vx", "vy", "mass", "flags", "id", "sph_mass", "rest_dens", "ax", "pvx"])
        rnd = GR.render_download(c)
    assert first == 0
    want = R.spawn_galaxy(g.as_dict(), R.galaxy_uniforms(seed, 40000, 12000))
    _check_galaxy(got, want, 40000)
    np.testing.assert_array_equal(rnd["color"], want["color"])
    np.testing.assert_array_equal(rnd["size"], want["size"])
    assert np.all(got["rest_dens"] == F(0.008)) and np.all(got["ax"] == 0) and np.all(got["pvx"] == 0)


def test_spawn_galaxy_appends_with_given_draws_then_steps():
    st, _ = configs.config0(n=5000)
    n0 = len(st["x"])
    g = GR.default_galaxy_params(visible=8000, dark=2000, first_id=n0, mass_visible=1.0, mass_dark=5.0)
    u = np.random.default_rng(17).random(3 * 8000 + 5 * 2000).astype(F)
    with Context(n0 + 10000) as c:
        c.upload(st)
        first = GR.spawn_galaxy(c, g, uniforms=u)
        got = c.download(["x", "y", "vx", "vy", "mass", "flags", "id"])
        p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0)
        c.step(p, 1)
        c.sync()
        after = c.download(["x", "id"])
        with pytest.raises(Exception):  # over capacity
            GR.spawn_galaxy(c, g)
    assert first == n0
    np.testing.assert_array_equal(got["x"][:n0], st["x"])
    want = R.spawn_galaxy(g.as_dict(), u)
    tail = {k: v[n0:] for k, v in got.items()}
    want["sph_mass"] = tail["sph_mass"] = np.zeros(1)
    _check_galaxy(tail, want, 8000)
    assert np.all(np.isfinite(after["x"])) and sorted(after["id"].tolist()) == list(range(n0 + 10000))


def test_spawn_galaxy_rejects_bad_draws():
    g = GR.default_galaxy_params(visible=10, dark=2)
    u = np.full(3 * 10 + 5 * 2, 0.5, F)
    u[4] = 1.5
    with Context(100) as c:
        with pytest.raises(Exception):
            GR.spawn_galaxy(c, g, uniforms=u)
