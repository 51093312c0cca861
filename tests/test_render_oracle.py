"""Pins of the render / export / spawner oracle (CPU): closed forms,
endpoints, the C library's own printf, the published splitmix64 outputs and
the spawner's distribution laws.  Also the host-side PLY formatter of the
library (ge_ply_format, host code) against the oracle text."""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

import render as R

F = np.float32


def test_color_lerp_endpoints_and_clamp():
    c1, c2 = R.pack([10, 20, 30, 40]), R.pack([250, 0, 130, 255])
    assert R.color_lerp(c1, c2, 0.0) == c1
    assert R.color_lerp(c1, c2, 1.0) == c2
    assert R.color_lerp(c1, c2, -3.0) == c1       # factor clamped below
    assert R.color_lerp(c1, c2, 7.5) == c2        # and above
    # truncation, not rounding: (1 - 0.5) * 0 + 0.5 * 255 = 127.5 -> 127
    assert R.channels(R.color_lerp(R.pack([0, 0, 0, 0]), R.pack([255, 255, 255, 255]), 0.5))[0] == 127
    # 0.25 of the way from 0 to 200 is exactly 50
    assert R.channels(R.color_lerp(R.pack([0, 0, 0, 0]), R.pack([200, 0, 0, 0]), 0.25))[0] == 50


def test_lerp_closed_form():
    assert R.lerp(0.62, 0.17, 0.0) == F(0.62)
    assert R.lerp(0.62, 0.17, 1.0) == F(0.62) + (F(0.17) - F(0.62))
    assert R.lerp(2.0, 6.0, 0.25) == F(3.0)


def _vis_inputs(n=600, seed=0):
    rng = np.random.default_rng(seed)
    flags = rng.choice([0, R.FLAG_SPH, R.FLAG_DARK, R.FLAG_SOLID, R.FLAG_UNIQUE, R.FLAG_UNIQUE | R.FLAG_SOLID],
                       n).astype(np.uint32)
    ax, ay, az = (rng.normal(0, 20, n).astype(F) for _ in range(3))
    col = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    pc = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    sc = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    size = rng.uniform(0.1, 2, n).astype(F)
    nb = rng.integers(0, 400, n).astype(np.int32)
    return flags, ax, ay, az, col, pc, sc, size, nb


def test_density_size_endpoints():
    flags, ax, ay, az, col, pc, sc, size, nb = _vis_inputs()
    nb[:3] = [0, 200, 1000]
    flags[:3] = 0
    _, _, _, s = R.visuals(flags, ax, ay, None, col, pc, sc, size, nb, dict(density_size=1, density_color=0))
    assert s[0] == F(0.62) and s[1] == F(0.62) + (F(0.17) - F(0.62)) and s[2] == s[1]
    dark_or_solid = (flags & (R.FLAG_DARK | R.FLAG_SOLID)) != 0
    np.testing.assert_array_equal(s[dark_or_solid], size[dark_or_solid])  # skipped
    # monotone: more neighbours never give a larger size
    ok = ~dark_or_solid
    o = np.argsort(nb[ok], kind="stable")
    assert np.all(np.diff(s[ok][o]) <= 0)


def test_density_color_endpoints_and_unique():
    flags, ax, ay, az, col, pc, sc, size, nb = _vis_inputs(seed=1)
    P, S = 0x64442800, 0x4B28509B
    c, p, s, _ = R.visuals(flags, ax, ay, None, col, pc, sc, size, nb, dict(density_color=1))
    dark = (flags & R.FLAG_DARK) != 0
    uniq = (flags & R.FLAG_UNIQUE) != 0
    np.testing.assert_array_equal(c[dark], col[dark])                 # dark matter keeps its colour
    lo = ~dark & ~uniq & (nb == 0)
    hi = ~dark & ~uniq & (nb >= 200)
    assert lo.any() and hi.any()
    assert np.all(c[lo] == P) and np.all(c[hi] == S)                  # the two ends of the ramp
    np.testing.assert_array_equal(p[~dark & ~uniq], P)
    np.testing.assert_array_equal(p[uniq], pc[uniq])                  # unique colours stay
    uhi = uniq & ~dark & (nb >= 200)
    np.testing.assert_array_equal(c[uhi], sc[uhi])


def test_force_size_and_color_use_the_acceleration_norm():
    flags, ax, ay, az, col, pc, sc, size, nb = _vis_inputs(seed=2)
    flags[:] = 0
    ax[:2], ay[:2], az[:2] = [0, 300], [0, 0], [0, 0]
    c, _, _, s = R.visuals(flags, ax, ay, az, col, pc, sc, size, nb,
                           dict(force_size=1, density_color=0, force_color=1))
    assert s[0] == F(0.62) and c[0] == 0x64442800                   # zero acceleration
    assert s[1] == R.lerp(0.62, 0.17, 1.0) and c[1] == 0x4B28509B   # saturated
    # a z component counts in 3D: |(3, 0, 4)| = 5
    flags2 = np.zeros(1, np.uint32)
    _, _, _, s3 = R.visuals(flags2, [3.0], [0.0], [4.0], [0], [0], [0], [1.0], [0],
                            dict(force_size=1, density_color=0))
    assert s3[0] == R.lerp(0.62, 0.17, F(5.0) / F(22.0))


def test_neighbor_counts_keep_skipped():
    off = np.array([0, 3, 3, 10, 12], np.int64)
    flags = np.array([0, R.FLAG_PINNED, R.FLAG_DARK, R.FLAG_SPH], np.uint32)
    prev = np.array([7, 8, 9, 10], np.int32)
    np.testing.assert_array_equal(R.neighbor_counts(off, flags, prev), [3, 8, 9, 2])


def test_splitmix64_published_outputs():
    # the first two outputs of splitmix64 seeded with 0 (state advances by
    # the golden gamma before each mix)
    with np.errstate(over="ignore"):
        z = R.splitmix64(np.array([0, 0x9E3779B97F4A7C15], np.uint64))
    assert int(z[0]) == 0xE220A8397B1DCDAF
    assert int(z[1]) == 0x6E789E6AA1B965F4


def test_galaxy_uniforms_layout():
    u = R.galaxy_uniforms(5, 4, 3)
    assert u.shape == (4 * 3 + 3 * 5,)
    assert np.all((u >= 0) & (u < 1))
    with np.errstate(over="ignore"):
        z = R.splitmix64(np.array([5 + 8 * 2 + 1, 5 + (1 << 60) + 8 * 1 + 4], np.uint64))
    assert u[2 * 3 + 1] == F(int(z[0]) >> 40) * F(2.0 ** -24)
    assert u[12 + 1 * 5 + 4] == F(int(z[1]) >> 40) * F(2.0 ** -24)


def _libc_printf(fmt, *args):
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    buf = ctypes.create_string_buffer(256)
    cargs = [ctypes.c_double(a) if isinstance(a, float) else ctypes.c_int(a) for a in args]
    libc.snprintf(buf, 256, fmt.encode(), *cargs)
    return buf.value.decode()


def test_ply_text_matches_the_c_library_printf():
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.normal(0, 1, 200) * 10.0 ** rng.integers(-8, 9, 200),
                           [0.0, -0.0, 1.0, 1234565.0, 1e-45, 3.4e38, 0.1, 123456.5, 1e6, 999999.5]]).astype(F)
    for v in vals:
        assert "%.6g" % float(v) == _libc_printf("%.6g", float(v)), v


def test_ply_vertices_2d_closed_form():
    x = np.array([0.0, 3840.0, 1920.0, 5.0], F)
    y = np.array([0.0, 2160.0, 1080.0, 5.0], F)
    flags = np.array([0, 0, 0, R.FLAG_DARK], np.uint32)
    col = np.array([R.pack([1, 2, 3, 4])] * 4, np.uint32)
    v = R.ply_vertices(x, y, None, flags, col, np.full(4, 0.5, F), 3840.0, 2160.0)
    assert len(v) == 3                                   # dark matter is left out
    ratio = F(F(3840) / F(2160))
    assert v["x"][0] == F(F(-1) * ratio) * F(10) and v["y"][0] == F(-10)
    assert v["x"][1] == F(ratio * F(10)) and v["y"][1] == F(10)
    assert v["x"][2] == 0 and v["y"][2] == 0 and np.all(v["z"] == 0)
    assert list(v[0][["r", "g", "b"]]) == [1, 2, 3]
    t = R.ply_text(v).decode().splitlines()
    assert t[2] == "element vertex 3" and t[10] == "end_header" and len(t) == 14
    assert t[13] == "0 0 0 1 2 3 0.5"


def test_library_ply_format_matches_oracle_text():
    # ge_ply_format is host code (no device involved)
    from galaxy_step import render as GR
    rng = np.random.default_rng(4)
    n = 70000  # above the threshold where the formatter splits the work over threads
    v = np.zeros(n, R.PLY_DTYPE)
    for k in ("x", "y", "z", "radius"):
        v[k] = (rng.normal(0, 1, n) * 10.0 ** rng.integers(-6, 7, n)).astype(F)
    for k in ("r", "g", "b"):
        v[k] = rng.integers(0, 256, n)
    assert GR.ply_format(v) == R.ply_text(v)
    assert GR.ply_format(v[:0]) == R.ply_text(v[:0])


def test_galaxy_laws():
    g = dict(visible=20000, dark=6000, center_x=100.0, center_y=-50.0, extra_vx=3.0, extra_vy=-1.0)
    u = R.galaxy_uniforms(11, g["visible"], g["dark"])
    p = R.spawn_galaxy(g, u)
    nv = g["visible"]
    dx, dy = p["x"][:nv] - 100.0, p["y"][:nv] + 50.0
    r = np.hypot(dx.astype(np.float64), dy)
    np.testing.assert_allclose(r, p["radius"][:nv], rtol=2e-5, atol=1e-3)
    # inverse CDF of the truncated exponential disc
    L, rmax = 90.0, 800.0
    cdf = (1 - np.exp(-np.sort(p["radius"][:nv]) / L)) / (1 - math.exp(-rmax / L))
    ks = np.max(np.abs(cdf - (np.arange(nv) + 0.5) / nv))
    assert ks < 1.63 / math.sqrt(nv)             # 99 % Kolmogorov-Smirnov bound
    # circular orbits: speed law, counter-clockwise (L_z > 0) after removing the slingshot
    vx, vy = p["vx"][:nv] - 3.0, p["vy"][:nv] + 1.0
    sp = np.hypot(vx.astype(np.float64), vy)
    np.testing.assert_allclose(sp, 8.5 * np.sqrt(1758.0 / (p["radius"][:nv] + 54.0)), rtol=1e-4)
    assert np.all(dx * vy - dy * vx > 0)
    np.testing.assert_allclose((dx * vx + dy * vy) / (r * sp), 0, atol=1e-3)
    # masses inside base * (1 +- scatter), sphMass = mass / 8.5e9
    assert p["mass"][:nv].min() >= 8.5e9 * 0.25 * (1 - 1e-6) and p["mass"][:nv].max() <= 8.5e9 * 1.75 * (1 + 1e-6)
    np.testing.assert_allclose(p["sph_mass"][:nv], p["mass"][:nv] / 8.5e9, rtol=1e-6)
    # dark halo: u = log(1 + (r / rc)^2) / log(1 + (R / rc)^2) is uniform
    rd = np.sort(p["radius"][nv:].astype(np.float64))
    ud = np.log1p((rd / 3.5) ** 2) / np.log1p((2000.0 / 3.5) ** 2)
    nd = len(rd)
    assert np.max(np.abs(ud - (np.arange(nd) + 0.5) / nd)) < 1.63 / math.sqrt(nd)
    assert rd.max() <= 2000.0 * (1 + 1e-5)
    # dark velocities are (k - 30) * 0.85 for integer k in [0, 59], plus the slingshot
    kx = (p["vx"][nv:] - 3.0) / 0.85 + 30
    assert np.all(np.abs(kx - np.round(kx)) < 1e-3) and kx.min() > -0.5 and kx.max() < 59.5
    assert np.all(p["flags"][nv:] == (R.FLAG_DARK | R.FLAG_UNIQUE)) and np.all(p["flags"][:nv] == R.FLAG_SPH)
