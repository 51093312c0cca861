"""Seeded synthetic inputs for the benchmark configurations.

The reference has no dataset; its galaxy spawner (GE/src/Particles/
particlesSpawning.cpp:130-252) samples an exponential radial profile with a
uniform angle and circular speeds.  These generators follow that pattern with
the physical constants of each workload (G = 1, explicit masses).
"""
import math

import numpy as np

from .lib import GE_FLAG_SPH


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def exp_disk(rng, n_disk, scale, clip, central_mass, G=1.0, noise=0.02,
             center=(0.0, 0.0), bulk_v=(0.0, 0.0)):
    """One exponential disk plus a central body (last particle).

    Radius r = -s ln(1 - u (1 - exp(-clip/s))) is Exp(scale) truncated at clip;
    speeds are circular from the enclosed mass (central body + inner disk
    particles) with N(0, noise * v_c) added per component.
    """
    u = rng.random(n_disk)
    r = -scale * np.log1p(-u * (1.0 - math.exp(-clip / scale)))
    ang = rng.random(n_disk) * (2.0 * math.pi)
    m = rng.uniform(0.5, 1.5, n_disk)
    order = np.argsort(r, kind="stable")
    enclosed = np.empty(n_disk)
    enclosed[order] = central_mass + np.cumsum(m[order]) - m[order]
    rr = np.maximum(r, 1e-6)
    vc = np.sqrt(G * enclosed / rr)
    c, s = np.cos(ang), np.sin(ang)
    x = r * c
    y = r * s
    vx = -vc * s + rng.normal(0.0, 1.0, n_disk) * noise * vc
    vy = vc * c + rng.normal(0.0, 1.0, n_disk) * noise * vc
    x = np.append(x, 0.0) + center[0]
    y = np.append(y, 0.0) + center[1]
    vx = np.append(vx, 0.0) + bulk_v[0]
    vy = np.append(vy, 0.0) + bulk_v[1]
    m = np.append(m, central_mass)
    return x, y, vx, vy, m


def _state(x, y, vx, vy, m, flags=None):
    n = len(x)
    st = {
        "x": np.asarray(x, np.float32), "y": np.asarray(y, np.float32),
        "vx": np.asarray(vx, np.float32), "vy": np.asarray(vy, np.float32),
        "mass": np.asarray(m, np.float32),
        "flags": np.zeros(n, np.uint32) if flags is None else np.asarray(flags, np.uint32),
        "id": np.arange(n, dtype=np.uint32),
    }
    return st


def config0(seed=0, n=20000, scale=100.0):
    """Plain 2D BH gravity: 20,000 particles, one exponential disk (scale 100,
    clip 600), masses U[0.5,1.5] + a central body of 5e4; G=1, eps=1,
    theta=0.8, dt=0.01, 100 steps."""
    rng = _rng(seed)
    st = _state(*exp_disk(rng, n - 1, scale, 600.0, 5e4))
    return st, dict(G=1.0, theta=0.8, softening=1.0, dt=0.01, steps=100)


def config1(seed=1, n=1_000_000, theta=0.7):
    """Two-disk collision: 2 x 500k (scale 150, clip 600) offset to
    (-1200,-300) / (+1200,+300) moving (+3,0) / (-3,0); eps=1, dt=0.005."""
    rng = _rng(seed)
    h = n // 2
    a = exp_disk(rng, h - 1, 150.0, 600.0, 5e4, center=(-1200.0, -300.0), bulk_v=(3.0, 0.0))
    b = exp_disk(rng, n - h - 1, 150.0, 600.0, 5e4, center=(1200.0, 300.0), bulk_v=(-3.0, 0.0))
    st = _state(*[np.concatenate([p, q]) for p, q in zip(a, b)])
    return st, dict(G=1.0, theta=theta, softening=1.0, dt=0.005, steps=200)


def uniform_disk(rng, n, radius):
    r = radius * np.sqrt(rng.random(n))
    ang = rng.random(n) * (2.0 * math.pi)
    return r * np.cos(ang), r * np.sin(ang)


def config2(seed=2, n_target=200_000, n_impactor=50_000, r_target=300.0, r_impactor=120.0,
            speed=8.0):
    """Gravity + SPH impact: uniform disks of 200k (R=300, at rest at the
    origin) and 50k (R=120) placed at (-440, 150) moving (+8, 0): impact
    parameter 0.5 R.  m=1, sphMass=1, restDens 0.008, stiff/visc/cohesion 1."""
    rng = _rng(seed)
    tx, ty = uniform_disk(rng, n_target, r_target)
    ix, iy = uniform_disk(rng, n_impactor, r_impactor)
    off_x = -(r_target + r_impactor + 20.0)
    off_y = 0.5 * r_target
    x = np.concatenate([tx, ix + off_x])
    y = np.concatenate([ty, iy + off_y])
    vx = np.concatenate([np.zeros(n_target), np.full(n_impactor, speed)])
    vy = np.zeros(n_target + n_impactor)
    n = n_target + n_impactor
    st = _state(x, y, vx, vy, np.ones(n), flags=np.full(n, GE_FLAG_SPH, np.uint32))
    st.update({
        "sph_mass": np.ones(n, np.float32), "rest_dens": np.full(n, 0.008, np.float32),
        "stiff": np.ones(n, np.float32), "visc": np.ones(n, np.float32),
        "cohesion": np.ones(n, np.float32), "dens": np.zeros(n, np.float32),
    })
    prm = dict(G=1.0, theta=0.7, softening=1.0, dt=0.002, steps=500, sph=1, sph_h=4.0,
               sph_cell=4.0, sph_unit_mass=0.03, viscosity=0.3, cohesion_coef=1.0,
               delta=19000.0, stiff_mult=1.0, sph_max_vel=250.0)
    return st, prm


def config3_sizes(seed=3, systems=512, lo=1000, hi=131072):
    rng = _rng(seed)
    n = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), systems))).astype(np.int64)
    scales = rng.uniform(50.0, 300.0, systems)
    return n, scales


def config3_system(k, n, scale, seed=3):
    """System k of the batched ensemble: a config-0 disk of n particles."""
    rng = _rng(seed * 1_000_003 + k)
    st = _state(*exp_disk(rng, int(n) - 1, float(scale), 600.0, 5e4))
    return st, dict(G=1.0, theta=0.7, softening=1.0, dt=0.01, steps=100)


def config4_stream(rate, window_s=60.0, seed=4, lo=2000, hi=500_000):
    """Poisson request stream: arrival time, particle count, steps, kind."""
    rng = _rng(seed * 7919 + int(rate))
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window_s:
            break
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        steps = int(rng.integers(10, 201))
        kind = "sph" if rng.random() < 0.3 else "gravity"
        out.append((t, n, steps, kind, int(rng.integers(0, 2**31))))
    return out


def config4_request(n, kind, seed):
    """One served request: a config-0 disk of n particles (kind "gravity"),
    or a config-2 impact scaled to n particles (kind "sph": 80% target, 20%
    impactor, radii scaled by sqrt(n / 250,000))."""
    if kind == "sph":
        nt = int(round(0.8 * n))
        k = math.sqrt(n / 250_000.0)
        st, prm = config2(seed=seed, n_target=nt, n_impactor=n - nt, r_target=300.0 * k,
                          r_impactor=120.0 * k)
        return st, prm
    st = _state(*exp_disk(_rng(seed), int(n) - 1, 100.0, 600.0, 5e4))
    return st, dict(G=1.0, theta=0.7, softening=1.0, dt=0.01)


def galaxy3d(rng, n_disk, scale, clip, central_mass, thickness=5.0, bulge=20.0, bulge_size=None, G=1.0,
             noise=0.02, center=(0.0, 0.0, 0.0), bulk_v=(0.0, 0.0, 0.0), tilt=0.0):
    """One 3D disk galaxy plus a central body (last particle), following the
    3D galaxy spawner (GE/src/Particles/particlesSpawning.cpp:642-704):
    exponential radius truncated at clip, uniform angle, vertical offset
    N(0, 1) * (thickness + bulge * exp(-r / bulge_size)), tangential speed in
    the disk plane (circular from the enclosed mass, G = 1, as the 2D
    workloads), then the disk rotated by `tilt` radians about the x axis."""
    bulge_size = scale * 0.5 if bulge_size is None else bulge_size
    u = rng.random(n_disk)
    r = -scale * np.log1p(-u * (1.0 - math.exp(-clip / scale)))
    ang = rng.random(n_disk) * (2.0 * math.pi)
    zoff = rng.normal(0.0, 1.0, n_disk) * (thickness + bulge * np.exp(-r / bulge_size))
    m = rng.uniform(0.5, 1.5, n_disk)
    order = np.argsort(r, kind="stable")
    enclosed = np.empty(n_disk)
    enclosed[order] = central_mass + np.cumsum(m[order]) - m[order]
    vc = np.sqrt(G * enclosed / np.maximum(r, 1e-6))
    c, s = np.cos(ang), np.sin(ang)
    x, y, z = r * c, r * s, zoff
    vx = -vc * s + rng.normal(0.0, 1.0, n_disk) * noise * vc
    vy = vc * c + rng.normal(0.0, 1.0, n_disk) * noise * vc
    vz = rng.normal(0.0, 1.0, n_disk) * noise * vc
    ct, st_ = math.cos(tilt), math.sin(tilt)
    y, z = y * ct - z * st_, y * st_ + z * ct
    vy, vz = vy * ct - vz * st_, vy * st_ + vz * ct
    cols = []
    for a, c0 in ((x, center[0]), (y, center[1]), (z, center[2])):
        cols.append(np.append(a, 0.0) + c0)
    vel = []
    for a, b0 in ((vx, bulk_v[0]), (vy, bulk_v[1]), (vz, bulk_v[2])):
        vel.append(np.append(a, 0.0) + b0)
    return cols[0], cols[1], cols[2], vel[0], vel[1], vel[2], np.append(m, central_mass)


def _state3(x, y, z, vx, vy, vz, m):
    n = len(x)
    f = lambda a: np.asarray(a, np.float32)  # noqa: E731
    return {"x": f(x), "y": f(y), "z": f(z), "vx": f(vx), "vy": f(vy), "vz": f(vz), "mass": f(m),
            "flags": np.zeros(n, np.uint32), "id": np.arange(n, dtype=np.uint32)}


def config3d(seed=5, n=1_000_000, theta=0.7):
    """3D two-galaxy collision: 2 x n/2 particles (scale 150, clip 600,
    central body 5e4) at (-1200, -300, -100) / (+1200, +300, +100) moving
    (+3, 0, 0) / (-3, 0, 0), the second disk tilted 30 degrees; G = 1,
    softening 1 (added unsquared, as the reference's 3D step), dt = 0.005."""
    rng = _rng(seed)
    h = n // 2
    a = galaxy3d(rng, h - 1, 150.0, 600.0, 5e4, center=(-1200.0, -300.0, -100.0), bulk_v=(3.0, 0.0, 0.0))
    b = galaxy3d(rng, n - h - 1, 150.0, 600.0, 5e4, center=(1200.0, 300.0, 100.0), bulk_v=(-3.0, 0.0, 0.0),
                 tilt=math.radians(30.0))
    st = _state3(*[np.concatenate([p, q]) for p, q in zip(a, b)])
    return st, dict(G=1.0, theta=theta, softening=1.0, dt=0.005, steps=200)


def galaxy3d_small(seed=6, n=20_000, scale=100.0):
    """One 3D galaxy of n particles (scale 100, clip 600, central 5e4)."""
    st = _state3(*galaxy3d(_rng(seed), n - 1, scale, 600.0, 5e4))
    return st, dict(G=1.0, theta=0.8, softening=1.0, dt=0.01, steps=100)


def uniform_ball(rng, n, radius):
    """n points uniform in a ball: r = R u^(1/3), isotropic direction."""
    r = radius * np.cbrt(rng.uniform(0.0, 1.0, n))
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(1.0 - cz * cz)
    return r * s * np.cos(ph), r * s * np.sin(ph), r * cz


def config3d_sph(seed=7, n_target=200_000, n_impactor=50_000, density=0.3, speed=8.0):
    """3D gravity + SPH impact: uniform balls of n_target (at rest at the
    origin) and n_impactor particles at `density` particles per unit volume,
    the impactor moving +x at `speed` with impact parameter half the target
    radius.  m = 1, sphMass = 1, stiff / visc / cohesion 1; restDens is set
    so the reference's predicted density (normalised by restDens) starts at
    restDens: restDens^2 = 2 * mass * density.  h = cell = 3 (SPH3D's
    radiusMultiplier and the AVX2 grid's cellSize), dt 0.002."""
    rng = _rng(seed)
    rt = (n_target / density / (4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
    ri = (n_impactor / density / (4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
    tx, ty, tz = uniform_ball(rng, n_target, rt)
    ix, iy, iz = uniform_ball(rng, n_impactor, ri)
    x = np.concatenate([tx, ix - (rt + ri + 6.0)])
    y = np.concatenate([ty, iy + 0.5 * rt])
    z = np.concatenate([tz, iz])
    n = n_target + n_impactor
    vx = np.concatenate([np.zeros(n_target), np.full(n_impactor, speed)])
    st = _state3(x, y, z, vx, np.zeros(n), np.zeros(n), np.ones(n))
    st["flags"] = np.full(n, GE_FLAG_SPH, np.uint32)
    unit = 0.03
    rest = math.sqrt(2.0 * unit * density)
    st.update({
        "sph_mass": np.ones(n, np.float32), "rest_dens": np.full(n, rest, np.float32),
        "stiff": np.ones(n, np.float32), "visc": np.ones(n, np.float32),
        "cohesion": np.ones(n, np.float32), "dens": np.zeros(n, np.float32),
    })
    prm = dict(G=1.0, theta=0.7, softening=1.0, dt=0.002, steps=500, sph=1, sph_h=3.0, sph_cell=3.0,
               sph_unit_mass=unit, viscosity=0.3, cohesion_coef=1.0, delta=19000.0, stiff_mult=1.0,
               sph_max_vel=250.0)
    return st, prm
