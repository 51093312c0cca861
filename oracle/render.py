"""CPU restatement of the steps either side of the physics frame: the render
consumers of the neighbour counts, the PLY export and the galaxy spawner.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product library
or its Python binding.  numpy in float32 with the reference's operation
order (no contraction), integer colour channels, Python's correctly rounded
'%.6g' for the PLY text.

Functions and the reference lines they follow:
  color_lerp        raylib ColorLerp (rtextures.c, raylib 5.5; the
                    reference fetches raylib at GIT_TAG master,
                    cmake/raylib.cmake): factor clamped to [0, 1], each
                    channel (unsigned char)((1 - f) * c1 + f * c2)
  lerp              raymath Lerp: start + amount * (end - start)
  neighbor_counts   NeighborSearchV2AVX2::neighborAmount
                    (GE/include/Particles/neighborSearch.h:779-796) and
                    NeighborSearchV2::neighborAmount (:585-598): skipped
                    particles keep their previous count
  visuals           DensitySize::sizeByDensity (densitySize.h:14-79), then
                    ColorVisuals::particlesColorVisuals' solid, density and
                    force modes (particleColorVisuals.h:111-277)
  ply_vertices      plyFileCreation (GE/src/globalLogic.cpp:146-190),
                    plyFileCreation3D (:192-242)
  ply_text          the buffer both write (header + "%.6g %.6g %.6g %d %d %d %.6g\\n")
  splitmix64        Steele, Lea and Flood (OOPSLA 2014), the counter-based
                    draws of the device spawner
  spawn_galaxy      ParticlesSpawning::particlesInitialConditions, galaxy
                    tool (GE/src/Particles/particlesSpawning.cpp:130-244)

Pins (tests/test_render_oracle.py): colour and size endpoints and
clamping, the integer truncation of ColorLerp, the PLY text against the C
library's printf, splitmix64 against its published first outputs, the
spawner's radial laws (inverse CDFs), orbital speed law, orbit sense and mass
range.  libm-dependent spawner values (log, exp, cos, sin, pow) are
compared with a tolerance, not pinned bit for bit.
"""
import numpy as np

F = np.float32
WHITE = 0xFFFFFFFF
FLAG_PINNED, FLAG_SPH, FLAG_DRAWN, FLAG_DARK, FLAG_UNIQUE, FLAG_SOLID = 1, 2, 4, 8, 16, 32
PI_F = F(3.14159265358979323846)

VISUAL_DEFAULTS = dict(count_neighbors=1, search=0, cell=3.0, radius=3.5, density_size=0, force_size=0,
                       min_size=0.17, max_size=0.62, size_acc=22.0, size_max_neighbors=200, size_multiplier=1.0,
                       solid_color=0, density_color=1, force_color=0, p_color=0x64442800, s_color=0x4B28509B,
                       color_max_neighbors=200, min_color_acc=0.0, max_color_acc=40.0)


def channels(c):
    c = np.asarray(c, dtype=np.uint32)
    return [((c >> np.uint32(8 * k)) & np.uint32(255)).astype(np.uint32) for k in range(4)]


def pack(ch):
    out = np.zeros(np.shape(ch[0]), np.uint32)
    for k in range(4):
        out |= (np.asarray(ch[k], np.uint32) & np.uint32(255)) << np.uint32(8 * k)
    return out


def color_lerp(c1, c2, f):
    f = np.asarray(f, dtype=F)
    f = np.where(f < F(0), F(0), np.where(f > F(1), F(1), f)).astype(F)
    a, b = channels(c1), channels(c2)
    out = []
    for k in range(4):
        v = ((F(1) - f) * a[k].astype(F) + f * b[k].astype(F)).astype(F)
        out.append(np.trunc(v).astype(np.uint32))
    return pack(out)


def lerp(a, b, t):
    a, b, t = F(a), F(b), np.asarray(t, F)
    return (a + t * (b - a)).astype(F)


def _clamp(v, lo, hi):
    v = np.asarray(v, F)
    return np.where(v < F(lo), F(lo), np.where(F(hi) < v, F(hi), v)).astype(F)


def _min1(v):
    v = np.asarray(v, F)
    return np.where(F(1) < v, F(1), v).astype(F)


def neighbor_counts(cand_off, flags, previous):
    """Candidate counts from CSR offsets; skipped particles keep `previous`."""
    cnt = np.diff(np.asarray(cand_off)).astype(np.int32)
    skip = (np.asarray(flags, np.uint32) & np.uint32(FLAG_PINNED | FLAG_DRAWN | FLAG_DARK)) != 0
    return np.where(skip, np.asarray(previous, np.int32), cnt).astype(np.int32)


def visuals(flags, ax, ay, az, color, pcolor, scolor, size, nbrs, v):
    """One pass of the consumers; returns (color, pcolor, scolor, size)."""
    v = dict(VISUAL_DEFAULTS, **v)
    flags = np.asarray(flags, np.uint32)
    dark = (flags & FLAG_DARK) != 0
    solid = (flags & FLAG_SOLID) != 0
    unique = (flags & FLAG_UNIQUE) != 0
    color, pcolor, scolor = (np.array(a, np.uint32) for a in (color, pcolor, scolor))
    size = np.array(size, F)
    nbrs = np.asarray(nbrs, np.int32)
    ax, ay = np.asarray(ax, F), np.asarray(ay, F)
    acc_sq = (ax * ax + ay * ay).astype(F)
    if az is not None:
        az = np.asarray(az, F)
        acc_sq = (acc_sq + az * az).astype(F)
    acc = np.sqrt(acc_sq).astype(F)
    mult = F(v["size_multiplier"])
    hi_s, lo_s = F(F(v["max_size"]) * mult), F(F(v["min_size"]) * mult)
    if v["force_size"]:
        norm = (_clamp(acc, 0.0, v["size_acc"]) / F(v["size_acc"])).astype(F)
        sel = ~(solid | dark)
        size[sel] = lerp(hi_s, lo_s, norm)[sel]
    if v["density_size"]:
        nd = _min1((nbrs.astype(F) / F(v["size_max_neighbors"])).astype(F))
        sel = ~(dark | solid)
        size[sel] = lerp(hi_s, lo_s, (nd * nd).astype(F))[sel]
    pc, sc = pcolor.copy(), scolor.copy()
    P, S = np.uint32(v["p_color"]), np.uint32(v["s_color"])
    if v["solid_color"]:
        pc = np.where(unique, pc, P).astype(np.uint32)
        color = pc.copy()
    if v["density_color"]:
        sel = ~dark
        pc = np.where(sel & ~unique, P, pc).astype(np.uint32)
        sc = np.where(sel & ~unique, S, sc).astype(np.uint32)
        inv = F(F(1) / F(v["color_max_neighbors"]))
        nd = _min1((nbrs.astype(F) * inv).astype(F))
        color = np.where(sel, color_lerp(pc, sc, nd), color).astype(np.uint32)
    if v["force_color"]:
        sel = ~dark
        norm = (_clamp(acc, v["min_color_acc"], v["max_color_acc"]) / F(v["max_color_acc"])).astype(F)
        pc = np.where(sel & ~unique, P, pc).astype(np.uint32)
        sc = np.where(sel & ~unique, S, sc).astype(np.uint32)
        color = np.where(sel, color_lerp(pc, sc, norm), color).astype(np.uint32)
    if v["solid_color"] or v["density_color"] or v["force_color"]:
        pcolor, scolor = pc, sc
    return color, pcolor, scolor, size


PLY_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1"), ("g", "u1"), ("b", "u1"),
                      ("pad", "u1"), ("radius", "<f4")])


def ply_vertices(x, y, z, flags, color, size, domain_w=None, domain_h=None):
    """Vertices of the non-dark particles in order; 2D when z is None."""
    keep = (np.asarray(flags, np.uint32) & FLAG_DARK) == 0
    x, y = np.asarray(x, F)[keep], np.asarray(y, F)[keep]
    out = np.zeros(len(x), PLY_DTYPE)
    if z is None:
        W, H = F(domain_w), F(domain_h)
        px = (((x / W).astype(F) * F(2)).astype(F) - F(1)).astype(F)
        py = (((y / H).astype(F) * F(2)).astype(F) - F(1)).astype(F)
        ratio = F(W / H)
        px = (px * ratio).astype(F)
        out["x"] = (px * F(10)).astype(F)
        out["y"] = (py * F(10)).astype(F)
        out["z"] = F(0)
    else:
        out["x"], out["y"], out["z"] = x, y, np.asarray(z, F)[keep]
    ch = channels(np.asarray(color, np.uint32)[keep])
    out["r"], out["g"], out["b"] = ch[0], ch[1], ch[2]
    out["radius"] = np.asarray(size, F)[keep]
    return out


PLY_HEADER = ("ply\nformat ascii 1.0\nelement vertex {}\nproperty float x\nproperty float y\nproperty float z\n"
              "property uchar red\nproperty uchar green\nproperty uchar blue\nproperty float radius\nend_header\n")


def ply_text(v):
    lines = [PLY_HEADER.format(len(v))]
    for q in v:
        lines.append("%.6g %.6g %.6g %d %d %d %.6g\n" % (float(q["x"]), float(q["y"]), float(q["z"]), int(q["r"]),
                                                         int(q["g"]), int(q["b"]), float(q["radius"])))
    return "".join(lines).encode()


_GOLD = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB
_MASK = (1 << 64) - 1


def splitmix64(x):
    """One output per counter value (uint64 array or int)."""
    z = (np.asarray(x, np.uint64) + np.uint64(_GOLD))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(_M1)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(_M2)
    return z ^ (z >> np.uint64(31))


def galaxy_uniforms(seed, visible, dark):
    """The device draws as the flat array the spawner accepts: 3 per visible
    particle, then 5 per dark particle."""
    def stream(s, count, per):
        k = np.arange(count, dtype=np.uint64)[:, None] * np.uint64(8) + np.arange(per, dtype=np.uint64)[None, :]
        ctr = (np.uint64(seed) + np.uint64(s << 60) + k) & np.uint64(_MASK)
        z = splitmix64(ctr)
        return ((z >> np.uint64(40)).astype(np.float64) * 2.0 ** -24).astype(F).ravel()
    with np.errstate(over="ignore"):
        return np.concatenate([stream(0, visible, 3), stream(1, dark, 5)])


GALAXY_DEFAULTS = dict(center_x=0.0, center_y=0.0, extra_vx=0.0, extra_vy=0.0, visible=40000, dark=12000,
                       outer_radius=200.0, scale_length=90.0, outer_radius_dm=2000.0, radius_core_dm=3.5,
                       mass_visible=8.5e9, mass_dark=1.416e11, mass_scatter=0.75, first_id=0)


def spawn_galaxy(g, u):
    """Particles of the galaxy tool from the flat draws u; dict of arrays."""
    g = dict(GALAXY_DEFAULTS, **g)
    nv, nd = int(g["visible"]), int(g["dark"])
    u = np.asarray(u, F)
    uv, ud = u[:3 * nv].reshape(nv, 3), u[3 * nv:].reshape(nd, 5)
    cx, cy, evx, evy = F(g["center_x"]), F(g["center_y"]), F(g["extra_vx"]), F(g["extra_vy"])
    L = F(g["scale_length"])
    max_r = F(F(g["outer_radius"]) + F(600))
    max_cum = F(F(1) - np.exp(F(-max_r / L)))
    rv = (uv[:, 0] * max_cum).astype(F)
    r = (-L * np.log((F(1) - rv).astype(F))).astype(F)
    r = np.where(r < F(0.01), F(0.01), r).astype(F)
    ang = ((uv[:, 1] * F(2)).astype(F) * PI_F).astype(F)
    dx, dy = np.cos(ang).astype(F), np.sin(ang).astype(F)
    x = (cx + (dx * r).astype(F)).astype(F)
    y = (cy + (dy * r).astype(F)).astype(F)
    speed = (F(10) * np.sqrt((F(1758) / (r + F(54)).astype(F)).astype(F))).astype(F)
    vx = ((((-dy) * speed).astype(F) * F(0.85)).astype(F) + evx).astype(F)
    vy = ((((dx) * speed).astype(F) * F(0.85)).astype(F) + evy).astype(F)
    sc = F(g["mass_scatter"])
    mv = (F(g["mass_visible"]) * (F(1) + ((uv[:, 2] * F(2)).astype(F) - F(1)).astype(F) * sc).astype(F)).astype(F)
    # dark halo: rc * sqrt((float)(pow(1 + (R / rc)^2, u) - 1)), pow in double
    rc = F(g["radius_core_dm"])
    ratio = F(F(g["outer_radius_dm"]) / rc)
    base = 1.0 + float(ratio) * float(ratio)
    inner = (np.power(base, ud[:, 0].astype(np.float64)) - 1.0).astype(F)
    rd = (rc * np.sqrt(inner).astype(F)).astype(F)
    angd = ((ud[:, 1] * F(2)).astype(F) * PI_F).astype(F)
    xd = (cx + (rd * np.cos(angd).astype(F)).astype(F)).astype(F)
    yd = (cy + (rd * np.sin(angd).astype(F)).astype(F)).astype(F)
    ix = np.minimum(np.trunc((ud[:, 2] * F(60)).astype(F)).astype(np.int32), 59)
    iy = np.minimum(np.trunc((ud[:, 3] * F(60)).astype(F)).astype(np.int32), 59)
    vxd = (((ix - 30).astype(F) * F(0.85)).astype(F) + evx).astype(F)
    vyd = (((iy - 30).astype(F) * F(0.85)).astype(F) + evy).astype(F)
    md = (F(g["mass_dark"]) * (F(1) + ((ud[:, 4] * F(2)).astype(F) - F(1)).astype(F) * sc).astype(F)).astype(F)
    n = nv + nd
    out = {
        "x": np.concatenate([x, xd]), "y": np.concatenate([y, yd]), "vx": np.concatenate([vx, vxd]),
        "vy": np.concatenate([vy, vyd]), "mass": np.concatenate([mv, md]),
        "flags": np.concatenate([np.full(nv, FLAG_SPH, np.uint32), np.full(nd, FLAG_DARK | FLAG_UNIQUE, np.uint32)]),
        "id": (np.uint32(g["first_id"]) + np.arange(n, dtype=np.uint32)).astype(np.uint32),
        "radius": np.concatenate([r, rd]),
        "color": np.concatenate([np.full(nv, 0x64808080, np.uint32), np.full(nd, 0x00808080, np.uint32)]),
        "size": np.full(n, 0.125, F),
    }
    out["sph_mass"] = (out["mass"] / F(8.5e9)).astype(F)
    return out
