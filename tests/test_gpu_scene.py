"""Scene files through the device: a save file in SaveSystem's version-190
layout is loaded into a context, stepped on the GPU, written back, and
reloaded.  The records are decoded with the independent layout restatement
of tests/test_scene.py.  Device fields must equal the device state exactly;
every other field (colours, sizes, labels, temperature, keys) must come back
unchanged for the particle of the same id; the stepped positions must match
the oracle stepping the same scene."""
import numpy as np
import pytest

import oracle as O
import oracle3d as O3
from galaxy_step import Context, configs
from galaxy_step.lib3d import Context3D
from galaxy_step.scene import Scene
from test_scene import SEP, random_records, record_dtype

pytestmark = pytest.mark.gpu

DEVICE_FIELDS = {"pos", "vel", "acc", "prevVel", "mass", "id", "isPinned", "isSPH", "isBeingDrawn",
                 "isDarkMatter", "predPos", "pressF", "press", "pressTmp", "dens", "predDens", "sphMass",
                 "restDens", "stiff", "visc", "cohesion"}


def _file(path, recs, yaml):
    head = np.array([190, 1, 2, 3, len(recs)], np.uint32).tobytes()
    path.write_bytes(yaml.encode() + SEP + head + recs.tobytes() + bytes(24))


def _records_2d(st, seed):
    n = len(st["x"])
    r = random_records(n, 2, seed)         # random rendering and bookkeeping fields
    r["pos"] = np.stack([st["x"], st["y"]], 1)
    r["vel"] = np.stack([st["vx"], st["vy"]], 1)
    for f in ("acc", "prevVel", "predPos", "pressF"):
        r[f] = 0.0
    r["mass"] = st["mass"]
    r["isSPH"] = 0
    r["isBeingDrawn"] = 0
    r["isDarkMatter"] = 0
    r["isPinned"] = 0
    r["isPinned"][::97] = 1
    return r


def _by_id(recs):
    return recs[np.argsort(recs["id"])]


def test_scene_2d_round_trip_through_the_device(tmp_path):
    st, _ = configs.config0(seed=12, n=6000)
    recs = _records_2d(st, 5)
    gm = np.float32(1.0 / 6.674e-11)
    yaml = (f"Theta: 0.800000012\n---\nSoftening: 1\n---\nGravityMultiplier: {gm:.9g}\n---\n"
            f"TimeMult: {np.float32(0.01 / 0.045):.9g}\n---\nSPHEnabled: false\n---\nis3DMode: false\n")
    src = tmp_path / "Save_0.bin"
    _file(src, recs, yaml)
    with Scene.load(src) as sc:
        p = sc.params()
        p.direct_below = 0
        n = sc.info()["count"]
        with Context(n) as c:
            sc.upload(c)
            c.step(p, 3)
            c.sync()
            dev = c.download(["x", "y", "vx", "vy", "id"])
            sc.download(c)
        out = tmp_path / "Save_1.bin"
        sc.save(out)
    with Scene.load(out) as back:
        assert back.info()["count"] == n and back.info()["global_id"] == max(int(recs["id"].max()) + 1, 1)
        got = np.frombuffer(back.records(), record_dtype(2))
    # device order in the file is the device order
    np.testing.assert_array_equal(got["id"], dev["id"])
    np.testing.assert_array_equal(got["pos"][:, 0], dev["x"])
    np.testing.assert_array_equal(got["vel"][:, 1], dev["vy"])
    g, r = _by_id(got), _by_id(recs)
    for name in recs.dtype.names:
        if name not in DEVICE_FIELDS:
            np.testing.assert_array_equal(g[name], r[name], err_msg=name)
    np.testing.assert_array_equal(g["isPinned"], r["isPinned"])
    assert np.all(g["pos"][r["isPinned"] == 1] == r["pos"][r["isPinned"] == 1])
    # the same scene stepped by the oracle
    ref = O.State({"x": recs["pos"][:, 0], "y": recs["pos"][:, 1], "vx": recs["vel"][:, 0],
                   "vy": recs["vel"][:, 1], "mass": recs["mass"], "id": recs["id"],
                   "flags": recs["isPinned"].astype(np.uint32)})
    ref.step(dict(G=float(p.G), theta=float(p.theta), softening=float(p.softening), dt=float(p.dt),
                  dt_start=float(p.dt), direct_below=0), nsteps=3)
    o = np.argsort(ref["id"])
    ext = float(np.ptp(ref["x"]))
    assert np.abs(g["pos"][:, 0] - ref["x"][o]).max() < 1e-4 * ext
    assert np.abs(g["pos"][:, 1] - ref["y"][o]).max() < 1e-4 * ext


def test_scene_3d_sph_round_trip_through_the_device(tmp_path):
    st, prm = configs.config3d_sph(n_target=2400, n_impactor=600)
    n = len(st["x"])
    r = random_records(n, 3, 6)
    r["pos"] = np.stack([st["x"], st["y"], st["z"]], 1)
    r["vel"] = np.stack([st["vx"], st["vy"], st["vz"]], 1)
    for f in ("acc", "prevVel", "predPos", "pressF"):
        r[f] = 0.0
    r["mass"] = st["mass"]
    for f, k in (("sphMass", "sph_mass"), ("restDens", "rest_dens"), ("stiff", "stiff"), ("visc", "visc"),
                 ("cohesion", "cohesion")):
        r[f] = st[k]
    r["dens"] = 0.0
    r["isSPH"] = 1
    r["isBeingDrawn"] = 0
    r["isDarkMatter"] = 0
    r["isPinned"] = 0
    gm = np.float32(1.0 / 6.674e-11)
    yaml = (f"Theta: 0.699999988\n---\nSoftening: 1\n---\nGravityMultiplier: {gm:.9g}\n---\n"
            f"TimeMult: {np.float32(0.002 / 0.045):.9g}\n---\nSPHEnabled: true\n---\nSPHRadiusMult: 3\n---\n"
            f"SPHMass: 0.0299999993\n---\nis3DMode: true\n")
    src = tmp_path / "Save_3d.bin"
    _file(src, r, yaml)
    with Scene.load(src) as sc:
        p = sc.params3()
        p.direct_below = 0
        assert p.sph == 1 and p.sph_h == 3.0
        with Context3D(n, sph=True) as c:
            sc.upload(c)
            c.step(p, 2)
            c.sync()
            dev = c.download(["x", "z", "press", "pred_dens", "pfx", "id"])
            sc.download(c)
        out = tmp_path / "Save_3d_out.bin"
        sc.save(out)
    with Scene.load(out) as back:
        got = np.frombuffer(back.records(), record_dtype(3))
    np.testing.assert_array_equal(got["id"], dev["id"])
    np.testing.assert_array_equal(got["pos"][:, 2], dev["z"])
    np.testing.assert_array_equal(got["press"], dev["press"])
    np.testing.assert_array_equal(got["predDens"], dev["pred_dens"])
    np.testing.assert_array_equal(got["pressF"][:, 0], dev["pfx"])
    g, rr = _by_id(got), _by_id(r)
    for name in r.dtype.names:
        if name not in DEVICE_FIELDS:
            np.testing.assert_array_equal(g[name], rr[name], err_msg=name)
    # the oracle's 3D step of the same scene
    ref = O3.State3({"x": r["pos"][:, 0], "y": r["pos"][:, 1], "z": r["pos"][:, 2], "vx": r["vel"][:, 0],
                     "vy": r["vel"][:, 1], "vz": r["vel"][:, 2], "mass": r["mass"], "id": r["id"],
                     "flags": np.full(n, 2, np.uint32), "sph_mass": r["sphMass"], "rest_dens": r["restDens"],
                     "stiff": r["stiff"], "visc": r["visc"], "cohesion": r["cohesion"]})
    q = {k: getattr(p, k) for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass",
                                     "viscosity", "cohesion_coef", "delta", "stiff_mult", "sph_max_vel")}
    q.update(dt_start=p.dt, direct_below=0)
    ref.step(q, 2)
    o = np.argsort(ref["id"])
    ext = float(np.ptp(ref["x"]))
    assert np.abs(g["pos"][:, 0] - ref["x"][o]).max() < 1e-4 * ext


def _rgba(c):
    """raylib Color {r, g, b, a} bytes as the packed RGBA8 word (r in the low byte)."""
    return np.ascontiguousarray(c).view("<u4").reshape(-1)


def test_scene_render_fields_through_the_device(tmp_path):
    # a context with render fields takes color, pColor, sColor, size,
    # neighbors, uniqueColor and isSolid from the file, the consumers update
    # them on the device, and the file gets the device values back
    from galaxy_step import render as GR
    import render as R
    st, _ = configs.config0(seed=13, n=5000)
    recs = _records_2d(st, 8)
    recs["isDarkMatter"][::11] = 1
    yaml = "Theta: 0.800000012\n---\nSoftening: 1\n---\nis3DMode: false\n"
    src = tmp_path / "Save_r.bin"
    _file(src, recs, yaml)
    v = GR.default_visual_params(density_size=1, force_color=0, count_neighbors=1)
    with Scene.load(src) as sc:
        n = sc.info()["count"]
        with Context(n, sph=True, render=True) as c:
            sc.upload(c)
            before = GR.render_download(c)
            fl = c.download(["flags", "x", "y", "ax", "ay"])
            GR.visuals(c, v)
            after = GR.render_download(c)
            sc.download(c)
        out = tmp_path / "Save_r_out.bin"
        sc.save(out)
    # upload: the record fields, colours as Color {r, g, b, a} bytes
    np.testing.assert_array_equal(before["color"], _rgba(recs["color"]))
    np.testing.assert_array_equal(before["pcolor"], _rgba(recs["pColor"]))
    np.testing.assert_array_equal(before["size"], recs["size"])
    np.testing.assert_array_equal(before["neighbors"], recs["neighbors"])
    uniq = (fl["flags"] & R.FLAG_UNIQUE) != 0
    np.testing.assert_array_equal(uniq, recs["uniqueColor"] == 1)
    np.testing.assert_array_equal((fl["flags"] & R.FLAG_SOLID) != 0, recs["isSolid"] == 1)
    # the consumers on the device, against the oracle
    off, _ = O.sph_candidates(fl["x"], fl["y"], fl["x"], fl["y"], 3.0)
    nb = R.neighbor_counts(off, fl["flags"], before["neighbors"])
    col, pc, sc_, size = R.visuals(fl["flags"], fl["ax"], fl["ay"], None, before["color"], before["pcolor"],
                                   before["scolor"], before["size"], nb, v.as_dict())
    np.testing.assert_array_equal(after["color"], col)
    np.testing.assert_array_equal(after["size"], size)
    # and back into the file (same order: no step ran)
    with Scene.load(out) as back:
        got = np.frombuffer(back.records(), record_dtype(2))
    np.testing.assert_array_equal(_rgba(got["color"]), after["color"])
    np.testing.assert_array_equal(_rgba(got["sColor"]), after["scolor"])
    np.testing.assert_array_equal(got["size"], after["size"])
    np.testing.assert_array_equal(got["neighbors"], after["neighbors"])
    np.testing.assert_array_equal(got["uniqueColor"], recs["uniqueColor"])
    np.testing.assert_array_equal(got["isSolid"], recs["isSolid"])


def test_scene_temperature_through_the_device(tmp_path):
    # a temperature scene: temp, sphLabel, ke, prevKe, isHotPoint and
    # hasSolidified go to a context with temperature fields, the frame runs
    # the coupling with the file's AmbientTemperature / HeatConductivity, and
    # the records get the device values back; one frame matches the oracle
    from galaxy_step import thermal as T
    import oracle as O
    st, _ = configs.config0(seed=14, n=6000)
    recs = _records_2d(st, 9)
    rng = np.random.default_rng(10)
    n = len(recs)
    recs["temp"] = rng.uniform(200.0, 1500.0, n).astype(np.float32)
    recs["sphLabel"] = rng.integers(0, 8, n).astype(np.uint32)
    recs["sphMass"] = rng.uniform(0.5, 1.5, n).astype(np.float32)
    recs["isPinned"] = 0
    gm = np.float32(1.0 / 6.674e-11)
    yaml = (f"Theta: 0.699999988\n---\nSoftening: 1\n---\nGravityMultiplier: {gm:.9g}\n---\n"
            f"TimeMult: {np.float32(0.01 / 0.045):.9g}\n---\nTemperatureSimulation: true\n---\n"
            f"AmbientTemperature: 300\n---\nHeatConductivityMultiplier: 0.0900000036\n---\nis3DMode: false\n")
    src = tmp_path / "Save_t.bin"
    _file(src, recs, yaml)
    with Scene.load(src) as sc:
        p = sc.params()
        tp = sc.thermal_params()
        assert p.temperature == 1 and tp.ambient_temp == 300.0 and tp.heat_conductivity == np.float32(0.09)
        p.direct_below = 0
        with Context(n, sph=True, temperature=True) as c:
            sc.upload(c)
            T.set_thermal(c, tp)
            c.step(p, 1)
            c.sync()
            dev = c.download(["id", "flags"])
            dev.update(T.thermal_download(c))
            sc.download(c)
        out = tmp_path / "Save_t_out.bin"
        sc.save(out)
    with Scene.load(out) as back:
        got = np.frombuffer(back.records(), record_dtype(2))
    np.testing.assert_array_equal(got["id"], dev["id"])
    np.testing.assert_array_equal(got["temp"], dev["temp"])
    np.testing.assert_array_equal(got["ke"], dev["ke"])
    np.testing.assert_array_equal(got["sphLabel"], dev["label"])
    np.testing.assert_array_equal(got["isHotPoint"], (dev["flags"] & 64) != 0)
    ref = O.State({"x": recs["pos"][:, 0], "y": recs["pos"][:, 1], "vx": recs["vel"][:, 0],
                   "vy": recs["vel"][:, 1], "mass": recs["mass"], "id": recs["id"],
                   "sph_mass": recs["sphMass"], "temp": recs["temp"], "label": recs["sphLabel"],
                   "flags": (recs["isHotPoint"].astype(np.uint32) * 64 + recs["hasSolidified"].astype(np.uint32) * 128),
                   "pvx": recs["prevVel"][:, 0], "pvy": recs["prevVel"][:, 1]}, sph=True, temperature=True)
    ref.step(dict(G=float(p.G), theta=float(p.theta), softening=float(p.softening), dt=float(p.dt),
                  dt_start=float(p.dt), direct_below=0, temperature=1, ambient_temp=300.0,
                  heat_conductivity=float(np.float32(0.09)), materials=O.reference_materials()), nsteps=1)
    o = np.argsort(ref["id"])
    g = _by_id(got)
    np.testing.assert_array_equal(g["temp"], ref["temp"][o])


def test_scene_constraints_through_the_device(tmp_path):
    # the 2D tail starts with the constraints (count + raw 36-byte records);
    # a context that holds constraints takes them, the frame solves them with
    # the file's constraint variables, and the tail gets the device's
    # constraints back with the rest of the tail (walls, shapes, lights) intact
    from galaxy_step import constraints as GC
    from test_gpu_constraints import _lattice
    st, cons = _lattice(20, 4)
    n = len(st["x"])
    recs = random_records(n, 2, 11)
    recs["pos"] = np.stack([st["x"], st["y"]], 1)
    recs["vel"] = np.stack([st["vx"], st["vy"]], 1)
    for f in ("acc", "prevVel", "predPos", "pressF"):
        recs[f] = 0.0
    recs["mass"] = st["mass"]
    recs["id"] = st["id"]
    for f in ("isPinned", "isSPH", "isBeingDrawn", "isDarkMatter", "isHotPoint"):
        recs[f] = 0
    rest_of_tail = bytes(range(40))
    tail = np.array([len(cons)], np.uint32).tobytes() + cons.tobytes() + rest_of_tail
    yaml = ("Theta: 0.699999988\n---\nSoftening: 1\n---\nGravityMultiplier: 0\n---\nParticleConstraints: true\n---\n"
            "ConstraintsStiffMultiplier: 0.5\n---\nis3DMode: false\n")
    src = tmp_path / "Save_c.bin"
    head = np.array([190, 1, 2, 3, n], np.uint32).tobytes()
    src.write_bytes(yaml.encode() + SEP + head + recs.tobytes() + tail)
    with Scene.load(src) as sc:
        cp, on = sc.constraint_params()
        assert on and cp.stiffness_mult == 0.5
        p = sc.params()
        p.constraints = 1
        p.direct_below = 0
        with Context(n, max_constraints=len(cons) + 8) as c:
            sc.upload(c)
            np.testing.assert_array_equal(GC.download(c), cons)
            GC.set_params(c, cp)
            c.step(p, 3)
            c.sync()
            dev = GC.download(c)
            sc.download(c)
        out = tmp_path / "Save_c_out.bin"
        sc.save(out)
    raw = out.read_bytes()
    body = raw[raw.index(SEP) + len(SEP):]
    rec_end = 20 + n * record_dtype(2).itemsize
    m = int(np.frombuffer(body[rec_end:rec_end + 4], np.uint32)[0])
    assert m == len(dev)
    got = np.frombuffer(body[rec_end + 4:rec_end + 4 + 36 * m], GC.CONSTRAINT_DTYPE)
    np.testing.assert_array_equal(got, dev)
    assert body[rec_end + 4 + 36 * m:] == rest_of_tail
    assert not np.array_equal(dev["displacement"], cons["displacement"])  # the solver ran
