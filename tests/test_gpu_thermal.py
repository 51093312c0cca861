"""GPU parity of the temperature coupling (include/ge_thermal.h) against the
oracle: node temperatures bit-exact on the same tree; one frame's
temperatures, kinetic energies, material phases and flags bit-exact (the
walk's acceptance and distances are the reference's, the walk's order of
node visits is each particle's own pre-order); three frames within 1e-5
relative (forces differ by the approximate reciprocal square root, so later
frames' velocities and positions drift apart)."""
import numpy as np
import pytest

import oracle as O
from galaxy_step import configs, default_params
from galaxy_step import thermal as T
from galaxy_step.lib import GE_FLAG_HOT_POINT, GE_FLAG_SOLIDIFIED, Context

pytestmark = pytest.mark.gpu

F = np.float32


def _scene(n, seed):
    st, _ = configs.config0(seed=seed, n=n)
    rng = np.random.default_rng(seed + 100)
    n = len(st["x"])
    st["temp"] = rng.uniform(100.0, 2000.0, n).astype(F)
    st["label"] = rng.integers(0, 9, n).astype(np.uint32)     # 0 and 8: no material
    st["sph_mass"] = rng.uniform(0.5, 2.0, n).astype(F)
    st["flags"] = np.where(rng.random(n) < 0.3, GE_FLAG_HOT_POINT, 0).astype(np.uint32)
    st["pvx"] = (st["vx"] + rng.normal(0, 2, n)).astype(F)
    st["pvy"] = (st["vy"] + rng.normal(0, 2, n)).astype(F)
    st["ax"] = rng.normal(0, 1, n).astype(F)
    st["ay"] = rng.normal(0, 1, n).astype(F)
    return st


def _params(**kw):
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0, temperature=1,
                       compensated=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _upload(c, st):
    c.upload(st)
    T.thermal_upload(c, st)


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


def test_node_temps_bitexact():
    st = _scene(30000, 1)
    n = len(st["x"])
    bb = O.bbox(st["x"], st["y"])
    perm = O.sort_perm(O.morton_keys(st["x"], st["y"], bb))
    nodes = O.build_tree(st["x"][perm], st["y"][perm], st["mass"][perm], bb)
    want = O.node_temps(nodes, st["temp"][perm])
    with Context(n, sph=True, temperature=True) as c:
        _upload(c, st)
        c.compute_morton_keys(bb)
        c.sort_by_morton_key()
        c.build_quadtree(bb)
        got = T.node_temps(c)
        t_sorted = T.thermal_download(c)["temp"]
    np.testing.assert_array_equal(t_sorted, st["temp"][perm])  # the sort moved the temperatures
    assert len(got) == len(want)
    np.testing.assert_array_equal(got, want)


KEYS = ("x", "y", "vx", "vy", "mass", "sph_mass", "rest_dens", "stiff", "visc", "cohesion", "flags", "id")


@pytest.mark.parametrize("G", [1.0, 0.0])
def test_one_frame_bitexact(G):
    st = _scene(20000, 2)
    n = len(st["x"])
    p = _params(G=G)
    s = O.State(st, sph=True, temperature=True)
    s.step(dict(G=G, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0, temperature=1,
                materials=O.reference_materials()), nsteps=1)
    ref = _by_id({k: s[k].copy() for k in KEYS + ("temp", "ke", "prev_ke", "label")})
    with Context(n, sph=True, temperature=True) as c:
        _upload(c, st)
        c.step(p, 1)
        c.sync()
        d = c.download(list(KEYS))
        d.update(T.thermal_download(c))
    got = _by_id(d)
    for k in ("temp", "ke", "prev_ke", "label", "mass", "sph_mass", "rest_dens", "stiff", "visc", "cohesion",
              "flags", "x", "y"):
        np.testing.assert_array_equal(got[k], ref[k], k)
    moved = got["temp"] != st["temp"][np.argsort(st["id"])]
    assert moved.mean() > 0.99
    hot = (got["flags"] & GE_FLAG_HOT_POINT) != 0
    solid = (got["flags"] & GE_FLAG_SOLIDIFIED) != 0
    assert hot.any() and solid.any()


def test_three_frames_with_sph_close():
    st, prm = configs.config2(n_target=12000, n_impactor=3000)
    n = len(st["x"])
    rng = np.random.default_rng(5)
    st["temp"] = rng.uniform(250.0, 1500.0, n).astype(F)
    # no material: a material phase sets mass = 8.5e9 * sphMass, which a
    # G = 1 scene cannot take (its later frames would not compare)
    st["label"] = np.zeros(n, np.uint32)
    p = _params(sph=1, sph_h=4.0, sph_cell=4.0, dt=0.002, dt_start=0.002, theta=0.7)
    s = O.State(st, sph=True, temperature=True)
    s.step(dict(G=1.0, theta=0.7, softening=1.0, dt=0.002, dt_start=0.002, direct_below=0, temperature=1, sph=1,
                sph_h=4.0, sph_cell=4.0, materials=O.reference_materials()), nsteps=3)
    ref = _by_id({k: s[k].copy() for k in ("temp", "ke", "x", "id")})
    with Context(n, sph=True, temperature=True) as c:
        _upload(c, st)
        c.step(p, 3)
        c.sync()
        d = c.download(["id", "x"])
        d.update(T.thermal_download(c, ["temp", "ke"]))
    got = _by_id(d)
    np.testing.assert_allclose(got["temp"], ref["temp"], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(got["ke"], ref["ke"], rtol=1e-3, atol=1e-6 * float(np.abs(ref["ke"]).max()))


def test_graph_replay_matches_eager():
    st = _scene(12000, 6)
    st["label"][:] = 0
    n = len(st["x"])
    out = []
    for graph in (0, 1):
        with Context(n, sph=True, temperature=True) as c:
            _upload(c, st)
            p = _params(use_graph=graph, dt_start=0.01)
            for _ in range(3):
                c.step(p, 2)
                c.sync()
            d = c.download(["id", "x"])
            d.update(T.thermal_download(c, ["temp"]))
            out.append(_by_id(d))
    np.testing.assert_array_equal(out[0]["temp"], out[1]["temp"])
    np.testing.assert_array_equal(out[0]["x"], out[1]["x"])


def test_custom_materials_and_staged_calculation():
    st = _scene(4000, 7)
    n = len(st["x"])
    mats = O.reference_materials()[:2]            # water and rock only
    mats[1].hot_point = 900.0
    tp = T.default_thermal_params(ambient_temp=300.0, ambient_heat_rate=0.5)
    with Context(n, sph=True, temperature=True) as c:
        _upload(c, st)
        lm = [T.Material.from_buffer_copy(bytes(m)) for m in mats]
        T.set_thermal(c, tp, lm)
        p = _params(dt=0.03)
        T.temperature_calculation(c, p)
        d = c.download(["sph_mass", "flags", "mass"])
        d.update(T.thermal_download(c))
    want = O.temperature(st, mats, ambient=300.0, rate=0.5, dt=0.03)
    for k in ("temp", "ke", "prev_ke", "sph_mass", "flags", "mass"):
        np.testing.assert_array_equal(d[k], want[k], k)


def test_thermal_rejections():
    st = _scene(1000, 8)
    with Context(1000, sph=True) as c:  # no temperature fields
        c.upload(st)
        with pytest.raises(Exception):
            c.step(_params(), 1)
    with Context(1000, sph=True, temperature=True) as c:
        _upload(c, st)
        with pytest.raises(Exception):  # the heat transfer follows the reference walk
            c.step(_params(accuracy=1), 1)
    with pytest.raises(Exception):       # the materials write SPH fields
        Context(1000, temperature=True)
