"""Pins of the temperature restatement (CPU): node temperatures by exact
integer totals, the walk's heat transfer against a closed form on a
two-cluster scene, temperatureCalculation's relaxation law and its phase
branches read from physics.cpp:567-652, and the ABI of include/ge_thermal.h
(exports, struct layout, the material table against the oracle's)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import oracle as O
from galaxy_step import configs

F = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _tree(n, seed):
    st, _ = configs.config0(seed=seed, n=n)
    bb = O.bbox(st["x"], st["y"])
    perm = O.sort_perm(O.morton_keys(st["x"], st["y"], bb))
    x, y, m = st["x"][perm], st["y"][perm], st["mass"][perm]
    return x, y, m, O.build_tree(x, y, m, bb)


def test_node_temps_exact_totals():
    x, y, m, nodes = _tree(6000, 1)
    rng = np.random.default_rng(2)
    t = rng.integers(1, 1000, len(x)).astype(F)       # integer sums stay exact in float32
    g = O.node_temps(nodes, t)
    assert g[0] == t.astype(np.float64).sum()
    tot = np.concatenate([[0.0], np.cumsum(t.astype(np.float64))])
    np.testing.assert_array_equal(g, (tot[nodes["end"]] - tot[nodes["start"]]).astype(F))


def test_node_temps_follow_the_fold_order():
    # float sums: a leaf folds its particles in order, an internal node its
    # children in the order q = 0, 2, 1, 3; reversing a leaf's order or
    # taking q = 0, 1, 2, 3 must change some node's value
    x, y, m, nodes = _tree(6000, 3)
    t = np.random.default_rng(4).uniform(1, 3000, len(x)).astype(F)
    g = O.node_temps(nodes, t)
    alt_q, alt_leaf = 0, 0
    for k, nd in enumerate(nodes):
        ch = nd["child"]
        if np.all(ch == 0xFFFFFFFF):
            s = F(0)
            for j in range(int(nd["end"]) - 1, int(nd["start"]) - 1, -1):
                s = F(s + t[j])
            alt_leaf += s != g[k]
        else:
            s = F(0)
            for q in (0, 1, 2, 3):
                if ch[q] != 0xFFFFFFFF:
                    s = F(s + g[ch[q]])
            alt_q += s != g[k]
    assert alt_q > 0 and alt_leaf > 0
    np.testing.assert_allclose(g[0], t.astype(np.float64).sum(), rtol=1e-5)


def test_walk_heat_closed_form_two_clusters():
    # particle 0 alone at the origin; three particles near (1000, 1000).  The
    # root is opened, particle 0's own quadrant is a leaf holding only itself
    # (no heat), and the cluster's quadrant is accepted: one transfer
    # t0' = t0 + K * (mean - t0) / |COM - x0| * dt (softening 0)
    x = np.array([0.0, 999.6, 1000.0, 1000.4], F)
    y = np.array([0.0, 1000.1, 999.7, 1000.3], F)
    m = np.ones(4, F)
    t = np.array([300.0, 1000.0, 2000.0, 4000.0], F)
    bb = O.bbox(x, y)
    perm = O.sort_perm(O.morton_keys(x, y, bb))
    assert perm[0] == 0
    nodes = O.build_tree(x[perm], y[perm], m[perm], bb)
    g = O.node_temps(nodes, t[perm])
    K, dt = 0.045, 0.5
    out = O.walk_heat(x[perm], y[perm], nodes, g, 0.5, 0.0, K, dt, t[perm])
    com = np.array([x[1:].astype(np.float64).mean(), y[1:].astype(np.float64).mean()])
    want = 300.0 + K * (np.mean([1000.0, 2000.0, 4000.0]) - 300.0) / np.hypot(*com) * dt
    assert abs(out[0] - want) < 1e-4
    # the cluster particles see particle 0's leaf (accepted) and open their
    # own leaf: each moves toward 300 and nothing else
    assert np.all(out[1:] < t[perm][1:])


def _mat(mid):
    return next(mm for mm in O.reference_materials() if mm.id == mid)


def _state(temp, label, v=(0.0, 0.0), pv=(0.0, 0.0), flags=0, sm=1.0):
    n = len(temp)
    return {"x": np.zeros(n, F), "y": np.zeros(n, F), "vx": np.full(n, v[0], F), "vy": np.full(n, v[1], F),
            "pvx": np.full(n, pv[0], F), "pvy": np.full(n, pv[1], F), "mass": np.ones(n, F),
            "flags": np.full(n, flags, np.uint32), "temp": np.array(temp, F),
            "label": np.array(label, np.uint32), "sph_mass": np.full(n, sm, F)}


def test_temperature_relaxes_toward_ambient():
    # at rest (vel = prevVel): only the cooling term, (t' - amb) = (t - amb) (1 - K rate dt)
    out = O.temperature(_state([500.0, 100.0, 274.0], [0, 0, 0]), [], ambient=274.0, rate=1.0, dt=0.5)
    np.testing.assert_allclose(out["temp"] - 274.0, np.array([226.0, -174.0, 0.0]) * (1 - 0.05 * 0.5), rtol=1e-6)
    assert np.all(out["ke"] == 0) and np.all(out["prev_ke"] == 0)
    # kinetic heating: q = |ke - prevKe| / (2 K sphMass + 1) is added before the cooling
    out = O.temperature(_state([274.0], [0], v=(3.0, 4.0), pv=(0.0, 0.0), sm=2.0), [], ambient=274.0, dt=0.0)
    assert out["ke"][0] == F(25.0) and out["prev_ke"][0] == 0
    np.testing.assert_allclose(out["temp"][0], 274.0 + 25.0 / (2 * 0.05 * 2.0 + 1.0), rtol=1e-6)


def test_temperature_phases():
    mats = O.reference_materials()
    water, rock = _mat(1), _mat(2)
    HOT, SOLID = O.GEO_HOT_POINT, O.GEO_SOLIDIFIED
    # water (cold point 273.2): above the boiling point -> hot values, no
    # hot flag; between -> base values and the hot flag; below freezing
    # with the hot flag -> cold values and solidified
    out = O.temperature(_state([400.0, 300.0, 200.0], [1, 1, 1], flags=0), mats, ambient=400.0, dt=0.0)
    assert out["sph_mass"][0] == F(water.hot_mass_mult) and out["rest_dens"][0] == F(water.hot_rest_dens)
    assert out["flags"][0] & HOT == 0
    assert out["sph_mass"][1] == F(water.mass_mult) and out["flags"][1] & HOT
    assert out["visc"][2] == F(water.cold_visc) and out["cohesion"][2] == F(water.cold_cohesion)
    np.testing.assert_array_equal(out["mass"], F(8.5e9) * out["sph_mass"])
    out = O.temperature(_state([200.0], [1], flags=HOT), mats, ambient=200.0, dt=0.0)
    assert out["flags"][0] & SOLID and not out["flags"][0] & HOT
    # rock (cold point 0): melting sets the hot flag; cooling below the hot
    # point with it solidifies
    out = O.temperature(_state([1400.0, 1000.0], [2, 2], flags=0), mats, ambient=1400.0, dt=0.0)
    assert out["stiff"][0] == F(rock.hot_stiff) and out["flags"][0] & HOT
    assert out["stiff"][1] == F(rock.stiff) and out["flags"][1] == 0
    out = O.temperature(_state([1000.0], [2], flags=HOT), mats, ambient=1000.0, dt=0.0)
    assert out["flags"][0] & SOLID and not out["flags"][0] & HOT
    # an unknown label keeps its SPH fields
    out = O.temperature(_state([5000.0], [42]), mats, ambient=5000.0, dt=0.0)
    assert out["sph_mass"][0] == 1.0 and out["flags"][0] == 0


def test_step_with_temperature_at_ambient_is_stationary():
    st, _ = configs.config0(seed=5, n=6000)
    st["vx"][:] = 0
    st["vy"][:] = 0
    st["temp"] = np.full(len(st["x"]), 274.0, F)
    s = O.State(st, sph=True, temperature=True)
    s.step(dict(G=0.0, theta=0.7, softening=1.0, dt=0.01, direct_below=0, temperature=1,
                materials=O.reference_materials()), nsteps=2)
    np.testing.assert_array_equal(s["temp"], F(274.0))


THERMAL_HEADER = os.path.join(ROOT, "include", "ge_thermal.h")


def test_library_exports_thermal_header():
    from galaxy_step import thermal
    lib = thermal.load_thermal_library()
    text = open(THERMAL_HEADER).read()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge_[a-z_0-9]+)\s*\(", text, re.M)))
    assert set(names) == set(thermal.THERMAL_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_thermal_struct_layout_and_material_table(tmp_path):
    from galaxy_step import thermal as T
    from galaxy_step import lib as L
    src = tmp_path / "szt.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_thermal.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(ge_material), sizeof(ge_thermal_params), '
                   'sizeof(ge_thermal_fields), offsetof(ge_material, is_plastic), '
                   'offsetof(ge_config, enable_temperature));return 0;}\n')
    exe = tmp_path / "szt"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [C.sizeof(T.Material), C.sizeof(T.ThermalParams), C.sizeof(T.ThermalFields),
                                     T.Material.is_plastic.offset, L.GeConfig.enable_temperature.offset]
    assert C.sizeof(T.Material) == C.sizeof(O.Material)
    lib_m = T.default_materials()
    ref_m = O.reference_materials()
    assert [m.id for m in lib_m] == list(range(1, 8))
    for a, b in zip(lib_m, ref_m):
        assert bytes(a) == bytes(b), a.id
    p = T.default_thermal_params()
    assert (p.heat_conductivity, p.ambient_temp, p.ambient_heat_rate) == (F(0.045), 274.0, 1.0)
    assert p.base_mass == F(8.5e9)
