"""Pins of the constraint restatement (CPU): a single spring's closed form
over the 15 substeps, momentum and centre-of-mass conservation, removal in
order, the triangle rule on the fifth frame, the breaking and plastic rules
read from physics.cpp:895-923, the periodic wrap; and the ABI of
include/ge_constraints.h."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import oracle as O

F = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pair_state(x1, x2, m1=1.0, m2=1.0, flags=(0, 0), labels=None):
    st = {"x": np.array([x1, x2], F), "y": np.zeros(2, F), "vx": np.zeros(2, F), "vy": np.zeros(2, F),
          "ax": np.zeros(2, F), "ay": np.zeros(2, F), "mass": np.array([m1, m2], F),
          "flags": np.array(flags, np.uint32), "id": np.array([7, 9], np.uint32)}
    if labels is not None:
        st["label"] = np.array(labels, np.uint32)
    return st


def _cons(rows):
    c = np.zeros(len(rows), O.CONSTRAINT_DTYPE)
    for k, r in enumerate(rows):
        for f, v in r.items():
            c[f][k] = v
    return c


def test_single_spring_closed_form():
    L0, d0, k, m1, m2 = 2.0, 0.3, 6.0, 1.0, 3.0
    st = _pair_state(0.0, L0 + d0, m1, m2)
    cons = _cons([dict(id1=7, id2=9, rest_length=L0, original_length=L0, stiffness=k, resistance=0.6,
                       plasticity_point=0.6)])
    s, out, col, fc = O.constraints(st, cons, [0])
    cf = k * 0.013333
    ds = d0 * (1 - cf) ** np.arange(15)           # displacement seen by each substep
    sep = float(s["x"][1] - s["x"][0])
    np.testing.assert_allclose(sep - L0, d0 * (1 - cf) ** 15, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(out["displacement"][0], ds[-1], rtol=1e-4)
    # the spring force uses pi.mass: acc_i = k * sum(d) (+x), acc_j = -k * sum(d) * m1 / m2
    np.testing.assert_allclose(s["ax"][0], k * ds.sum(), rtol=1e-4)
    np.testing.assert_allclose(s["ax"][1], -k * ds.sum() * m1 / m2, rtol=1e-4)
    # momentum of the forces and the mass centre are conserved
    assert abs(m1 * s["ax"][0] + m2 * s["ax"][1]) < 1e-4 * abs(m1 * s["ax"][0])
    np.testing.assert_allclose(m1 * s["x"][0] + m2 * s["x"][1], m2 * (L0 + d0), rtol=1e-6)
    assert fc == 1


def test_removal_keeps_order_and_counts_frames():
    st = _pair_state(0.0, 2.0)
    cons = _cons([dict(id1=7, id2=9, rest_length=2.0, stiffness=1.0),
                  dict(id1=7, id2=99, rest_length=2.0, stiffness=1.0),           # missing particle
                  dict(id1=9, id2=7, rest_length=2.0, stiffness=2.0, is_broken=1),
                  dict(id1=9, id2=7, rest_length=2.0, stiffness=3.0)])
    _, out, col, fc = O.constraints(st, cons, [0, 1, 2, 3], forces=0)
    assert list(out["stiffness"]) == [1.0, 3.0] and list(col) == [0, 3] and fc == 1
    _, out, _, fc = O.constraints(st, cons[1:3], [0, 1], forces=0, frame_count=4)
    assert len(out) == 0 and fc == 4               # nothing left: the counter stays


def test_triangle_rule_on_the_fifth_frame():
    # a triangle 0-1-2 plus a dangling edge 2-3
    st = {"x": np.array([0, 1, 0.5, 3], F), "y": np.array([0, 0, 0.8, 0.8], F), "vx": np.zeros(4, F),
          "vy": np.zeros(4, F), "ax": np.zeros(4, F), "ay": np.zeros(4, F), "mass": np.ones(4, F),
          "flags": np.zeros(4, np.uint32), "id": np.arange(4, dtype=np.uint32)}
    rows = [dict(id1=a, id2=b, rest_length=1.0, stiffness=1.0) for a, b in ((0, 1), (1, 2), (0, 2), (2, 3))]
    _, out, _, fc = O.constraints(st, _cons(rows), [0, 1, 2, 0], forces=0, frame_count=3)
    assert fc == 4 and not out["is_broken"].any()
    _, out, _, fc = O.constraints(st, _cons(rows), [0, 1, 2, 0], forces=0, frame_count=4)
    assert fc == 5 and list(out["is_broken"]) == [0, 0, 0, 1]


def test_breaking_and_plastic_rules():
    mats = O.reference_materials()
    water, iron = 1, 3
    # non-plastic (water): |d| >= resistance * resistance_mult * rest breaks, then no forces
    st = _pair_state(0.0, 2.2, labels=(water, water))
    cons = _cons([dict(id1=7, id2=9, rest_length=2.0, original_length=2.0, stiffness=60.0, resistance=0.06)])
    s, out, _, _ = O.constraints(st, cons, [0], materials=mats)
    assert out["is_broken"][0] and s["ax"][0] == 0 and s["x"][1] == F(2.2)
    # unbreakable: the same stretch holds
    s, out, _, _ = O.constraints(st, cons, [0], materials=mats, unbreakable=1)
    assert not out["is_broken"][0] and s["ax"][0] > 0
    # plastic (iron): past plasticityPoint * originalLength the rest length yields
    st = _pair_state(0.0, 2.3, labels=(iron, iron))
    cons = _cons([dict(id1=7, id2=9, rest_length=2.0, original_length=2.0, stiffness=1.0, resistance=0.24,
                       plasticity_point=0.108)])
    _, out, _, _ = O.constraints(st, cons, [0], materials=mats, forces=0)
    assert out["is_plastic"][0] and out["rest_length"][0] > 2.0
    # a hot point breaks any material constraint
    st = _pair_state(0.0, 2.0, flags=(O.GEO_HOT_POINT, 0), labels=(iron, iron))
    _, out, _, _ = O.constraints(st, cons, [0], materials=mats)
    assert out["is_broken"][0]
    # without materials nothing breaks (idToMaterial yields null)
    st = _pair_state(0.0, 9.0, labels=(0, 0))
    _, out, _, _ = O.constraints(st, cons, [0], materials=mats)
    assert not out["is_broken"][0]


def test_periodic_wrap_takes_the_short_way():
    W = 100.0
    st = _pair_state(1.0, 99.0)
    cons = _cons([dict(id1=7, id2=9, rest_length=1.0, original_length=1.0, stiffness=10.0)])
    _, out, _, _ = O.constraints(st, cons, [0], forces=0, periodic=1, domain_w=W, domain_h=W)
    assert abs(out["displacement"][0] - 1.0) < 1e-5   # |delta| = 2 through the boundary


HEADER = os.path.join(ROOT, "include", "ge_constraints.h")


def test_library_exports_and_layout(tmp_path):
    from galaxy_step import constraints as GC
    from galaxy_step import lib as L
    lib = GC.load_constraint_library()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge_[a-z_0-9]+)\s*\(", open(HEADER).read(),
                                  re.M)))
    assert set(names) == set(GC.CONSTRAINT_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
    src = tmp_path / "szc.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_constraints.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(ge_constraint), '
                   'sizeof(ge_constraint_params), offsetof(ge_constraint, is_plastic), '
                   'offsetof(ge_config, max_constraints), offsetof(ge_step_params, constraints));return 0;}\n')
    exe = tmp_path / "szc"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [36, C.sizeof(GC.ConstraintParams), GC.Constraint.is_plastic.offset,
                                     L.GeConfig.max_constraints.offset, L.StepParams.constraints.offset]
    assert GC.CONSTRAINT_DTYPE.itemsize == C.sizeof(GC.Constraint) == O.CONSTRAINT_DTYPE.itemsize
    p = GC.default_constraint_params()
    assert (p.stiffness_mult, p.resistance_mult, p.unbreakable, p.forces) == (1.0, 1.0, 0, 1)
    assert p.damping == F(0.001) and p.correction_ratio == F(0.013333)


# ---- createConstraints' automatic branch (physics.cpp:654-766) -------------

SOL = O.GEO_SOLIDIFIED


def _cstate(x, y, ids, flags, labels=None):
    n = len(x)
    st = {"x": np.asarray(x, F), "y": np.asarray(y, F), "mass": np.ones(n, F), "id": np.asarray(ids, np.uint32),
          "flags": np.asarray(flags, np.uint32)}
    if labels is not None:
        st["label"] = np.asarray(labels, np.uint32)
    return st


def test_create_constraints_hand_placed():
    """A (id 5) and B (id 7) 2 apart; C (id 3) 1.5 beyond B, 3.5 from A
    (3.5^2 = 12.25 > 12); D (id 1) 3.3 from A (inside sqrt 12) but two cells of
    3 away, so never a candidate.  Only A and C are solidified."""
    st = _cstate([0.5, 2.5, 4.0, -2.95], [0.5, 0.5, 0.5, 0.5], [5, 7, 3, 1], [SOL, 0, SOL, 0])
    flags, cons = O.create_constraints(st)
    np.testing.assert_array_equal(flags, [0, 0, 0, 0])  # hasSolidified consumed
    assert [(int(c["id1"]), int(c["id2"])) for c in cons] == [(5, 7), (3, 7)]
    np.testing.assert_array_equal(cons["rest_length"], F([2.0, 1.5]))
    np.testing.assert_array_equal(cons["original_length"], cons["rest_length"])
    # no materials: stiffness 60, resistance and plastic point 0.6
    np.testing.assert_array_equal(cons["stiffness"], F([60, 60]))
    np.testing.assert_array_equal(cons["resistance"], F([0.6, 0.6]))
    np.testing.assert_array_equal(cons["plasticity_point"], F([0.6, 0.6]))
    assert not cons["is_broken"].any() and not cons["is_plastic"].any() and not cons["displacement"].any()
    # a lower id on the other side: only the particle with the lower id creates the pair
    st2 = _cstate([0.5, 2.5], [0.5, 0.5], [9, 4], [SOL, 0])
    _, cons2 = O.create_constraints(st2)
    assert len(cons2) == 0


def test_create_constraints_materials():
    """Rock (2, cold point 0) next to water (1, cold point 273.2) is skipped
    from the rock's side; rock-rock takes rock's stiffness (60) and the mean
    resistance (0.211); iron-rock takes iron's stiffness and the means; water
    creating toward rock is allowed (the rule looks at i's cold point)."""
    mats = O.reference_materials()
    by_id = {m.id: m for m in mats}
    x = [0.5, 1.5, 2.5, 3.5]
    st = _cstate(x, [0.5] * 4, [1, 2, 3, 4], [SOL, SOL, SOL, 0], labels=[2, 1, 3, 2])
    _, cons = O.create_constraints(st, mats)
    got = [(int(c["id1"]), int(c["id2"])) for c in cons]
    # i = id1 (rock): j = id2 water skipped, id3 iron ok, id4 rock ok
    # i = id2 (water): j = id3 iron, id4 rock (water's cold point is not 0)
    # i = id3 (iron): j = id4 rock
    assert got == [(1, 3), (1, 4), (2, 3), (2, 4), (3, 4)]
    k = {p: c for p, c in zip(got, cons)}
    rock, water, iron = by_id[2], by_id[1], by_id[3]
    assert k[(1, 4)]["stiffness"] == F(rock.constraint_stiffness)
    assert k[(1, 4)]["resistance"] == F((F(rock.constraint_resistance) + F(rock.constraint_resistance)) * F(0.5))
    assert k[(1, 3)]["resistance"] == F((F(rock.constraint_resistance) + F(iron.constraint_resistance)) * F(0.5))
    assert k[(1, 3)]["plasticity_point"] == F(
        (F(rock.constraint_plastic_point) + F(iron.constraint_plastic_point)) * F(0.5))
    assert k[(2, 3)]["stiffness"] == F(water.constraint_stiffness)
    assert k[(3, 4)]["stiffness"] == F(iron.constraint_stiffness)


def _brute_create(st, mats, cell=3.0):
    """O(N^2) restatement written independently: every ordered pair that the
    rules admit, ordered by (i, cell dx, cell dy, j)."""
    x, y, ids, fl = st["x"], st["y"], st["id"], st["flags"]
    lab = st.get("label")
    mat = {m.id: m for m in mats}
    cx = np.floor(x / F(cell)).astype(np.int64)
    cy = np.floor(y / F(cell)).astype(np.int64)
    rows = []
    for i in np.nonzero(fl & SOL)[0]:
        dx, dy = x - x[i], y - y[i]
        d2 = dx * dx + dy * dy
        ok = (np.abs(cx - cx[i]) <= 1) & (np.abs(cy - cy[i]) <= 1) & (np.arange(len(x)) != i) & (d2 <= F(12))
        ok &= ids[i] < ids
        mi = mat.get(int(lab[i])) if lab is not None else None
        for j in np.nonzero(ok)[0]:
            mj = mat.get(int(lab[j])) if lab is not None else None
            if mi is not None and mj is not None and mi.cold_point == 0.0 and mj.cold_point != 0.0:
                continue
            both = mi is not None and mj is not None
            rows.append(((int(i), int(cx[j] - cx[i]), int(cy[j] - cy[i]), int(j)),
                         (int(ids[i]), int(ids[j]), F(np.sqrt(d2[j])),
                          F(mi.constraint_stiffness) if both else F(60),
                          (F(mi.constraint_resistance) + F(mj.constraint_resistance)) * F(0.5) if both else F(0.6),
                          (F(mi.constraint_plastic_point) + F(mj.constraint_plastic_point)) * F(0.5) if both else F(0.6))))
    rows.sort(key=lambda r: r[0])
    return [r[1] for r in rows]


def test_create_constraints_lattice_against_brute_force():
    rng = np.random.default_rng(3)
    side = 40
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel() * 1.7 + (gy.ravel() % 2) * 0.85 + rng.normal(0, 0.15, side * side) - 20).astype(F)
    y = (gy.ravel() * 1.47 + rng.normal(0, 0.15, side * side) - 15).astype(F)
    n = len(x)
    ids = rng.permutation(n).astype(np.uint32) + 100
    flags = np.where(rng.random(n) < 0.3, SOL, 0).astype(np.uint32) | np.where(rng.random(n) < 0.2, 2, 0).astype(np.uint32)
    labels = rng.integers(0, 8, n).astype(np.uint32)
    st = _cstate(x, y, ids, flags, labels=labels)
    mats = O.reference_materials()
    out_flags, cons = O.create_constraints(st, mats)
    want = _brute_create(st, mats)
    assert len(want) > 500
    got = [(int(c["id1"]), int(c["id2"]), c["rest_length"], c["stiffness"], c["resistance"], c["plasticity_point"])
           for c in cons]
    assert got == want
    np.testing.assert_array_equal(out_flags, flags & ~np.uint32(SOL))
