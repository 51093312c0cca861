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
