"""Pins of the merger restatement (CPU): the merge formulas of
physics.cpp:986-1021 on two particles, the serial greedy order on a chain,
dark matter and distance filters, the swap-with-back deletion order, and the
ABI of include/ge_merger.h."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import oracle as O

F = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _st(x, y, m, v=None, flags=None):
    n = len(x)
    v = v if v is not None else [(0.0, 0.0)] * n
    return {"x": np.array(x, F), "y": np.array(y, F), "vx": np.array([a for a, _ in v], F),
            "vy": np.array([b for _, b in v], F), "mass": np.array(m, F),
            "flags": np.array(flags if flags is not None else [0] * n, np.uint32),
            "id": np.arange(10, 10 + n, dtype=np.uint32)}


def test_two_particles_merge_formulas():
    # totalRadius = size * 16 * 0.6 = 0.96 for size 0.1: particles 1.5 apart overlap (<= 1.92)
    st = _st([0.0, 1.5], [0.0, 0.0], [3.0, 1.0], v=[(1.0, 0.0), (-2.0, 4.0)])
    s, size, ps, nb = O.merge(st, [0.1, 0.1], [0.3, 0.4], [0, 0])
    assert len(s["x"]) == 1 and s["id"][0] == 10          # the heavier (first) stays
    assert s["mass"][0] == 4.0
    np.testing.assert_allclose([s["vx"][0], s["vy"][0]], [(3 * 1 - 2) / 4, 4 / 4], rtol=1e-7)
    np.testing.assert_allclose(ps[0], 0.4 + (np.sqrt(0.3 ** 2 + 0.4 ** 2) - 0.4) * 0.25, rtol=1e-6)
    assert nb[0] == 1                                     # UpdateNeighbors' count (within 10)
    # too far apart: nothing merges; dark matter never merges
    s, *_ = O.merge(_st([0.0, 2.5], [0.0, 0.0], [3.0, 1.0]), [0.1, 0.1], [0.1, 0.1], [0, 0])
    assert len(s["x"]) == 2
    s, _, _, nb = O.merge(_st([0.0, 0.5], [0.0, 0.0], [3.0, 1.0], flags=[8, 0]), [0.1, 0.1], [0.1, 0.1], [7, 0])
    assert len(s["x"]) == 2 and nb[0] == 7 and nb[1] == 0  # a dark particle keeps its count


def test_serial_order_on_a_chain():
    # A - B - C in a row, each pair 1.5 apart (A-C 3.0 apart: no overlap).
    # The loop starts at the last index: C meets B first; B (mass 2) absorbs C
    # (mass 1), then A (mass 5) absorbs B: one particle with mass 8 is left.
    st = _st([0.0, 1.5, 3.0], [0.0, 0.0, 0.0], [5.0, 2.0, 1.0])
    s, *_ = O.merge(st, [0.1] * 3, [0.1] * 3, [0] * 3)
    assert len(s["x"]) == 1 and s["mass"][0] == 8.0 and s["id"][0] == 10
    # C heaviest: C absorbs B, and A (mass 1) does not overlap C: two left
    st = _st([0.0, 1.5, 3.0], [0.0, 0.0, 0.0], [1.0, 2.0, 5.0])
    s, *_ = O.merge(st, [0.1] * 3, [0.1] * 3, [0] * 3)
    assert sorted(s["mass"].tolist()) == [1.0, 7.0]


def test_swap_with_back_order_and_conservation():
    # particles 1 and 3 merge into 0 and 2; deletion swaps with the back
    x = [0.0, 0.5, 100.0, 100.5, 200.0]
    st = _st(x, [0.0] * 5, [4.0, 1.0, 4.0, 1.0, 1.0], v=[(1.0, 1.0), (2.0, -1.0), (0.0, 3.0), (5.0, 0.0), (0.0, 0.0)])
    s, *_ = O.merge(st, [0.1] * 5, [0.1] * 5, [0] * 5)
    assert list(s["id"]) == [10, 14, 12]                  # 13 popped (last), then 11 swapped with 14
    p0 = sum(m * v for m, v in zip(st["mass"], st["vx"]))
    np.testing.assert_allclose((s["mass"] * s["vx"]).sum(), p0, rtol=1e-6)
    assert s["mass"].sum() == st["mass"].sum()


HEADER = os.path.join(ROOT, "include", "ge_merger.h")


def test_library_exports_and_layout(tmp_path):
    from galaxy_step import merger as GM
    lib = GM.load_merger_library()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge_[a-z_0-9]+)\s*\(", open(HEADER).read(),
                                  re.M)))
    assert set(names) == set(GM.MERGER_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
    src = tmp_path / "szm.c"
    src.write_text('#include <stdio.h>\n#include "ge_merger.h"\n'
                   'int main(void){printf("%zu\\n", sizeof(ge_merger_params));return 0;}\n')
    exe = tmp_path / "szm"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    assert int(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout) == \
        C.sizeof(GM.MergerParams)
    m = GM.default_merger_params()
    assert (m.density_radius, m.cell_size, m.texture_half_size, m.size_multiplier) == (10.0, 3.0, 16.0, F(0.6))
