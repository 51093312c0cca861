"""GPU parity of the merger (include/ge_merger.h) against the oracle's serial
UpdateNeighbors + mergerSolver: survivors, their order (swap-with-back),
masses, velocities, previousSize and neighbour counts bit-exact on crowded
scenes with dark matter; repeated merges between gravity frames; the cell
size rule."""
import numpy as np
import pytest

import oracle as O
from galaxy_step import default_params
from galaxy_step import merger as GM
from galaxy_step import render as GR
from galaxy_step.lib import GE_FLAG_DARK, Context

pytestmark = pytest.mark.gpu

F = np.float32


def _crowd(n, seed, box=60.0):
    rng = np.random.default_rng(seed)
    st = {"x": rng.uniform(-box, box, n).astype(F), "y": rng.uniform(-box, box, n).astype(F),
          "vx": rng.normal(0, 1, n).astype(F), "vy": rng.normal(0, 1, n).astype(F),
          "mass": rng.choice([1.0, 2.0, 3.0, 5.0], n).astype(F),
          "flags": np.where(rng.random(n) < 0.1, GE_FLAG_DARK, 0).astype(np.uint32),
          "id": rng.permutation(np.arange(7, 7 + 2 * n, 2)).astype(np.uint32)}
    rf = {"size": rng.uniform(0.05, 0.3, n).astype(F), "previous_size": rng.uniform(0.05, 0.3, n).astype(F),
          "neighbors": rng.integers(0, 50, n).astype(np.int32)}
    return st, rf


KEYS = ("x", "y", "vx", "vy", "mass", "flags", "id")


def _merge_both(c, m):
    d = c.download(list(KEYS))
    r = GR.render_download(c, ["size", "previous_size", "neighbors"])
    want, size, ps, nb = O.merge(d, r["size"], r["previous_size"], r["neighbors"], **m.as_dict())
    removed = GM.merge(c, m)
    got = c.download(list(KEYS))
    gr = GR.render_download(c, ["size", "previous_size", "neighbors"])
    assert removed == len(d["x"]) - len(want["x"])
    for k in KEYS:
        np.testing.assert_array_equal(got[k], want[k], k)
    np.testing.assert_array_equal(gr["size"], size)
    np.testing.assert_array_equal(gr["previous_size"], ps)
    np.testing.assert_array_equal(gr["neighbors"], nb)
    return removed


@pytest.mark.parametrize("seed,cell", [(1, 3.0), (2, 5.5)])
def test_merge_matches_serial_oracle(seed, cell):
    st, rf = _crowd(6000, seed)
    n = len(st["x"])
    m = GM.default_merger_params(cell_size=cell)
    with Context(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        removed = _merge_both(c, m)
    assert removed > 100


def test_merges_between_gravity_frames():
    st, rf = _crowd(8000, 3, box=90.0)
    n = len(st["x"])
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.02, dt_start=0.02, direct_below=0)
    total = 0
    with Context(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        for _ in range(3):
            m = GM.default_merger_params()
            m.cell_size = GM.cell_size(c, m)
            total += _merge_both(c, m)
            c.step(p, 2)
            c.sync()
    assert total > 0


def test_cell_size_rule_and_rejections():
    st, rf = _crowd(2000, 4)
    n = len(st["x"])
    with Context(n, render=True) as c:
        c.upload(st)
        GR.render_upload(c, rf)
        got = GM.cell_size(c)
    live = (st["flags"] & GE_FLAG_DARK) == 0
    d = (rf["size"][live] * F(16.0) * F(0.6) * F(2.0)).astype(F).max()
    assert got == float(np.clip(d, 2.0, 80.0))
    with Context(100) as c:  # no render fields
        c.upload({k: v[:100] for k, v in st.items()})
        with pytest.raises(Exception):
            GM.merge(c)
