"""Batched form: several independent systems stepped together in one context.

Each system keeps its own bounding box, Morton order and tree, and no force
crosses systems, so a system's result in the mix must be bit-identical to the
same system stepped alone (same code path, single-system context).  Against
the oracle the batch is held to the single-system tolerances.
"""
import numpy as np
import pytest

import oracle as O
from galaxy_step import Context, configs, default_params

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "vx", "vy", "ax", "ay", "id"]


def _systems(sizes, seed=3):
    out = []
    for k, n in enumerate(sizes):
        st, _ = configs.config3_system(k, n, 50.0 + 37.0 * k, seed=seed)
        out.append(st)
    return out


def _concat(systems):
    keys = ["x", "y", "vx", "vy", "mass", "flags"]
    st = {k: np.concatenate([s[k] for s in systems]) for k in keys}
    # ids unique across the batch
    st["id"] = np.arange(sum(len(s["x"]) for s in systems), dtype=np.uint32)
    return st


def _params(**kw):
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _split(d, off):
    return [{k: v[off[s]:off[s + 1]] for k, v in d.items()} for s in range(len(off) - 1)]


@pytest.mark.parametrize("accuracy", [0, 1, 2])
def test_batch_bitidentical_to_alone(accuracy):
    sizes = [7000, 1, 2, 130, 64, 65, 20000, 4097, 3]
    systems = _systems(sizes)
    batch = _concat(systems)
    p = _params(accuracy=accuracy)
    steps = 4
    with Context(len(batch["x"]), max_systems=len(sizes)) as c:
        c.upload(batch, counts=sizes)
        c.step(p, steps)
        c.sync()
        off = c.system_offsets()
        got = _split(c.download(FIELDS), off)
    np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(sizes)]))
    base = 0
    for s, st in enumerate(systems):
        n = sizes[s]
        st = dict(st)
        st["id"] = np.arange(base, base + n, dtype=np.uint32)
        base += n
        with Context(n) as a:
            a.upload(st)
            a.step(p, steps)
            a.sync()
            alone = a.download(FIELDS)
        for f in FIELDS:
            np.testing.assert_array_equal(got[s][f], alone[f], err_msg=f"system {s} field {f}")


def test_batch_matches_oracle():
    sizes = [5000, 9000, 6000]
    systems = _systems(sizes, seed=5)
    batch = _concat(systems)
    p = _params(theta=0.7)
    with Context(len(batch["x"]), max_systems=3) as c:
        c.upload(batch, counts=sizes)
        c.step(p, 3)
        c.sync()
        off = c.system_offsets()
        got = _split(c.download(FIELDS), off)
    for s, st in enumerate(systems):
        ref = O.State(st)
        ref.step(dict(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0), nsteps=3)
        o_ref = np.argsort(ref["id"])
        o_got = np.argsort(got[s]["id"] - off[s])
        ext = float(np.ptp(ref["x"]))
        assert np.abs(got[s]["x"][o_got] - ref["x"][o_ref]).max() < 1e-4 * ext
        assert np.abs(got[s]["vy"][o_got] - ref["vy"][o_ref]).max() < 1e-3 * float(np.abs(ref["vy"]).max())


def test_batch_many_small_systems():
    rng = np.random.default_rng(9)
    sizes = list(rng.integers(1, 300, 200))
    systems = _systems(sizes, seed=7)
    batch = _concat(systems)
    p = _params()
    with Context(len(batch["x"]), max_systems=len(sizes)) as c:
        c.upload(batch, counts=sizes)
        c.step(p, 2)
        c.sync()
        d = c.download(["x", "vx", "id"])
    assert np.all(np.isfinite(d["x"])) and np.all(np.isfinite(d["vx"]))
    assert sorted(d["id"].tolist()) == list(range(len(batch["x"])))


def test_batch_rejects_bad_layouts():
    with Context(100, max_systems=2) as c:
        with pytest.raises(Exception):
            c.upload({"x": np.zeros(10), "y": np.zeros(10), "mass": np.ones(10)}, counts=[10, 0])
        with pytest.raises(Exception):
            c.upload({"x": np.zeros(9), "y": np.zeros(9), "mass": np.ones(9)}, counts=[3, 3, 3])
