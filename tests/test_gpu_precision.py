"""GPU tests of the precision form (ge_config.precise_positions): positions
kept as a home cell of side 1/16 plus a float offset from its centre.

The form has no counterpart in the reference; what it must show is that it
wins back the accuracy that float positions lose far from the origin.  The
measure is a trajectory: a small cluster integrated for a few hundred
frames, once near the origin and once displaced to (6000.3, -4000.7), where
a float's spacing is 4.9e-4 (the cluster's radius is 1).  The truth is the
same kick-drift-kick scheme integrated in FP64 with numpy, with the same
force law (the direct sum with the softening added unsquared,
physics.cpp:392-434, or the softened kernel the walk evaluates).  Near the
origin both forms track the truth; far from it the plain form drifts by
about a float spacing per drift and the precision form does not.  The
velocities measure the trajectory cleanly; the downloaded positions are the
floats nearest the precise ones, so they can sit up to half a float spacing
(2.4e-4 at 6000) from the truth however exact the trajectory.
"""
import numpy as np
import pytest

from galaxy_step import Context, configs, default_params

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "vx", "vy", "ax", "ay", "id"]


def _params(**kw):
    p = default_params(G=1.0, theta=0.5, softening=0.05, dt=0.002, dt_start=0.002, direct_below=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _cluster(n=192, seed=3, center=(0.0, 0.0)):
    """A Plummer-like cluster of radius ~1, total mass 1, roughly virial."""
    rng = np.random.default_rng(seed)
    r = 0.6 / np.sqrt(np.maximum(rng.random(n) ** (-2.0 / 3.0) - 1.0, 1e-3))
    r = np.minimum(r, 3.0)
    a = rng.random(n) * 2 * np.pi
    x = r * np.cos(a)
    y = r * np.sin(a)
    v = 0.7 / np.sqrt(1.0 + r * r)
    vx = -v * np.sin(a) + 0.05 * rng.standard_normal(n)
    vy = v * np.cos(a) + 0.05 * rng.standard_normal(n)
    m = np.full(n, 1.0 / n)
    st = {
        "x": (x + center[0]).astype(np.float32),
        "y": (y + center[1]).astype(np.float32),
        "vx": vx.astype(np.float32),
        "vy": vy.astype(np.float32),
        "mass": m.astype(np.float32),
        "id": np.arange(n, dtype=np.uint32),
    }
    return st


def _truth(st, steps, G, eps, dt, squared):
    """FP64 kick-drift-kick with the frame's order: the uploaded acceleration
    is zero, then per frame v += a dt/2, x += v dt, a = F(x), v += a dt/2."""
    x = st["x"].astype(np.float64)
    y = st["y"].astype(np.float64)
    vx = st["vx"].astype(np.float64)
    vy = st["vy"].astype(np.float64)
    m = st["mass"].astype(np.float64)
    ax = np.zeros_like(x)
    ay = np.zeros_like(y)
    for _ in range(steps):
        vx += ax * dt * 0.5
        vy += ay * dt * 0.5
        x += vx * dt
        y += vy * dt
        dx = x[None, :] - x[:, None]
        dy = y[None, :] - y[:, None]
        d2 = dx * dx + dy * dy + (eps * eps if squared else eps)
        inv3 = d2 ** -1.5
        np.fill_diagonal(inv3, 0.0)
        ax = G * (dx * inv3 * m[None, :]).sum(axis=1)
        ay = G * (dy * inv3 * m[None, :]).sum(axis=1)
        vx += ax * dt * 0.5
        vy += ay * dt * 0.5
    return x, y, vx, vy


def _run(st, p, steps, precise):
    n = len(st["x"])
    with Context(n, precise_positions=precise) as c:
        c.upload(st)
        c.step(p, steps)
        c.sync()
        out = c.download(FIELDS)
    order = np.argsort(out["id"])
    return tuple(out[f][order].astype(np.float64) for f in ("x", "y", "vx", "vy"))


def _errs(got, truth):
    """(largest coordinate error, largest velocity error)"""
    pos = max(float(np.max(np.abs(got[0] - truth[0]))), float(np.max(np.abs(got[1] - truth[1]))))
    vel = float(np.max(np.hypot(got[2] - truth[2], got[3] - truth[3])))
    return pos, vel


FAR = (6000.3, -4000.7)


@pytest.mark.parametrize("form", ["direct", "walk"])
def test_precise_positions_win_back_accuracy_far_from_origin(form):
    steps = 300
    kw = dict(naive=1) if form == "direct" else dict(accuracy=2)
    p = _params(**kw)
    squared = form == "walk"  # the walk softens with eps^2, naiveGravity with eps
    res = {}
    for center in [(0.0, 0.0), FAR]:
        st = _cluster(center=center)
        truth = _truth(st, steps, p.G, p.softening, p.dt, squared)
        plain = _errs(_run(st, p, steps, False), truth)
        prec = _errs(_run(st, p, steps, True), truth)
        res[center] = (plain, prec)
        print(form, center, "position / velocity error: plain %.3g / %.3g  precise %.3g / %.3g"
              % (plain[0], plain[1], prec[0], prec[1]))
    (plain0, prec0), (plain_far, prec_far) = res[(0.0, 0.0)], res[FAR]
    # near the origin the two forms track the truth alike
    assert max(prec0) < 2e-4 and max(plain0) < 2e-4
    # far away the plain form loses about a float spacing per drift ...
    assert plain_far[1] > 1e-3 and plain_far[0] > 1e-3
    # ... and the precision form keeps its near-origin accuracy: velocities as
    # near the origin, positions within half a float spacing of the truth
    assert prec_far[1] < 2e-4 and prec_far[1] < plain_far[1] / 20
    half = 0.5 * float(np.spacing(np.float32(abs(FAR[0]) + 10.0)))
    assert prec_far[0] < half + 1e-5


def test_precise_matches_plain_near_origin():
    st, _ = configs.config0(seed=31, n=20000)
    p = _params(theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, accuracy=2)
    a = _run(st, p, 5, False)[:2]
    b = _run(st, p, 5, True)[:2]
    ext = float(np.ptp(st["x"]))
    assert np.max(np.abs(a[0] - b[0])) < 1e-4 * ext
    assert np.max(np.abs(a[1] - b[1])) < 1e-4 * ext


def test_precise_form_restarts_from_written_positions():
    """A position written by the host restarts its home form from that float:
    a context that ran before and a fresh one, given the same state, then
    compute the same frames bit for bit."""
    st = _cluster(center=FAR)
    n = len(st["x"])
    p = _params(accuracy=2)
    with Context(n, precise_positions=True) as a, Context(n, precise_positions=True) as b:
        a.upload(st)
        a.step(p, 40)
        a.sync()
        mid = a.download(FIELDS + ["mass"])
        # every position moved by one float spacing: none keeps its view
        mid["x"] = np.nextafter(mid["x"], np.float32(np.inf))
        mid["y"] = np.nextafter(mid["y"], np.float32(-np.inf))
        a.upload(mid)
        b.upload(mid)
        a.step(p, 40)
        b.step(p, 40)
        a.sync()
        b.sync()
        ra, rb = a.download(FIELDS), b.download(FIELDS)
    for f in FIELDS:
        np.testing.assert_array_equal(ra[f], rb[f], err_msg=f)
