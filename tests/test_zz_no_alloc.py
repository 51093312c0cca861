"""ge_diagnostics_noalloc, the entry point that allocates nothing, against
ge_diagnostics: the same bits on consecutive calls that share the fixed
device scratch (the second call on a smaller input), and a refusal, before
any launch, of a potential sum larger than that scratch holds."""
import numpy as np
import pytest

from galaxy_step import Context, GeError, configs, default_params

pytestmark = pytest.mark.gpu

FIELDS = ("mass", "px", "py", "lz", "kinetic", "potential", "comx", "comy")


def _params():
    return default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01)


def _same(a, b, what):
    for f in FIELDS:
        assert a[f] == b[f], f"{what}: {f} {a[f]!r} != {b[f]!r}"


def test_noalloc_equals_allocating_form_on_consecutive_calls():
    p = _params()
    st, _ = configs.config0(n=20000)
    with Context(20000) as c:
        c.upload(st)
        c.step(p, 2)
        c.sync()
        for pot in (True, False):
            _same(c.diagnostics(0, p, potential=pot), c.diagnostics(0, p, potential=pot, noalloc=True),
                  f"20000 particles, potential={pot}")
    # second context, smaller systems: the scratch still holds the larger
    # call's partials, which must not leak into these sums
    sizes = [3000, 1200, 2500]
    systems = [configs.config3_system(k, m, 60.0 + 50 * k, seed=4)[0] for k, m in enumerate(sizes)]
    batch = {k: np.concatenate([s[k] for s in systems]) for k in ("x", "y", "vx", "vy", "mass", "id")}
    with Context(sum(sizes), max_systems=3) as c:
        c.upload(batch, counts=sizes)
        c.step(p, 1)
        c.sync()
        for sys in (0, 1, 2):
            _same(c.diagnostics(sys, p, potential=True), c.diagnostics(sys, p, potential=True, noalloc=True),
                  f"system {sys} of {sizes}")


def test_noalloc_limits():
    p = _params()
    n = 4 * 1024 * 1024 + 1  # one particle more than the potential scratch covers
    rng = np.random.default_rng(3)
    st = {"x": rng.uniform(-1e3, 1e3, n).astype(np.float32), "y": rng.uniform(-1e3, 1e3, n).astype(np.float32),
          "vx": rng.normal(0, 1, n).astype(np.float32), "vy": rng.normal(0, 1, n).astype(np.float32),
          "mass": rng.uniform(0.5, 1.5, n).astype(np.float32)}
    with Context(n) as c:
        c.upload(st)
        with pytest.raises(GeError) as e:
            c.diagnostics(0, p, potential=True, noalloc=True)
        assert e.value.code == -5  # GE_E_CAPACITY
        # the moments have no size limit (grid-stride blocks, at most 1024)
        _same(c.diagnostics(0, p), c.diagnostics(0, p, noalloc=True), f"{n} particles, moments")
