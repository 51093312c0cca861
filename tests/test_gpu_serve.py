"""GPU tests of the host layer: the autotuner and the continuous-batching
request server (include/ge_serve.h).

Each served request must produce the same result as the same request run
alone (ge_upload + ge_step on a fresh context with the server's parameters).
Gravity requests share one batched context with systems joining and leaving
between rounds; SPH requests run on their own contexts concurrently.  The
tolerance is zero: the batched step is per-system independent, so the
results are bit-identical, and the test asserts exactly that.
"""
import numpy as np
import pytest

from galaxy_step import Context, configs, default_params
from galaxy_step import serve as S

pytestmark = pytest.mark.gpu

OUT = ("x", "y", "vx", "vy", "ax", "ay", "id")


def _gravity_params(**kw):
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0,
                       accuracy=1, compensated=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _sph_params():
    _, prm = configs.config2(n_target=100, n_impactor=20)
    p = default_params(direct_below=0)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(p, k, prm[k])
    p.dt_start = p.dt
    return p


def _requests(seed=11, n_grav=8, n_sph=3):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_grav + n_sph):
        sph = k >= n_grav
        n = int(rng.integers(1500, 4000 if sph else 12000))
        st, _ = configs.config4_request(n, "sph" if sph else "gravity", seed=100 + k)
        reqs.append({"state": st, "steps": int(rng.integers(2, 6 if sph else 10)), "sph": int(sph),
                     "arrival_s": float(rng.uniform(0.0, 0.03))})
    return reqs


def _alone(rq, pg, ps):
    st = rq["state"]
    n = len(st["x"])
    with Context(n, sph=bool(rq["sph"])) as c:
        c.upload(st)
        c.step(ps if rq["sph"] else pg, rq["steps"])
        c.sync()
        return c.download(list(OUT))


def _cfg(**kw):
    base = dict(batch_capacity=1 << 17, max_systems=16, sph_slots=2, sph_capacity=8192, round_frames=3)
    base.update(kw)
    return S.default_server_config(**base)


@pytest.mark.parametrize("batching", [1, 0])
def test_served_results_equal_alone(batching):
    reqs = _requests()
    pg, ps = _gravity_params(), _sph_params()
    with S.Server(_cfg(batching=batching), pg, ps) as srv:
        outs, results, stats = srv.run(reqs, OUT)
    assert stats["completed"] == len(reqs)
    assert all(r["status"] == 0 for r in results)
    assert 0 < stats["p50_s"] <= stats["p99_s"]
    if batching:
        assert stats["max_resident"] > 1
    else:
        assert stats["max_resident"] == 1
    for k, rq in enumerate(reqs):
        ref = _alone(rq, pg, ps)
        for f in OUT:
            np.testing.assert_array_equal(outs[k][f], ref[f], err_msg=f"request {k} field {f}")


def test_server_rejects_oversized_and_empty_requests():
    st, _ = configs.config4_request(3000, "gravity", seed=5)
    reqs = [{"state": st, "steps": 3, "sph": 0, "arrival_s": 0.0},
            {"state": st, "steps": 0, "sph": 0, "arrival_s": 0.0},
            {"state": st, "steps": 3, "sph": 1, "arrival_s": 0.0}]  # no SPH fields, no slot big enough
    with S.Server(_cfg(sph_slots=0), _gravity_params(), None) as srv:
        outs, results, stats = srv.run(reqs, OUT)
    assert results[0]["status"] == 0
    assert results[1]["status"] != 0 and results[2]["status"] != 0
    assert stats["completed"] == 1


def test_autotune_choice_and_state_restored():
    sizes = [5000, 1200, 9000]
    systems = [configs.config3_system(k, m, 80.0 + 40 * k, seed=8)[0] for k, m in enumerate(sizes)]
    batch = {k: np.concatenate([s[k] for s in systems]) for k in ("x", "y", "vx", "vy", "mass", "id")}
    n = len(batch["x"])
    p = _gravity_params()
    with Context(n, max_systems=3) as a, Context(n, max_systems=3) as b:
        a.upload(batch, counts=sizes)
        b.upload(batch, counts=sizes)
        a.step(p, 2)
        b.step(p, 2)
        leaf, block, r = S.autotune(a, p, frames=2)
        assert leaf in (8, 16, 32) and block in (64, 128, 256)
        assert r.ms_per_frame > 0 and r.reference_ms > 0
        assert r.ms_per_frame == min(min(row) for row in r.table_ms)
        a.step(p, 3)
        b.step(p, 3)
        a.sync()
        b.sync()
        da, db = a.download(list(OUT)), b.download(list(OUT))
    for f in OUT:
        np.testing.assert_array_equal(da[f], db[f], err_msg=f)


def test_run_after_an_aborted_run_is_unaffected():
    """A run stopped by the round limit while batch rounds and SPH chunks are
    in flight leaves nothing behind: the next run on the same server gives
    every request the result it has alone."""
    reqs = _requests(seed=23)
    pg, ps = _gravity_params(), _sph_params()
    with S.Server(_cfg(round_frames=1), pg, ps) as srv:
        srv.set_round_limit(2)
        with pytest.raises(S.GeError) as e:
            srv.run(reqs, OUT)
        assert e.value.code == -5  # GE_E_CAPACITY: round limit reached
        srv.set_round_limit(0)
        outs, results, stats = srv.run(reqs, OUT)
    assert stats["completed"] == len(reqs)
    assert all(r["status"] == 0 for r in results)
    for k, rq in enumerate(reqs):
        ref = _alone(rq, pg, ps)
        for f in OUT:
            np.testing.assert_array_equal(outs[k][f], ref[f], err_msg=f"request {k} field {f}")
