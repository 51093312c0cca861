"""The benchmark workloads at their stated sizes on the GPU.

cfg1  1M two-disk collision: 200 frames of the reference walk against the
      oracle's 200 frames; the hexadecapole form's force error against the FP64
      direct sum at theta 0.5, 0.7 and 1.0.
cfg2  250k gravity + SPH impact: two frames against the oracle (SPH fields
      within 1e-4 of the field scale), momentum of the SPH forces conserved.
cfg3  512-system ensemble (12.8M particles) for 100 frames in one context:
      eight sampled systems bit-identical to the same system run alone, and
      every system's force-normalised error against the FP64 direct sum.
cfg4  Poisson stream at 200 requests/s with sizes up to 500k particles and 30 %
      gravity + SPH: every request completes and a sample, including the
      largest of each kind, is bit-identical to the request run alone.
"""
import math

import numpy as np
import pytest

import oracle as O
from galaxy_step import Context, configs, default_params
from galaxy_step import serve as S

pytestmark = pytest.mark.gpu


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: v[o] for k, v in d.items()}


def _errors(ax, ay, rx, ry):
    e2 = (ax - rx) ** 2 + (ay - ry) ** 2
    a2 = rx ** 2 + ry ** 2
    return float(np.sqrt(np.mean(e2 / a2))), float(np.sqrt(np.mean(e2) / np.mean(a2)))


# ---- cfg1 -------------------------------------------------------------------

def test_cfg1_200_frames_match_oracle():
    """The reference walk (theta 0.7, eps 1, dt 0.005) for the configuration's
    200 frames.  The GPU's rsqrt and fused multiply-adds round differently
    from the oracle's, and the collision is chaotic near the disks' central
    bodies (mass 5e4, softening 1): there any rounding difference grows
    without bound over 200 frames.  The yardstick is therefore the oracle's
    own sensitivity: the oracle run again from inputs that differ by one ulp
    in x.  At every percentile from the median to 99.9 the GPU's distance
    from the oracle must stay within twice the oracle's distance from itself,
    and the median within 1e-6 of the initial extent L."""
    st, prm = configs.config1(n=1_000_000)
    n = len(st["x"])
    ext0 = float(max(np.ptp(st["x"]), np.ptp(st["y"])))
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=prm["dt"], dt_start=prm["dt"], accuracy=0,
                       compensated=0)
    with Context(n) as c:
        c.upload(st)
        c.step(p, prm["steps"])
        c.sync()
        got = _by_id(c.download(["x", "y", "id"]))
    op = dict(G=1.0, theta=0.7, softening=1.0, dt=prm["dt"], dt_start=prm["dt"])

    def oracle_run(state):
        s = O.State(state)
        s.step(op, nsteps=prm["steps"])
        return _by_id({k: s[k].copy() for k in ("x", "y", "id")})

    ref = oracle_run(st)
    nudged = dict(st, x=np.nextafter(np.asarray(st["x"], np.float32), np.float32(np.inf)))
    ref2 = oracle_run(nudged)
    d_go = np.hypot(got["x"] - ref["x"], got["y"] - ref["y"])
    d_oo = np.hypot(ref2["x"] - ref["x"], ref2["y"] - ref["y"])
    pct = [50, 90, 99, 99.9]
    q_go, q_oo = np.percentile(d_go, pct), np.percentile(d_oo, pct)
    print(f"cfg1 200 frames, |dpos| at percentiles {pct}: GPU vs oracle {np.array2string(q_go, precision=2)}, "
          f"oracle vs oracle one ulp apart {np.array2string(q_oo, precision=2)} (L = {ext0:.0f})")
    assert q_go[0] < 1e-6 * ext0
    assert np.all(q_go <= 2.0 * q_oo + 1e-6 * ext0), (q_go, q_oo)


@pytest.mark.parametrize("theta,bound", [(0.5, 3e-4), (0.7, 1e-3), (1.0, 4e-3)])
def test_cfg1_hexadecapole_theta_sweep_error(theta, bound):
    """The headline form (moments to the hexadecapole, Kahan accumulation) at
    each theta of the sweep: RMS per-particle force error against the FP64
    direct sum on 4096 sampled particles; at theta 0.7 the 1e-3 target."""
    st, prm = configs.config1(n=1_000_000)
    n = len(st["x"])
    p = default_params(G=1.0, theta=theta, softening=1.0, dt=prm["dt"], dt_start=prm["dt"], accuracy=2,
                       compensated=1)
    with Context(n) as c:
        c.upload(st)
        c.step(p, 2)
        c.compute_gravity(p)
        d = c.download(["ax", "ay"])
        t = np.sort(np.random.default_rng(1).choice(n, 4096, replace=False)).astype(np.uint32)
        rx, ry = c.direct_fp64(t, p)
    per, norm = _errors(d["ax"][t], d["ay"][t], rx, ry)
    print(f"cfg1 hexadecapole theta={theta}: RMS per particle {per:.3e}, force-normalised {norm:.3e}")
    assert per < bound


# ---- cfg2 -------------------------------------------------------------------

def test_cfg2_full_size_frames_match_oracle():
    st, prm = configs.config2()
    n = len(st["x"])
    assert n == 250_000
    p = default_params(direct_below=0)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(p, k, prm[k])
    p.dt_start = p.dt
    fields = ("x", "y", "vx", "vy", "pred_dens", "press", "pfx", "pfy", "sph_mass", "id")
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, 2)
        c.sync()
        got = _by_id(c.download(list(fields)))
    s = O.State(st, sph=True)
    s.step(dict(prm, dt_start=prm["dt"], direct_below=0), nsteps=2)
    ref = _by_id({k: s[k].copy() for k in fields})
    ext = float(np.ptp(ref["x"]))
    assert np.abs(got["x"] - ref["x"]).max() < 1e-4 * ext
    assert np.abs(got["y"] - ref["y"]).max() < 1e-4 * ext
    for f in ("pred_dens", "press", "pfx", "pfy"):
        scale = float(np.abs(ref[f]).max())
        assert scale > 0, f
        err = np.abs(got[f].astype(np.float64) - ref[f])
        bad = np.nonzero(err > 1e-4 * scale + 1e-4 * np.abs(ref[f]))[0]
        assert bad.size == 0, (f, bad[:8], got[f][bad[:8]], ref[f][bad[:8]])
    # SPH pair forces cancel: momentum sum_i sphMass_i pressF_i is conserved
    fx = (got["pfx"].astype(np.float64) * got["sph_mass"]).sum()
    fy = (got["pfy"].astype(np.float64) * got["sph_mass"]).sum()
    scale = float(np.abs(got["pfx"] * got["sph_mass"]).sum() + np.abs(got["pfy"] * got["sph_mass"]).sum())
    assert abs(fx) < 1e-4 * scale and abs(fy) < 1e-4 * scale


# ---- cfg3 -------------------------------------------------------------------

def test_cfg3_ensemble_of_512_systems():
    sizes, scales = configs.config3_sizes(systems=512)
    systems = [configs.config3_system(k, int(sizes[k]), float(scales[k]))[0] for k in range(len(sizes))]
    keys = ("x", "y", "vx", "vy", "mass", "id")
    batch = {k: np.concatenate([s[k] for s in systems]) for k in keys}
    total = int(sizes.sum())
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, compensated=1, accuracy=2,
                       direct_below=0, leaf_capacity=16, use_graph=1)
    steps = 100
    rng = np.random.default_rng(7)
    sample = sorted(set(rng.choice(len(sizes), 6, replace=False).tolist()) |
                    {int(np.argmax(sizes)), int(np.argmin(sizes))})
    out = ("x", "y", "vx", "vy", "ax", "ay", "id")
    with Context(total, max_systems=len(sizes)) as c:
        c.upload(batch, counts=sizes)
        c.step(p, steps)
        c.sync()
        batched = {s: c.download_system(s, list(out)) for s in sample}
        # per-system force error of the final state vs the FP64 direct sum
        off = c.system_offsets()
        c.compute_gravity(p)
        d = c.download(["ax", "ay"])
        tgt, owner = [], []
        for s in range(len(sizes)):
            m = int(off[s + 1] - off[s])
            pick = rng.choice(m, min(256, m), replace=False) + off[s]
            tgt.append(pick)
            owner.append(np.full(len(pick), s))
        tgt = np.concatenate(tgt).astype(np.uint32)
        owner = np.concatenate(owner)
        rx, ry = c.direct_fp64(tgt, p)
    e2 = (d["ax"][tgt] - rx) ** 2 + (d["ay"][tgt] - ry) ** 2
    a2 = rx ** 2 + ry ** 2
    norm_sys = np.sqrt(np.bincount(owner, e2) / np.bincount(owner, a2))
    per_sys = np.sqrt(np.bincount(owner, e2 / a2) / np.bincount(owner))
    print(f"cfg3: force-normalised error per system: median {np.median(norm_sys):.2e}, max {norm_sys.max():.2e}; "
          f"per particle: median {np.median(per_sys):.2e}, max {per_sys.max():.2e}")
    assert norm_sys.max() < 1e-4
    assert np.median(per_sys) < 1e-4
    # each sampled system alone on its own context: the same bits
    with Context(int(sizes.max())) as c:
        for s in sample:
            c.upload(systems[s])
            c.step(p, steps)
            c.sync()
            alone = c.download(list(out))
            for f in out:
                np.testing.assert_array_equal(batched[s][f], alone[f], err_msg=f"system {s} ({sizes[s]}) {f}")


# ---- cfg4 -------------------------------------------------------------------

def test_cfg4_poisson_stream_results_equal_alone():
    _, sp = configs.config2(n_target=10, n_impactor=10)
    pg = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, compensated=1, accuracy=2,
                        direct_below=0, leaf_capacity=16)
    ps = default_params(direct_below=0, compensated=1, accuracy=2, leaf_capacity=16)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(ps, k, sp[k])
    ps.dt_start = ps.dt
    rng = np.random.default_rng(12)
    reqs = []
    for k in range(40):
        n = int(round(math.exp(rng.uniform(math.log(2000), math.log(500_000)))))
        if k == 0:
            n = 500_000  # the largest size of the stream, gravity
        if k == 1:
            n = 499_000  # and nearly the largest, gravity + SPH
        kind = "sph" if (k == 1 or (k > 1 and rng.random() < 0.3)) else "gravity"
        st, _ = configs.config4_request(n, kind, seed=int(rng.integers(0, 2 ** 31)))
        reqs.append({"state": st, "steps": int(rng.integers(10, 201)), "sph": int(kind == "sph")})
    arr = np.cumsum(np.random.default_rng(5).exponential(1.0 / 200.0, len(reqs)))
    for r, t in zip(reqs, arr):
        r["arrival_s"] = float(t)
    out = ("x", "y", "vx", "vy", "id")
    cfg = S.default_server_config(batching=1, sph_slots=6, sph_capacity=1 << 19, batch_capacity=1 << 25,
                                  max_systems=512, round_frames=8, use_graph=1)
    with S.Server(cfg, pg, ps) as srv:
        outs, results, stats = srv.run(reqs, out)
    assert all(r["status"] == 0 for r in results)
    assert stats["completed"] == len(reqs) and stats["max_resident"] > 1
    print(f"cfg4 200 req/s: {stats['solves_per_s']:.1f} solves/s, p50 {stats['p50_s']:.3f} s, "
          f"p99 {stats['p99_s']:.3f} s")
    sample = [0, 1] + sorted(np.random.default_rng(2).choice(np.arange(2, len(reqs)), 4, replace=False).tolist())
    for k in sample:
        rq = reqs[k]
        n = len(rq["state"]["x"])
        with Context(n, sph=bool(rq["sph"])) as c:
            c.upload(rq["state"])
            c.step(ps if rq["sph"] else pg, rq["steps"])
            c.sync()
            alone = c.download(list(out))
        for f in out:
            np.testing.assert_array_equal(outs[k][f], alone[f], err_msg=f"request {k} (n={n}, sph={rq['sph']}) {f}")
