#!/usr/bin/env python3
"""Benchmark of the 2D physics step on one MI355X.

Headline (`value`): particle-updates/s of the 1,000,000-particle two-disk
collision at theta = 0.7 (eps = 1, dt = 0.005, G = 1, seeded synthetic input)
in the accuracy form: moments to the hexadecapole, Kahan-compensated FP32
accumulation, leaf capacity and walk block size picked by the autotuner.  Its
RMS force error against an FP64 direct sum (4096 sampled particles, on the
device) is reported beside it, together with:

  - the plain form: the reference's monopole walk and leaf rule on the same
    input (the parity path), with its own error;
  - the CPU baseline: the oracle's OpenMP C restatement of the reference step
    on this host, on a bounded number of frames of the same workload;
  - the batched form: 512 independent systems of 1,000 .. 131,072 particles
    in one context, 100 steps, against the same systems solved one by one;
  - the served form: Poisson request streams at 50, 200 and 800 requests/s
    through the continuous-batching server, against the plain form of the
    server (one request on the device at a time); p50/p99 time to
    completion, solves/s and particle-updates/s.

A step is one full frame through the C ABI (ge_step): integrateStart +
bounding box, Morton keys, radix sort, gather, quadtree build, node moments,
tree walk with the fused integrateEnd kick.  State stays resident in HBM; the
timed region covers exactly K frames between two device synchronisations,
with hipEvent pairs around every phase on the context's stream, from which
the roofline of the dominant kernel (the walk) is computed.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "galaxy-engine_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# Algorithmic counts, committed once (bench/roofline_constants.json): 14 FP32
# flops per gravity interaction (accepted node or leaf partner), SPH flops per
# pair per pass, compulsory HBM bytes per particle and per node of each phase,
# and the MI355X peaks.  The expression-level flops of the accuracy modes'
# node terms (32, 107) are reported beside the 14-flop figure, labelled.
with open(os.path.join(ROOT, "bench", "roofline_constants.json")) as _f:
    CONST = json.load(_f)
FLOPS_PAIR = CONST["gravity"]["flops_per_interaction"]
FLOPS_PAIR3 = CONST["gravity3d"]["flops_per_interaction"]
FLOPS_NODE = {int(k): v for k, v in CONST["gravity"]["expression_flops_per_node"].items()}
PEAK_FP32_TFLOPS = CONST["peaks"]["fp32_flops_per_s"] / 1e12
PEAK_HBM_GBS = CONST["peaks"]["hbm_bytes_per_s"] / 1e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--theta", type=float, default=0.7)
    ap.add_argument("--accuracy", type=int, default=2,
                    help="0 reference walk, 1 quadrupole, 3 octupole, 2 hexadecapole")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="bound on the CPU-baseline sample (0 disables it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-sph", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--only-3d", action="store_true", help="run only the 3D leg (profiling) and print its JSON")
    ap.add_argument("--only-sph", action="store_true", help="run only the 2D SPH impact leg (profiling)")
    ap.add_argument("--only-sph3d", action="store_true", help="run only the 3D SPH impact leg (profiling)")
    ap.add_argument("--only-couplings", action="store_true",
                    help="run only the leg of the steps either side and the couplings")
    ap.add_argument("--no-couplings", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rates", type=str, default="10,50,200,800")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain elementwise copy of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calibrate_copy(mib):
    """A copy kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE:
    out = -in over `mib` MiB of float32 reads exactly mib MiB and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise copy", "bytes_read": n * 4, "bytes_written": n * 4,
                      "launches": 5}), flush=True)


def traffic_from_profiles(accuracy, block):
    """HBM bytes per walk launch measured with rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, calibrated on a copy kernel) for this form and block size."""
    path = os.path.join(ROOT, "profiles", "bh_walk_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("accuracy") != accuracy:
        return None
    return t.get("bytes_per_launch_by_block", {}).get(str(block or 256))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """The host cores this process may use: OMP_NUM_THREADS when set (the
    GPU host's share), else the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(st, prm, budget_s):
    """The oracle's OpenMP restatement of the reference step on this host, at
    the host's thread count and at half of it (the reference's default,
    main.cpp:20-22), each for about budget_s / 2 seconds of frames."""
    import oracle as O
    full = _host_threads()
    n = len(st["x"])
    p = dict(prm, dt_start=prm["dt"])
    runs = {}
    for threads in (full, max(1, full // 2)):
        O.set_threads(threads)
        s = O.State(st)
        t0 = time.perf_counter()
        steps = 0
        while True:
            s.step(p, 1)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget_s / 2 or steps >= 30:
                break
        runs[threads] = (n * steps / el, steps, el)
    O.set_threads(full)
    v, steps, el = runs[full]
    hv, hsteps, hel = runs[max(1, full // 2)]
    return {"value": v, "unit": "particle-updates/s", "cores": full, "kind": "port",
            "sample": f"{steps} full frames of the same {n}-particle workload (theta={prm['theta']}, "
                      f"reference walk) through the OpenMP C restatement of the reference step, {el:.1f} s",
            "half_threads": {"value": hv, "cores": max(1, full // 2), "sample": f"{hsteps} frames, {hel:.1f} s",
                             "note": "the reference's default thread count, half the host's"},
            "cpu_model": _cpu_model(), "compiler": O.build_info()}


def step_roofline(phase_ms, ms_per_step, n, m_nodes, interactions, accuracy, sph_flops=None):
    """T* = sum over phases of max(B / HBM peak, F / FP32 peak) from the
    committed constants; achieved = T* / T_measured."""
    bpp, bpn = CONST["bytes_per_particle"], CONST["bytes_per_node"]
    bw, fl = CONST["peaks"]["hbm_bytes_per_s"], CONST["peaks"]["fp32_flops_per_s"]
    rows = {}
    tstar = 0.0
    for ph, ms in phase_ms.items():
        if ph not in bpp and ph != "gravity" and ph != "sph":
            continue
        if ph == "gravity":
            b = bpp["walk"] * n + bpn["walk"] * m_nodes + (bpn["multipole"] * m_nodes if accuracy != 0 else 0)
            f = FLOPS_PAIR * interactions
        elif ph == "sph":
            b = bpp["sph_passes"] * n
            f = sph_flops or 0
        else:
            b = bpp[ph] * n + bpn.get(ph, 0) * m_nodes
            f = 0
        if ph == "sph_grid":
            b *= 2  # two grids per frame
        t = max(b / bw, f / fl) * 1e3
        rows[ph] = {"bytes": int(b), "flops": int(f), "t_star_ms": round(t, 5), "measured_ms": ms,
                    "frac": round(t / ms, 4) if ms > 0 else None}
        tstar += t
    return {"t_star_ms": round(tstar, 5), "t_measured_ms": ms_per_step, "achieved": tstar / ms_per_step,
            "phases": rows,
            "formula": "T* = sum_p max(B_p / 8.0e12 B/s, F_p / 157.3e12 flop/s), constants in "
                       "bench/roofline_constants.json"}


def force_errors(c, p, n, samples=4096, seed=0):
    """Per-particle and force-normalised RMS error of the current accelerations
    against the FP64 direct sum on the device, on `samples` sampled particles."""
    import numpy as np
    c.compute_gravity(p)
    d = c.download(["ax", "ay"])
    t = np.sort(np.random.default_rng(seed).choice(n, min(samples, n), replace=False)).astype(np.uint32)
    rx, ry = c.direct_fp64(t, p)
    e2 = (d["ax"][t] - rx) ** 2 + (d["ay"][t] - ry) ** 2
    a2 = rx ** 2 + ry ** 2
    return {"rms_per_particle": float(np.sqrt(np.mean(e2 / a2))),
            "rms_force_normalised": float(np.sqrt(np.mean(e2) / np.mean(a2))),
            "samples": int(len(t))}


def single_system(a, st, prm, accuracy, tune):
    """Times K frames of one form on the 1M system; returns a dict."""
    import numpy as np
    from galaxy_step import Context, default_params
    from galaxy_step import serve as S
    n = len(st["x"])
    p = default_params(G=prm["G"], theta=a.theta, softening=prm["softening"], dt=prm["dt"],
                       dt_start=prm["dt"], compensated=1, accuracy=accuracy, direct_below=0)
    out = {}
    with Context(n) as c:
        tu0 = time.perf_counter()
        c.upload(st)
        out["upload_ms"] = (time.perf_counter() - tu0) * 1e3
        if tune:
            leaf, block, r = S.autotune(c, p, frames=3)
            p.leaf_capacity, p.walk_block = leaf, block
            out["autotune"] = {"leaf_capacity": leaf, "walk_block": block, "ms_per_frame": r.ms_per_frame,
                               "reference_leaf_rule_ms": r.reference_ms,
                               "table_ms": [[round(v, 4) for v in row] for row in r.table_ms]}
        # every pass starts from the uploaded state, so all passes time the
        # same frames
        # eager frames (one launch sequence per frame), no timers inside
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        te0 = time.perf_counter()
        c.step(p, a.steps)
        c.sync()
        te1 = time.perf_counter()
        # the same frames replayed through a hipGraph (ge_step's capture path):
        # the timed form
        c.upload(st)
        p.use_graph = 1
        c.step(p, a.warmup)
        c.sync()
        t0 = time.perf_counter()
        c.step(p, a.steps)
        c.sync()
        t1 = time.perf_counter()
        p.use_graph = 0
        # per-phase device time from HIP events around each phase, in a pass
        # of its own (the events add launch gaps, so it is not the timed pass)
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        c.profile_enable(True)
        c.step(p, a.steps)
        c.sync()
        phases = c.profile_read()
        c.profile_enable(False)
        td0 = time.perf_counter()
        _ = c.download(["x", "y", "vx", "vy", "mass", "id"])
        out["download_ms"] = (time.perf_counter() - td0) * 1e3
        out["errors"] = force_errors(c, p, n)
        nodes_acc, pairs = c.interaction_counts()
        out["node_count"] = c.node_count()
    el = t1 - t0
    out["ms_per_step"] = el * 1e3 / a.steps
    out["value"] = n * a.steps / el
    out["graph_ms_per_step"] = out["ms_per_step"]
    out["eager_ms_per_step"] = (te1 - te0) * 1e3 / a.steps
    out["phases"] = phases
    grav_ms, grav_launches = phases["gravity"]
    out["walk_ms"] = grav_ms / max(grav_launches, 1)
    alg = FLOPS_PAIR * (nodes_acc + pairs)
    expr = FLOPS_NODE[accuracy] * nodes_acc + FLOPS_PAIR * pairs
    out["walk_tflops"] = alg / (out["walk_ms"] * 1e-3) / 1e12
    out["walk_tflops_expression"] = expr / (out["walk_ms"] * 1e-3) / 1e12
    out["interactions"] = {"nodes": nodes_acc, "pairs": pairs, "per_particle": (nodes_acc + pairs) / n,
                           "flops": alg, "expression_flops": expr}
    out["params"] = {"accuracy": accuracy, "leaf_capacity": p.leaf_capacity, "walk_block": p.walk_block}
    return out


def sph_impact(a):
    """cfg2: gravity + SPH planetary impact, 250,000 particles, h = cell = 4,
    the reference's SPH constants, theta 0.7, dt 0.002 (parity form: the
    reference's walk and leaf rule)."""
    from galaxy_step import Context, configs, default_params
    st, prm = configs.config2()
    n = len(st["x"])
    p = default_params(direct_below=0, compensated=1)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(p, k, prm[k])
    p.dt_start = p.dt
    steps = max(a.steps, 10)
    with Context(n, sph=True) as c:
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        c.profile_enable(True)
        t0 = time.perf_counter()
        c.step(p, steps)
        c.sync()
        t1 = time.perf_counter()
        ph = c.profile_read()
        c.profile_enable(False)
        d = c.download(["x", "y", "ppx", "ppy", "flags"])
        nodes_acc, pairs = c.interaction_counts()
        m_nodes = c.node_count()
    el = t1 - t0
    pc = sph_pair_counts(d, prm["sph_h"])
    fpp = CONST["sph_flops_per_pair"]
    sph_flops = (fpp["viscosity_cohesion"] * pc["pos"] + fpp["density"] * (pc["pred"] + pc["visitors"])
                 + fpp["pressure"] * pc["pred"])
    phase_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items() if v[1]}
    sph_ms = phase_ms.get("sph", 0.0)
    step = step_roofline(phase_ms, el * 1e3 / steps, n, m_nodes, nodes_acc + pairs, 0, sph_flops=sph_flops)
    return {"workload": f"gravity + SPH impact, {n} particles (200k + 50k), h = cell = 4, theta 0.7, dt 0.002",
            "particle_updates_per_s": n * steps / el, "ms_per_step": el * 1e3 / steps, "steps": steps,
            "phase_ms": phase_ms,
            "sph_pairs": pc,
            "sph_passes": {"flops": int(sph_flops), "ms": sph_ms,
                           "tflops": sph_flops / (sph_ms * 1e-3) / 1e12 if sph_ms else None,
                           "frac_fp32": sph_flops / (sph_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS if sph_ms else None,
                           "gbps_compulsory": CONST["bytes_per_particle"]["sph_passes"] * n / (sph_ms * 1e-3) / 1e9
                           if sph_ms else None,
                           "note": "algorithmic flops: 44 per ordered pair within h at pos (viscosity and "
                                   "cohesion), 16 per pair within h at predPos including self (density), 32 per "
                                   "ordered pair within h at predPos (pressure), pairs counted on the last "
                                   "frame's state"},
            "step_roofline": step}


def sph_pair_counts(d, h):
    """Ordered pairs (i, j), j != i, of SPH visitors (SPH, not being drawn)
    closer than h at pos and at predPos, counted with a k-d tree."""
    import numpy as np
    from scipy.spatial import cKDTree
    vis = ((d["flags"] & 2) != 0) & ((d["flags"] & 4) == 0)
    out = {"visitors": int(vis.sum())}
    for key, (fx, fy) in (("pos", ("x", "y")), ("pred", ("ppx", "ppy"))):
        pts = np.stack([d[fx][vis], d[fy][vis]], 1).astype(np.float64)
        # query_pairs keeps r <= h; the passes keep r < h (equal distances are measure-zero here)
        out[key] = 2 * len(cKDTree(pts).query_pairs(float(h), output_type="ndarray"))
    return out


def three_d(a, cpu_seconds):
    """The 3D gravity frame (mode3D's gravity path): two 500k-particle 3D
    galaxies, theta 0.7, the reference's monopole walk, octree and leaf rule
    (softening 1 added unsquared), Kahan-compensated accumulation."""
    import numpy as np
    from galaxy_step import configs
    from galaxy_step.lib3d import Context3D, default_params3
    st, prm = configs.config3d(n=a.n, theta=a.theta)
    n = len(st["x"])
    p = default_params3(G=prm["G"], theta=a.theta, softening=prm["softening"], dt=prm["dt"], dt_start=prm["dt"],
                        direct_below=0, compensated=1)
    steps = max(a.steps, 5)
    with Context3D(n) as c:
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        t0 = time.perf_counter()
        c.step(p, steps)
        c.sync()
        t1 = time.perf_counter()
        # the walk alone (synchronous call around one launch, interactions counted)
        walk = []
        for _ in range(5):
            w0 = time.perf_counter()
            c.compute_gravity(p)
            walk.append(time.perf_counter() - w0)
        inter = c.interaction_counts()
        d = c.download(["ax", "ay", "az"])
        t = np.sort(np.random.default_rng(0).choice(n, 4096, replace=False)).astype(np.uint32)
        rx, ry, rz = c.direct_fp64(t, p.G, p.softening)
        nodes = c.node_count()
    e2 = (d["ax"][t] - rx) ** 2 + (d["ay"][t] - ry) ** 2 + (d["az"][t] - rz) ** 2
    a2 = rx ** 2 + ry ** 2 + rz ** 2
    el = t1 - t0
    walk_ms = min(walk) * 1e3
    flops = FLOPS_PAIR3 * inter[0]
    out = {"workload": f"3D Barnes-Hut frame, two-galaxy collision ({n} particles), theta {a.theta}, "
                       "reference octree and walk",
           "particle_updates_per_s": n * steps / el, "ms_per_step": el * 1e3 / steps, "steps": steps,
           "walk_ms": walk_ms, "walk_tflops": flops / (walk_ms * 1e-3) / 1e12,
           "interactions_per_particle": inter[0] / n, "accepted_nodes": inter[1], "node_count": nodes,
           "force_error_vs_fp64": {"rms_per_particle": float(np.sqrt(np.mean(e2 / a2))),
                                   "rms_force_normalised": float(np.sqrt(np.mean(e2) / np.mean(a2))),
                                   "samples": int(len(t))}}
    if cpu_seconds > 0:
        import oracle3d as O3
        s = O3.State3(st)
        q = dict(prm, theta=a.theta, dt_start=prm["dt"], direct_below=0)
        c0 = time.perf_counter()
        k = 0
        while True:
            s.step(q, 1)
            k += 1
            if time.perf_counter() - c0 >= cpu_seconds or k >= 20:
                break
        cel = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": n * k / cel, "unit": "particle-updates/s", "kind": "port",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "sample": f"{k} full 3D frames of the same workload through the OpenMP C "
                                         f"restatement, {cel:.1f} s"}
    return out


def sph3d(a, cpu_seconds):
    """3D gravity + SPH3D impact (mode3D with isSPHEnabled): balls of 200k and
    50k particles at 0.3 per unit volume, h = cell = 3, theta 0.7, dt 0.002,
    the reference's octree walk and SPH3D passes."""
    from galaxy_step import configs
    from galaxy_step.lib3d import Context3D, default_params3
    st, prm = configs.config3d_sph()
    n = len(st["x"])
    p = default_params3(direct_below=0, compensated=1)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(p, k, prm[k])
    p.dt_start = p.dt
    steps = max(a.steps, 10)
    with Context3D(n, sph=True) as c:
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        t0 = time.perf_counter()
        c.step(p, steps)
        c.sync()
        t1 = time.perf_counter()
        cnt = c.neighbor_counts(p)
    el = t1 - t0
    out = {"workload": f"3D gravity + SPH3D impact, {n} particles (200k + 50k balls), h = cell = 3, "
                       "theta 0.7, dt 0.002",
           "particle_updates_per_s": n * steps / el, "ms_per_step": el * 1e3 / steps, "steps": steps,
           "candidates_per_particle": float(cnt.mean())}
    if cpu_seconds > 0:
        import oracle3d as O3
        s = O3.State3(st)
        q = {k: getattr(p, k) for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell",
                                         "sph_unit_mass", "viscosity", "cohesion_coef", "delta", "stiff_mult",
                                         "sph_max_vel")}
        q.update(dt_start=p.dt, direct_below=0)
        c0 = time.perf_counter()
        k = 0
        while True:
            s.step(q, 1)
            k += 1
            if time.perf_counter() - c0 >= cpu_seconds or k >= 5:
                break
        cel = time.perf_counter() - c0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": n * k / cel, "unit": "particle-updates/s", "kind": "port",
                               "cores": threads,
                               "sample": f"{k} full 3D gravity + SPH3D frames of the same workload through the C "
                                         f"restatement ({threads} OpenMP threads for gravity, the SPH passes on "
                                         f"one thread), {cel:.1f} s"}
    return out


def batched(a, accuracy):
    """cfg3: the batched form against the same systems one by one."""
    import numpy as np
    from galaxy_step import Context, configs, default_params
    sizes, scales = configs.config3_sizes(systems=a.batch_systems)
    systems = [configs.config3_system(k, int(sizes[k]), float(scales[k]))[0] for k in range(len(sizes))]
    keys = ("x", "y", "vx", "vy", "mass")
    batch = {k: np.concatenate([s[k] for s in systems]) for k in keys}
    total = int(sizes.sum())
    p = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, compensated=1,
                       accuracy=accuracy, direct_below=0, leaf_capacity=16, use_graph=1)
    steps = a.batch_steps
    log(f"batched: {len(sizes)} systems, {total} particles, {steps} steps")
    from galaxy_step import serve as S
    with Context(total, max_systems=len(sizes)) as c:
        c.upload(batch, counts=sizes)
        # leaf capacity in {8, 16, 32} and walk block in {64, 128, 256} picked on the ensemble itself
        leaf, block, tr = S.autotune(c, p, frames=3)
        p.leaf_capacity, p.walk_block = leaf, block
        tune = {"leaf_capacity": leaf, "walk_block": block, "ms_per_frame": tr.ms_per_frame,
                "table_ms": [[round(v, 4) for v in row] for row in tr.table_ms]}
        c.upload(batch, counts=sizes)
        c.step(p, 4)  # warm the code paths and the graph
        c.sync()
        t0 = time.perf_counter()
        c.upload(batch, counts=sizes)
        c.step(p, steps)
        c.sync()
        _ = c.download(["x", "y", "vx", "vy"])
        t1 = time.perf_counter()
        # per-system force error vs FP64 direct sum on up to 256 particles each
        off = c.system_offsets()
        c.compute_gravity(p)
        d = c.download(["ax", "ay"])
        rng = np.random.default_rng(3)
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
    rel2 = e2 / a2
    per_sys = np.sqrt(np.bincount(owner, rel2) / np.bincount(owner))
    norm_sys = np.sqrt(np.bincount(owner, e2) / np.bincount(owner, a2))
    # the system's per-particle RMS without its largest single term: a
    # central body near equilibrium (|a| far below the system's median) can
    # carry most of it
    worst = int(np.argmax(per_sys))
    w_rel = np.sort(rel2[owner == worst])
    worst_rest = float(np.sqrt(w_rel[:-1].mean())) if len(w_rel) > 1 else 0.0
    t_batch = t1 - t0
    # plain form: every system alone on one context, same parameters
    t0 = time.perf_counter()
    with Context(int(sizes.max())) as c:
        for s in systems:
            c.upload(s)
            c.step(p, steps)
            c.sync()
            _ = c.download(["x", "y", "vx", "vy"])
    t_plain = time.perf_counter() - t0
    ns = len(sizes)
    return {"workload": f"{ns} systems, 1,000..131,072 particles (log-uniform), {total} particles, "
                        f"{steps} steps, theta 0.7",
            "batched_solves_per_s": ns / t_batch, "plain_solves_per_s": ns / t_plain,
            "speedup": t_plain / t_batch,
            "batched_particle_updates_per_s": total * steps / t_batch,
            "plain_particle_updates_per_s": total * steps / t_plain,
            "batched_s": t_batch, "plain_s": t_plain,
            "force_error_rms_per_particle": {"median_system": float(np.median(per_sys)),
                                             "max_system": float(per_sys.max()),
                                             "max_system_without_its_largest_term": worst_rest,
                                             "max_system_particles": int(sizes[worst])},
            "force_error_rms_normalised": {"median_system": float(np.median(norm_sys)),
                                           "max_system": float(norm_sys.max())},
            "params": {"accuracy": accuracy, "leaf_capacity": p.leaf_capacity, "walk_block": p.walk_block,
                       "graphs": True},
            "autotune": tune}


def _lattice(side, seed):
    """A jittered square lattice with right, up and up-right springs (random
    ids, shuffled constraint order), as the constraint parity tests build it."""
    import numpy as np
    from galaxy_step import constraints as GC
    F = np.float32
    rng = np.random.default_rng(seed)
    n = side * side
    gx, gy = np.meshgrid(np.arange(side), np.arange(side), indexing="xy")
    x = (gx.ravel() + rng.normal(0, 0.08, n)).astype(F)
    y = (gy.ravel() + rng.normal(0, 0.08, n)).astype(F)
    ids = rng.permutation(np.arange(100, 100 + 3 * n, 3)).astype(np.uint32)
    a = np.arange(n).reshape(side, side)
    rows = np.concatenate([np.stack([a[:, :-1].ravel(), a[:, 1:].ravel()], 1),
                           np.stack([a[:-1, :].ravel(), a[1:, :].ravel()], 1),
                           np.stack([a[:-1, :-1].ravel(), a[1:, 1:].ravel()], 1)])
    rng.shuffle(rows)
    cons = np.zeros(len(rows), GC.CONSTRAINT_DTYPE)
    cons["id1"], cons["id2"] = ids[rows[:, 0]], ids[rows[:, 1]]
    d = np.hypot(x[rows[:, 1]] - x[rows[:, 0]], y[rows[:, 1]] - y[rows[:, 0]]).astype(F)
    cons["rest_length"] = (d * rng.uniform(0.9, 1.1, len(rows))).astype(F)
    cons["original_length"] = d
    cons["stiffness"] = rng.uniform(20, 60, len(rows)).astype(F)
    cons["resistance"] = rng.uniform(0.05, 0.3, len(rows)).astype(F)
    cons["plasticity_point"] = (cons["resistance"] * F(0.5)).astype(F)
    st = {"x": x, "y": y, "vx": rng.normal(0, 0.5, n).astype(F), "vy": rng.normal(0, 0.5, n).astype(F),
          "ax": np.zeros(n, F), "ay": np.zeros(n, F),
          "mass": rng.uniform(0.5, 2.0, n).astype(F), "id": ids, "flags": np.zeros(n, np.uint32),
          "label": rng.choice([0, 1, 3, 5], n).astype(np.uint32), "temp": np.full(n, 300.0, F)}
    return st, cons


def _crowd(n, seed, box):
    """Particles crowded enough to overlap (render sizes 0.05-0.3), 10 % dark."""
    import numpy as np
    from galaxy_step.lib import GE_FLAG_DARK
    F = np.float32
    rng = np.random.default_rng(seed)
    st = {"x": rng.uniform(-box, box, n).astype(F), "y": rng.uniform(-box, box, n).astype(F),
          "vx": rng.normal(0, 1, n).astype(F), "vy": rng.normal(0, 1, n).astype(F),
          "mass": rng.choice([1.0, 2.0, 3.0, 5.0], n).astype(F),
          "flags": np.where(rng.random(n) < 0.1, GE_FLAG_DARK, 0).astype(np.uint32),
          "id": rng.permutation(np.arange(7, 7 + 2 * n, 2)).astype(np.uint32)}
    rf = {"size": rng.uniform(0.05, 0.3, n).astype(F), "previous_size": rng.uniform(0.05, 0.3, n).astype(F),
          "neighbors": np.zeros(n, np.int32)}
    return st, rf


def couplings(a, cpu_seconds):
    """The steps either side of the frame and the couplings, each timed on the
    device (wall time around synchronised calls, best of several) and, where
    the oracle restates them in C, on the host on a bounded sample."""
    import numpy as np
    from galaxy_step import Context, configs, default_params
    from galaxy_step import constraints as GC
    from galaxy_step import merger as GM
    from galaxy_step import render as GR
    from galaxy_step import thermal as T

    def best(fn, reps, sync):
        fn()
        sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            sync()
            ts.append(time.perf_counter() - t0)
        return min(ts) * 1e3

    out = {}
    # render fields: neighbour counts, size and colour modes; PLY vertices
    st, prm = configs.config1(n=a.n, theta=a.theta)
    n = len(st["x"])
    p = default_params(G=prm["G"], theta=a.theta, softening=prm["softening"], dt=prm["dt"], dt_start=prm["dt"],
                       direct_below=0)
    with Context(n, sph=True, render=True) as c:  # neighbour counts run on the SPH grid
        c.upload(st)
        c.step(p, 1)
        c.sync()
        v = GR.default_visual_params(count_neighbors=1, density_size=1, density_color=1)
        ms = best(lambda: GR.visuals(c, v), 5, c.sync)
        out["visuals"] = {"workload": f"{n} particles of the two-disk frame: neighbour counts (3x3 cells of 3, "
                                      "radius 3.5), size by density, density colours",
                          "ms": ms, "particles_per_s": n / (ms * 1e-3)}
        ms = best(lambda: GR.ply_vertices(c), 3, c.sync)
        out["ply_vertices"] = {"workload": f"{n} particles: order-keeping compaction of the visible ones into "
                                           "20-byte PLY vertices on the device, copied to the host",
                               "ms": ms, "vertices_per_s": n / (ms * 1e-3)}
    # the galaxy spawner: 800k visible + 200k dark particles appended per call
    g = GR.default_galaxy_params(visible=800_000, dark=200_000)
    spawn = []
    for k in range(3):
        with Context(1_000_000, render=True) as c:
            GR.spawn_galaxy(c, GR.default_galaxy_params(visible=1000, dark=100), seed=99)
            c.sync()
            with Context(1_000_000, render=True) as c2:
                t0 = time.perf_counter()
                GR.spawn_galaxy(c2, g, seed=k)
                c2.sync()
                spawn.append(time.perf_counter() - t0)
    ms = min(spawn) * 1e3
    out["spawn_galaxy"] = {"workload": "1,000,000 particles (800k visible, 200k dark) appended with counter-based "
                                       "draws", "ms": ms, "particles_per_s": 1e6 / (ms * 1e-3)}
    # temperature coupling on the SPH impact: frames with and without it
    st, prm = configs.config2()
    n = len(st["x"])
    st["temp"] = np.random.default_rng(8).uniform(250.0, 1500.0, n).astype(np.float32)
    st["label"] = np.zeros(n, np.uint32)
    frames = {}
    for temp in (0, 1):
        pt = default_params(direct_below=0, compensated=1, temperature=temp)
        for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
                  "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
            setattr(pt, k, prm[k])
        pt.dt_start = pt.dt
        with Context(n, sph=True, temperature=True) as c:
            c.upload(st)
            T.thermal_upload(c, st)
            frames[temp] = best(lambda: c.step(pt, 5), 3, c.sync) / 5
    out["temperature"] = {"workload": f"gravity + SPH impact, {n} particles: node temperatures, heat transfer "
                                      "in the walk and temperatureCalculation each frame",
                          "ms_per_frame": frames[1], "ms_per_frame_without": frames[0],
                          "added_ms": frames[1] - frames[0]}
    # constraints: Physics::constraints on a 400 x 400 spring lattice
    st, cons = _lattice(400, 1)
    n = len(st["x"])
    pc = default_params(G=0.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, direct_below=0, constraints=1)
    with Context(n, sph=True, temperature=True, max_constraints=len(cons)) as c:
        c.upload(st)
        T.thermal_upload(c, st)
        GC.upload(c, cons)
        ms = best(lambda: GC.solve(c, pc), 5, c.sync)
        left = len(GC.download(c))
    out["constraints"] = {"workload": f"{n}-particle lattice, {len(cons)} springs: removal, 15 coloured substeps, "
                                      "breaking and plastic rules", "ms": ms,
                          "constraint_substeps_per_s": 15 * len(cons) / (ms * 1e-3), "constraints_left": left}
    # the merger on a crowded million
    # density 0.055 per unit area: well merged, but below percolation of the
    # overlap graph (its largest component has about a hundred particles)
    st, rf = _crowd(1_000_000, 3, 2130.0)
    n = len(st["x"])
    m = GM.default_merger_params()
    ts, removed = [], 0
    for _ in range(3):
        with Context(n, render=True) as c:
            c.upload(st)
            GR.render_upload(c, rf)
            c.sync()
            t0 = time.perf_counter()
            removed = GM.merge(c, m)
            c.sync()
            ts.append(time.perf_counter() - t0)
    ms = min(ts) * 1e3
    out["merger"] = {"workload": f"{n} crowded particles (10 % dark): UpdateNeighbors grid and the greedy "
                                 "merge loop", "ms": ms, "particles_per_s": n / (ms * 1e-3), "merged": removed}
    if cpu_seconds > 0:
        import oracle as O
        ss, srf = _crowd(100_000, 3, 2130.0 * (0.1 ** 0.5))
        t0 = time.perf_counter()
        O.merge(ss, srf["size"], srf["previous_size"], srf["neighbors"], **m.as_dict())
        el = time.perf_counter() - t0
        out["merger"]["cpu_baseline"] = {"particles_per_s": 1e5 / el, "cores": 1, "kind": "port",
                                         "sample": "100,000 particles at the same density through the serial C "
                                                   f"restatement, {el:.2f} s"}
        st2, cons2 = _lattice(100, 1)
        colors = np.arange(len(cons2), dtype=np.uint32)  # any order: a serial loop
        t0 = time.perf_counter()
        O.constraints(st2, cons2, colors, materials=O.reference_materials())
        el = time.perf_counter() - t0
        out["constraints"]["cpu_baseline"] = {"constraint_substeps_per_s": 15 * len(cons2) / el, "cores": 1,
                                              "kind": "port",
                                              "sample": f"100 x 100 lattice ({len(cons2)} springs), one solve through "
                                                        f"the serial C restatement, {el:.2f} s"}
    return out


def served(a, accuracy):
    """cfg4: Poisson request streams through the server, batched and plain."""
    import numpy as np
    from galaxy_step import configs, default_params
    from galaxy_step import serve as S
    pg = default_params(G=1.0, theta=0.7, softening=1.0, dt=0.01, dt_start=0.01, compensated=1,
                        accuracy=accuracy, direct_below=0, leaf_capacity=16)
    _, sp = configs.config2(n_target=10, n_impactor=10)
    ps = default_params(direct_below=0, compensated=1, accuracy=accuracy, leaf_capacity=16)
    for k in ("G", "theta", "softening", "dt", "sph", "sph_h", "sph_cell", "sph_unit_mass", "viscosity",
              "cohesion_coef", "delta", "stiff_mult", "sph_max_vel"):
        setattr(ps, k, sp[k])
    ps.dt_start = ps.dt
    # one request mix (sizes log-uniform in [2,000, 500,000], steps uniform in
    # [10, 200], 30 % gravity + SPH), replayed with each rate's arrivals
    rng = np.random.default_rng(4)
    mix = []
    for k in range(a.serve_requests):
        n = int(round(math.exp(rng.uniform(math.log(2000), math.log(500_000)))))
        steps = int(rng.integers(10, 201))
        kind = "sph" if rng.random() < 0.3 else "gravity"
        st, _ = configs.config4_request(n, kind, seed=int(rng.integers(0, 2 ** 31)))
        mix.append({"state": st, "steps": steps, "sph": int(kind == "sph")})
    results = []
    for rate in [float(r) for r in a.serve_rates.split(",")]:
        arr = np.cumsum(np.random.default_rng(int(rate) * 7 + 1).exponential(1.0 / rate, len(mix)))
        reqs = [dict(m, arrival_s=float(t)) for m, t in zip(mix, arr)]
        row = {"rate_per_s": rate, "requests": len(reqs),
               "sph_requests": sum(r["sph"] for r in reqs),
               "particle_updates": int(sum(len(r["state"]["x"]) * r["steps"] for r in reqs))}
        for form, batching in (("served", 1), ("plain", 0)):
            cfg = S.default_server_config(batching=batching, sph_slots=6, sph_capacity=1 << 19,
                                          batch_capacity=1 << 25, max_systems=512, round_frames=8, use_graph=1)
            with S.Server(cfg, pg, ps) as srv:
                _, res, stats = srv.run(reqs, ("x",))
            bad = sum(1 for r in res if r["status"] != 0)
            row[form] = {"p50_s": stats["p50_s"], "p99_s": stats["p99_s"], "mean_s": stats["mean_s"],
                         "solves_per_s": stats["solves_per_s"], "particle_updates_per_s": stats["updates_per_s"],
                         "wall_s": stats["wall_s"], "max_resident": stats["max_resident"],
                         "device_bytes": stats["device_bytes"], "failed": bad}
            log(f"served rate {rate}: {form} p50 {stats['p50_s']:.3f}s p99 {stats['p99_s']:.3f}s "
                f"{stats['solves_per_s']:.1f} solves/s")
        row["solves_speedup"] = row["served"]["solves_per_s"] / max(row["plain"]["solves_per_s"], 1e-12)
        results.append(row)
    return results


def main():
    a = parse()
    if a.calibrate_copy_mib:
        calibrate_copy(a.calibrate_copy_mib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(independent systems shard with no communication)", file=sys.stderr)
        sys.exit(2)
    from galaxy_step import configs

    if a.only_3d:
        print(json.dumps({"three_d": three_d(a, 0.0 if a.no_cpu else min(a.cpu_seconds, 10.0))}), flush=True)
        return
    if a.only_sph:
        print(json.dumps({"sph_impact": sph_impact(a)}), flush=True)
        return
    if a.only_sph3d:
        print(json.dumps({"sph3d_impact": sph3d(a, 0.0 if a.no_cpu else min(a.cpu_seconds, 8.0))}), flush=True)
        return
    if a.only_couplings:
        print(json.dumps({"couplings": couplings(a, 0.0 if a.no_cpu else a.cpu_seconds)}), flush=True)
        return
    st, prm = configs.config1(n=a.n, theta=a.theta)
    n = len(st["x"])
    log(f"n={n} theta={a.theta} accuracy={a.accuracy} warmup={a.warmup} steps={a.steps}")
    head = single_system(a, st, prm, a.accuracy, tune=True)
    log(f"headline: {head['ms_per_step']:.3f} ms/step (graph replay; eager {head['eager_ms_per_step']:.3f}), walk "
        f"{head['walk_ms']:.3f} ms, errors {head['errors']}, tune {head.get('autotune', {}).get('leaf_capacity')}")
    plain = single_system(a, st, prm, 0, tune=False) if a.accuracy != 0 else head
    log(f"plain: {plain['ms_per_step']:.3f} ms/step, errors {plain['errors']}")
    cpu = None
    if not a.no_cpu and a.cpu_seconds > 0:
        cpu = cpu_baseline(st, dict(prm, theta=a.theta), a.cpu_seconds)
        log(f"cpu: {cpu['value']:.3e} particle-updates/s")
    sph = None if a.no_sph else sph_impact(a)
    if sph:
        log(f"sph impact: {sph['ms_per_step']:.3f} ms/step, phases {sph['phase_ms']}")
    d3 = None if a.no_3d else three_d(a, 0.0 if a.no_cpu else min(a.cpu_seconds, 10.0))
    if d3:
        log(f"3d: {d3['ms_per_step']:.3f} ms/step, walk {d3['walk_ms']:.3f} ms ({d3['walk_tflops']:.1f} TFLOP/s), "
            f"errors {d3['force_error_vs_fp64']}, cpu {d3.get('cpu_baseline', {}).get('value')}")
    s3 = None if (a.no_sph or a.no_3d) else sph3d(a, 0.0 if a.no_cpu else min(a.cpu_seconds, 8.0))
    if s3:
        log(f"sph3d: {s3['ms_per_step']:.3f} ms/step, cpu {s3.get('cpu_baseline', {}).get('value')}")
    bat = None if a.no_batch else batched(a, a.accuracy)
    if bat:
        log(f"batched: {bat['batched_solves_per_s']:.1f} solves/s vs plain {bat['plain_solves_per_s']:.1f} "
            f"({bat['speedup']:.2f}x)")
    srv = None if a.no_serve else served(a, a.accuracy)
    cpl = None
    if not a.no_couplings:
        try:
            cpl = couplings(a, 0.0 if a.no_cpu else a.cpu_seconds)
            log("couplings: " + ", ".join(f"{k} {v.get('ms', v.get('ms_per_frame', 0.0)):.2f} ms"
                                          for k, v in cpl.items()))
        except Exception as e:  # an auxiliary leg: report it and keep the headline line
            cpl = {"error": f"{type(e).__name__}: {e}"}
            log(f"couplings leg failed: {cpl['error']}")
    phase_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in head["phases"].items() if v[1]}
    achieved = head["walk_tflops"]
    inter = head["interactions"]
    step = step_roofline(phase_ms, head["ms_per_step"], n, head["node_count"], inter["nodes"] + inter["pairs"],
                         head["params"]["accuracy"])
    out = {
        "metric": "particle-updates/s, 1M-body BH step theta=0.7; batched solves/s + p99 latency",
        "value": head["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-disk generator)",
        "config": {"workload": "2D Barnes-Hut frame, two-disk collision (1,000,000 particles)",
                   "particles": n, "theta": a.theta, "softening": prm["softening"], "dt": prm["dt"],
                   "G": prm["G"], "compensated": True, "accuracy": head["params"]["accuracy"],
                   "leaf_capacity": head["params"]["leaf_capacity"], "walk_block": head["params"]["walk_block"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles(head["params"]["accuracy"], head["params"]["walk_block"]),
                     "traffic_source": "profiles/bh_walk_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE per walk "
                                       "launch at this block size, calibrated on a copy kernel (not measured "
                                       "in this run)",
                     "kernel": ("k_bh_mp (accuracy-mode tree walk)" if head["params"]["accuracy"]
                                else "k_bh (reference tree walk)"),
                     "kernel_ms": head["walk_ms"],
                     "interactions": inter, "flops_per_interaction": FLOPS_PAIR,
                     "expression": {"achieved": head["walk_tflops_expression"],
                                    "frac": head["walk_tflops_expression"] / PEAK_FP32_TFLOPS,
                                    "flops_per_node": FLOPS_NODE[head["params"]["accuracy"]],
                                    "flops_per_pair": FLOPS_PAIR,
                                    "note": "flops of the expression the walk evaluates (a node to the "
                                            "hexadecapole is 107); not the roofline's algorithmic count"},
                     "step": step},
        "force_error_vs_fp64": head["errors"],
        "graph_ms_per_step": head["graph_ms_per_step"],
        "eager_ms_per_step": head["eager_ms_per_step"],
        "timing": "K frames replayed through ge_step's hipGraph capture (use_graph), bracketed by "
                  "device syncs; eager_ms_per_step is the same frames launched one by one; phase_ms comes "
                  "from a separate pass with HIP events around each phase",
        "host_transfer": {"upload_ms": head["upload_ms"], "download_ms": head["download_ms"],
                          "note": "one upload of all fields and one download of x, y, vx, vy, mass, id "
                                  "(pageable host memory); frames never cross PCIe"},
        "autotune": head.get("autotune"),
        "phase_ms": phase_ms,
        "plain_form": {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                       "force_error_vs_fp64": plain["errors"], "walk_ms": plain["walk_ms"],
                       "walk_tflops": plain["walk_tflops"], "walk_frac": plain["walk_tflops"] / PEAK_FP32_TFLOPS,
                       "note": "reference monopole walk, opening test and leaf rule (the parity path)"},
        "cpu_baseline": cpu,
        "sph_impact": sph,
        "three_d": d3,
        "sph3d_impact": s3,
        "couplings": cpl,
        "batched": bat,
        "served": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
