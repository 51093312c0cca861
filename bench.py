#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the 2D Barnes-Hut physics step on one
MI355X, on the 1,000,000-particle two-disk collision at theta = 0.7
(eps = 1, dt = 0.005, G = 1, synthetic seeded input).

A step is one full frame through the C ABI (ge_step): integrateStart +
bounding box, Morton keys, radix sort, gather, quadtree build, BH walk with
the fused integrateEnd kick.  The state stays resident in HBM; the timed
region covers exactly K frames between two device synchronisations.

The JSON line also carries the roofline of the dominant kernel (the BH walk:
VALU FP32 bound, 14 flops per interaction, interaction count measured by the
kernel itself) and a CPU baseline: the oracle's OpenMP restatement of the
reference step timed on this host on a bounded sample of the same workload.
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

FLOPS_PER_INTERACTION = 14          # dx, dy, r^2+eps^2, rsqrt, r^-3, G*M*r^-3, two accumulates
PEAK_FP32_TFLOPS = 157.3            # MI355X vector FP32
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--theta", type=float, default=0.7)
    ap.add_argument("--compensated", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="bound on the CPU-baseline sample (0 disables it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain elementwise copy of this many MiB (HBM counter calibration)")
    return ap.parse_args()


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


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "bh_walk_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(st, prm, budget_s):
    import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    s = O.State(st)
    p = dict(prm, dt_start=prm["dt"])
    t0 = time.perf_counter()
    steps = 0
    while True:
        s.step(p, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    n = len(st["x"])
    return {"value": n * steps / el, "unit": "particle-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} full frames of the same {n}-particle workload (theta={prm['theta']}) "
                      f"through the OpenMP C restatement of the reference step, {el:.1f} s"}


def main():
    a = parse()
    if a.calibrate_copy_mib:
        calibrate_copy(a.calibrate_copy_mib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(independent systems shard with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from galaxy_step import Context, configs, default_params

    st, prm = configs.config1(n=a.n, theta=a.theta)
    n = len(st["x"])
    p = default_params(G=prm["G"], theta=a.theta, softening=prm["softening"], dt=prm["dt"],
                       dt_start=prm["dt"], compensated=a.compensated)
    print(f"[bench] n={n} theta={a.theta} warmup={a.warmup} steps={a.steps}", file=sys.stderr, flush=True)
    with Context(n) as c:
        c.upload(st)
        c.step(p, a.warmup)
        c.sync()
        c.profile_enable(True)
        t0 = time.perf_counter()
        c.step(p, a.steps)
        c.sync()
        t1 = time.perf_counter()
        phases = c.profile_read()
        c.profile_enable(False)
        # interaction count of the current tree (the walk counts itself)
        c.compute_gravity(p)
        inter = c.interactions()
        nodes = c.node_count()
    el = t1 - t0
    ms_step = el * 1e3 / a.steps
    value = n * a.steps / el
    grav_ms, grav_launches = phases["gravity"]
    bh_s = grav_ms * 1e-3 / max(grav_launches, 1)
    achieved = FLOPS_PER_INTERACTION * inter / bh_s / 1e12
    phase_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in phases.items() if v[1]}
    print(f"[bench] {ms_step:.3f} ms/step, BH walk {bh_s * 1e3:.3f} ms, I/N={inter / n:.1f}, "
          f"nodes={nodes}, phases(ms/launch)={phase_ms}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu and a.cpu_seconds > 0:
        cpu = cpu_baseline(st, dict(prm, theta=a.theta), a.cpu_seconds)
    out = {
        "metric": "particle-updates/s, 1M-body BH step theta=0.7",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-disk generator)",
        "config": {"workload": "2D Barnes-Hut frame, two-disk collision",
                   "particles": n, "theta": a.theta, "softening": prm["softening"],
                   "dt": prm["dt"], "G": prm["G"], "compensated": bool(a.compensated),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles(),
                     "kernel": "k_bh (Barnes-Hut walk)", "kernel_ms": bh_s * 1e3,
                     "interactions": inter, "flops_per_interaction": FLOPS_PER_INTERACTION},
        "phase_ms": phase_ms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
