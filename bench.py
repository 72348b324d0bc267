#!/usr/bin/env python3
"""Throughput of the APIC step on one MI355X: particle-steps/s on the 1024^2
dam break (dx = 1/1024, 8 jittered particles per fluid cell, 2,005,736
particles, dt = 2.5e-4), the reference's own solver rules.

A step is one Apic::update(dt) (src/apic.rs:2429) over all particles, with
every array resident in HBM.  W untimed warm-up updates, then K timed updates
bracketed by device synchronisation; value = sum of particles x substeps / wall
seconds.  HIP events recorded by the library on its own stream time every
stage and the two transfer kernels inside the timed region; the dominant
kernel's achieved bandwidth is reported against the 8 TB/s HBM peak.  The CPU
baseline is the oracle (FP32, one thread) on the first steps of the same
workload, timed on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "apic-fluid-simulation_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 1024² grid, 8 ppc; served mixed-size p50/p99 step latency"
WORKLOAD = "cfg2_dam_break_1024"


def algorithmic_bytes(n_p, n_alloc, n_f, n_c):
    """Compulsory HBM bytes per launch of the two transfer kernels.

    g2p_kernel: per particle lane read x, y, time residual (12 B), write the 9
    particle fields (36 B) and the next sort record (8 B); per fluid cell read
    one u and one v face value (8 B) and their open flags (2 B).
    p2g_owner_kernel: per particle read the 9 fields (36 B); per grid cell
    write its u face and v face (4 + 4 B), their valid flags (1 + 1 B), the
    density (4 B) and the fluid flag (1 B) -- the reference rewrites every
    face and cell in this stage.
    """
    return {
        "kernel_g2p": 48 * n_alloc + 8 * n_p + 10 * n_f,
        "kernel_p2g": 36 * n_p + 15 * n_c,
    }


KERNEL_NAMES = {"kernel_g2p": "g2p_kernel", "kernel_p2g": "p2g_owner_kernel"}


def cpu_baseline(xy, cfg, budget_s):
    """The oracle (FP32 restatement of the reference step, one thread) on the
    first steps of the same workload; steps chosen to fill about budget_s."""
    from oracle import pyoracle as po
    o = po.Oracle(cfg["width"], cfg["height"], cfg["dx"], po.bench_props(), xy=xy)
    n = xy.shape[0]
    t0 = time.perf_counter()
    o.update(cfg["dt"])
    first = time.perf_counter() - t0
    steps = max(2, min(int(budget_s / max(first, 1e-3)), 200))
    total, substeps = first, o.stats()["substeps"]
    for _ in range(steps - 1):
        t0 = time.perf_counter()
        o.update(cfg["dt"])
        total += time.perf_counter() - t0
        substeps += o.stats()["substeps"]
    return {
        "value": n * substeps / total,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{WORKLOAD}: first {steps} updates of dt={cfg['dt']} from the seeded scene, "
                  f"oracle FP32 restatement, one thread, {platform.processor() or platform.machine()}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=15)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one simulation per GPU and no "
              "multi-GPU path exists (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions, WorldProperties

    cfg = scenes.CONFIGS[WORKLOAD]
    xy = scenes.config_scene(WORKLOAD)
    n_p = xy.shape[0]
    props = WorldProperties(-9.81, 0.1, 5.0, 1e-5)
    g = Apic.with_particles(cfg["width"], cfg["height"], cfg["dx"], props, xy)
    opts = DebugOptions(False)
    dt = cfg["dt"]
    for _ in range(args.warmup):
        g.update(opts, dt)
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    g.set_profiling(1)  # HIP events around the two transfer kernels only
    g.get_max_particle_velocity()  # synchronises the library's stream
    substeps, iters, fluid = 0, [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.update(opts, dt)
        s = g.stats()
        substeps += s["substeps"]
        iters.append(s["pcg_iterations_total"])
        fluid.append(s["fluid_cells"])
    g.get_max_particle_velocity()
    elapsed = time.perf_counter() - t0
    timers = g.timers()
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
    # stage breakdown from a short separate pass with events around every stage
    g.set_profiling(2)
    stage_steps = min(args.steps, 20)
    stage_substeps = 0
    for _ in range(stage_steps):
        g.update(opts, dt)
        stage_substeps += g.stats()["substeps"]
    stage_timers = g.timers()
    g.set_profiling(0)

    value = n_p * substeps / elapsed
    n_alloc = g.field_length("POS_X")
    n_f = float(np.mean(fluid))
    algo = algorithmic_bytes(n_p, n_alloc, n_f, cfg["width"] * cfg["height"])
    kernels = {}
    for k in ("kernel_g2p", "kernel_p2g"):
        ms, cnt = timers[k]
        avg_ms = ms / max(cnt, 1)
        kernels[k] = {"avg_ms": avg_ms, "launches": cnt, "bytes": algo[k],
                      "gbs": algo[k] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    d = kernels[dom]
    stages = {k: round(v[0] / max(stage_substeps, 1), 4) for k, v in stage_timers.items()
              if not k.startswith("kernel_")}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded dam break (SplitMix64 jitter), no dataset",
        "config": {"workload": WORKLOAD, "grid": [cfg["width"], cfg["height"]], "dx": cfg["dx"],
                   "particles": n_p, "ppc": cfg["ppc"], "dt": dt, "seed": cfg["seed"],
                   "solver": "reference rules (abs 1e-5, 100-iteration cap)", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": d["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (d["gbs"] / HBM_PEAK_GBS) if d["gbs"] else None, "traffic": None,
                     "bytes_per_launch": d["bytes"], "avg_launch_ms": d["avg_ms"]},
        "kernels": kernels,
        "stage_ms_per_step": stages,
        "pcg_iterations_mean": float(np.mean(iters)),
        "substeps": substeps,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(xy, cfg, args.cpu_seconds)
    g.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
