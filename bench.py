#!/usr/bin/env python3
"""Throughput of the APIC step on one MI355X: particle-steps/s on the 1024^2
dam break (dx = 1/1024, 8 jittered particles per fluid cell, 2,005,736
particles, dt = 2.5e-4, PCG tolerance 1e-6 relative: solver mode 1 with FP64
reductions and a 500-iteration cap).  The same workload under the
reference's own solver rules (absolute 1e-5 exits) is reported beside it.

A step is one Apic::update(dt) (src/apic.rs:2429) over all particles, with
every array resident in HBM.  W untimed warm-up updates, then K timed updates
bracketed by device synchronisation; value = sum of particles x substeps / wall
seconds.  HIP events recorded by the library on its own stream time the P2G
and G2P kernels, and one SpMV launch and one whole PCG iteration per solve
(iteration 1; events around all of them would add their own cost to the
step) inside the timed region; the unit with the largest share of the step
is the roofline unit, its achieved bandwidth (algorithmic bytes per launch /
average launch time) reported against the 8 TB/s HBM peak.  Every timed
solve's exit reason and residual are recorded, the FP64 true residual
||b - A p|| / ||b|| of the last one is computed on the host from the
library's own system and pressure, and the line says whether every solve
met the tolerance ("valid").  A short second pass times every stage.  The
CPU baseline is the oracle (FP32) on the first steps of the same workload,
timed on this host on all the CPUs the process may use (the threaded mode:
the reference's parallel decomposition) and on one.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
                       [--solver relative|reference]
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


def algorithmic_bytes(n_p, n_alloc, n_f, n_c, levels):
    """Compulsory HBM bytes per launch of the timed kernels, and per PCG
    iteration.

    G2p: per particle lane read x, y, time residual (12 B), write the 9
    particle fields (36 B); per particle the sort's permutation entry (4 B:
    the first substep reads its records through it, writing them in sorted
    order) and the next sort record (8 B); per fluid cell read one u and one v
    face value (8 B) and their open flags (2 B).
    P2G (P2gBlock and P2gSeam): per particle read the 9 fields (36 B) and its
    permutation entry (4 B: the records are read through the sort's
    permutation); per grid cell write its u face and v face (4 + 4 B), their
    valid flags (1 + 1 B), the density (4 B) and the fluid flag (1 B) -- the
    reference rewrites every face and cell in this stage.  (The blocks' halo
    records, 2 KB per 16 x 16 block written and read back, are overhead.)
    SpmvDot: per fluid cell read z, the old search vector, diag,
    plus_x, plus_y (20 B) and write the new search vector and A s (8 B); the
    fluid flag of every cell (1 B).
    PCG iteration (SpmvDot, the update and the V-cycle with the next r.z):
    68 B per base fluid cell (SpMV 24, update and residual norm 28, search
    update 16), 8 B per base cell (the dense r.z of src/apic.rs:1936), and
    per multigrid level l, base included, 89 B per fluid cell F_l (zeroing,
    two smoothing sweeps of 28 B, restriction 21, prolongation 12) and 4 B
    per cell C_l (its fluid flags as words).
    """
    return {
        "kernel_g2p": 48 * n_alloc + 12 * n_p + 10 * n_f,
        "kernel_p2g": 40 * n_p + 15 * n_c,
        "kernel_spmv": 28 * n_f + n_c,
        "kernel_pcg_iteration": 68 * n_f + 8 * n_c + sum(89 * f + 4 * c for f, c in levels),
    }


def stage_bytes(n_p, n_c, n_f, levels, iters, width, height):
    """Algorithmic bytes of each stage of one substep (compulsory traffic of
    the stage's inputs and outputs): the sort 16 N_p for the keys and 16 N_p
    per radix pass of 8 bits over 2 ceil(log2 max(W, H)) key bits (the
    records' gather into sorted order is not a pass of its own: P2G reads
    them through the permutation and G2P writes them in sorted order); P2G
    40 N_p + 14 N_c; the pressure system (three smoothing passes, assembly)
    92 N_f; the multigrid hierarchy 24 N_f; the PCG `iters` iterations of
    algorithmic_bytes' PCG iteration; the gradient 34 N_c; the four
    extrapolation passes 80 N_c; G2P 52 N_p + 8 N_c."""
    key_bits = 2 * max(1, (max(width, height) - 1).bit_length())
    passes = -(-key_bits // 8)
    it = 68 * n_f + 8 * n_c + sum(89 * f + 4 * c for f, c in levels)
    return {"sort": (16 + 16 * passes) * n_p, "p2g": 40 * n_p + 14 * n_c, "pressure_build": 92 * n_f,
            "mg_build": 24 * n_f, "pcg": iters * it, "gradient": 34 * n_c, "extrapolate": 80 * n_c,
            "g2p": 52 * n_p + 8 * n_c}


def stage_rates(stage_ms, n_p, n_c, n_f, levels, iters, width, height):
    """{stage: {"ms": per substep, "gbs": algorithmic GB/s, "frac": of the HBM peak}}."""
    sb = stage_bytes(n_p, n_c, n_f, levels, iters, width, height)
    out = {}
    for k, ms in stage_ms.items():
        if k in sb and ms > 0:
            gbs = sb[k] / (ms * 1e-3) / 1e9
            out[k] = {"ms": ms, "bytes": sb[k], "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


KERNEL_NAMES = {"kernel_g2p": "one_kernel<G2p<2> >", "kernel_p2g": "P2G (P2gBlock, P2gSeam)",
                "kernel_spmv": "one_kernel<SpmvDot<1024> >",
                "kernel_pcg_iteration": "PCG iteration (SpmvDot<1024>, VDown<32, 512, true>, 2 x VDown2, "
                                        "BottomLds, 2 x VUp2, VUp<32, 512>)"}
# the kernels of one PCG iteration of the 1024^2 workload and their launches
PCG_ITERATION_KERNELS = {"one_kernel<SpmvDot<1024> >": 1, "one_kernel<VDown<32, 512, true> >": 1,
                         "one_kernel<VDown2>": 2, "one_kernel<BottomLds>": 1, "one_kernel<VUp2>": 2,
                         "one_kernel<VUp<32, 512> >": 1}
# the launches of a unit timed as a whole
UNIT_KERNELS = {"kernel_p2g": {"one_kernel<P2gBlock>": 1, "one_kernel<P2gSeam>": 1},
                "kernel_pcg_iteration": PCG_ITERATION_KERNELS}
BATCH_WORKLOAD = "cfg4_batched_256"
LARGE_WORKLOAD = "cfg3_falling_block_4096x2048"


SOLVERS = {
    # the workload's "PCG tol 1e-6": relative 2-norm, FP64 reductions, 500-iteration cap
    "relative": dict(mode=1, rel_tol=1e-6, max_iterations=500, fp64_reductions=True, recompute_every=25),
    # the reference's own rules (absolute 1e-5 exits, 100-iteration cap)
    "reference": dict(mode=0),
}
SOLVER_NAMES = {"relative": "relative 1e-6 (FP64 reductions, residual recompute every 25, cap 500)",
                "reference": "reference rules (absolute 1e-5, 100-iteration cap)"}


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def baseline_threads():
    """Threads for the all-cores CPU baseline: the CPUs this process may use,
    capped by OMP_NUM_THREADS when the environment sets it."""
    _, _, usable = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return min(usable, int(env))
    return usable


def cpu_baseline(xy, cfg, budget_s, solver, threads):
    """The oracle, the FP32 restatement of the reference step, on the first
    steps of the same workload, timed on this host.  threads > 1 runs its
    threaded mode: the reference's rayon loops as OpenMP loops over the same
    units (8-particle chunks with the reference's minimum lengths, CAS
    scatters in P2G, chunked reductions); one thread is the serial
    restatement.  Steps are chosen to fill about budget_s."""
    from oracle import pyoracle as po
    o = po.Oracle(cfg["width"], cfg["height"], cfg["dx"], po.bench_props(), xy=xy)
    o.set_threads(threads)
    sv = SOLVERS[solver]
    if sv["mode"] == 1:
        o.set_solver(1, sv["rel_tol"], sv["max_iterations"], sv["fp64_reductions"], sv["recompute_every"])
    n = xy.shape[0]
    t0 = time.perf_counter()
    o.update(cfg["dt"])
    first = time.perf_counter() - t0
    steps = max(2, min(int(budget_s / max(first, 1e-3)), cfg.get("steps", 200)))
    total, substeps = first, o.stats()["substeps"]
    for _ in range(steps - 1):
        t0 = time.perf_counter()
        o.update(cfg["dt"])
        total += time.perf_counter() - t0
        substeps += o.stats()["substeps"]
    model, logical, usable = host_cpu()
    return {
        "value": n * substeps / total,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "ms_per_step": total / steps * 1e3,
        "cpu_model": model,
        "logical_cpus": logical,
        "usable_cpus": usable,
        "sample": f"{WORKLOAD}: first {steps} of its {cfg.get('steps', 200)} updates of dt={cfg['dt']} from the seeded "
                  f"scene, solver {SOLVER_NAMES[solver]}; the oracle (FP32 restatement of the reference step) "
                  + ("serial, one thread" if threads == 1 else
                     f"threaded on {threads} OpenMP threads with the reference's parallel decomposition")
                  + f", on {model}",
    }


def true_residual(g):
    """||b - A p||_2 / ||b||_2 of the last pressure solve, in FP64 on the host,
    for the FP64-accumulated pressure and for the FP32 pressure the step
    uses: b is the relative mode's copy of the right-hand side, A the
    assembled 5-point operator (DIAG, PLUS_X, PLUS_Y) over the fluid cells."""
    import numpy as np
    W, H = g.width, g.height
    b = g.read("RHS").astype(np.float64)
    diag = g.read("DIAG").astype(np.float64)
    px = g.read("PLUS_X").astype(np.float64)
    py = g.read("PLUS_Y").astype(np.float64)
    idx = np.nonzero(g.read("FLUID"))[0] + W
    nb = np.linalg.norm(b[idx])
    out = {}
    for name, p in (("fp64_pressure", g.read("PRESSURE_F64")), ("fp32_pressure", g.read("PRESSURE"))):
        pp = np.zeros(W * H + 2 * W)
        pp[W:W + W * H] = p.astype(np.float64)
        ap = diag[idx] * pp[idx] + px[idx] * pp[idx + 1] + px[idx - 1] * pp[idx - 1] + \
            py[idx] * pp[idx + W] + py[idx - W] * pp[idx - W]
        out[name] = float(np.linalg.norm(b[idx] - ap) / nb) if nb > 0 else 0.0
    return out


def level_counts(g):
    """(fluid cells, cells) of every multigrid level, base grid first."""
    out = []
    for lvl in range(g.dims()["num_levels"] + 1):
        fl = g.read("FLUID", lvl)
        out.append((int(fl.sum()), int(fl.size)))
    return out


def timed_run(Apic, DebugOptions, props, cfg, xy, solver, steps, warmup, stage_steps, verify=True):
    """W warm-up updates, K timed updates (HIP events around the two transfer
    kernels, one SpMV and one whole PCG iteration per solve only), the exit
    state of every timed solve, the true residual of the last one, then a
    short pass with events around every stage."""
    import numpy as np
    g = Apic.with_particles(cfg["width"], cfg["height"], cfg["dx"], props, xy)
    g.set_solver_config(**SOLVERS[solver])
    opts = DebugOptions(False)
    dt = cfg["dt"]
    for _ in range(warmup):
        g.update(opts, dt)
    # launch shapes for this grid (radix digit width, multigrid tile side and
    # level pairing), timed on the warmed-up state; they change no result
    tuning = g.autotune(reps=5)
    print(f"[bench] autotuned: radix_bits {tuning['radix_bits']}, few_tiles {tuning['few_tiles']} "
          f"(sort {tuning['sort_ms']:.3f} ms, V-cycle {tuning['vcycle_ms']:.3f} ms)", file=sys.stderr, flush=True)
    g.set_profiling(1)  # HIP events around the timed kernels only
    g.get_max_particle_velocity()  # synchronises the library's stream
    substeps, iters, fluid, rel, reasons = 0, [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        g.update(opts, dt)
        st = g.stats()
        substeps += st["substeps"]
        iters.append(st["pcg_iterations_total"])
        fluid.append(st["fluid_cells"])
        rel.append(st["pcg_rel_residual"])
        reasons.append(st["pcg_exit_reason"])
    g.get_max_particle_velocity()
    elapsed = time.perf_counter() - t0
    timers = g.timers()
    check = true_residual(g) if verify and SOLVERS[solver]["mode"] == 1 else None
    levels = level_counts(g)
    g.set_profiling(2)
    stage_substeps = 0
    for _ in range(stage_steps):
        g.update(opts, dt)
        stage_substeps += g.stats()["substeps"]
    stage_timers = g.timers()
    g.set_profiling(0)
    n_alloc = g.field_length("POS_X")
    g.close()
    return dict(elapsed=elapsed, substeps=substeps, iters=iters, fluid=fluid, rel=rel, reasons=reasons,
                timers=timers, stage_timers=stage_timers, stage_substeps=stage_substeps, n_alloc=n_alloc,
                true_residual=check, levels=levels, tuning=tuning)


def batched_run(count=256, side_min=64, side_max=512, seed=4, steps=10, warmup=3, seq_rounds=2, dt=0.002):
    """Independent dam breaks of mixed sizes stepped together by one batch
    update per step, against the same simulations stepped one at a time with
    Apic.update (each update finished before the next simulation starts).
    Solver: relative 1e-6 with FP64 reductions.  Per step: wall time of the
    update plus a stream synchronisation; value = sum of particles x substeps
    / the summed step times."""
    import numpy as np
    from FLIP import scenes
    from FLIP.apic import Apic, Batch, DebugOptions, WorldProperties
    t_setup = time.perf_counter()
    mix = scenes.instance_mix(count, side_min, side_max, seed)
    n_p = np.array([m["xy"].shape[0] for m in mix], dtype=np.int64)
    print(f"[bench] batched workload: {count} simulations, {int(n_p.sum())} particles "
          f"({time.perf_counter() - t_setup:.1f}s to seed)", file=sys.stderr, flush=True)

    def make(create, m):
        s = create(m["width"], m["height"], m["dx"], WorldProperties(m["gravity"], 0.1, 5.0, 1e-5), m["xy"])
        s.set_gravity_x(m["gravity_x"])
        s.set_solver_config(**SOLVERS["relative"])
        return s

    opts = DebugOptions(False)
    batch = Batch()
    t0 = time.perf_counter()
    sims = []
    for i, m in enumerate(mix):
        sims.append(make(batch.create_sim, m))
        if (i + 1) % 64 == 0 or i + 1 == count:
            print(f"[bench] created {i + 1}/{count} ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    for w in range(warmup):
        t0 = time.perf_counter()
        batch.update(sims, opts, dt)
        sims[0].get_max_particle_velocity()  # synchronises the batch's stream
        print(f"[bench] batched warm-up step {w}: {time.perf_counter() - t0:.3f}s", file=sys.stderr, flush=True)
    lat, ps, iters = [], 0, []
    for k in range(steps):
        t0 = time.perf_counter()
        batch.update(sims, opts, dt)
        sims[0].get_max_particle_velocity()
        lat.append(time.perf_counter() - t0)
        st = [s.stats() for s in sims]
        ps += int(sum(n * x["substeps"] for n, x in zip(n_p, st)))
        iters.append(max(x["pcg_iterations_total"] for x in st))
        print(f"[bench] batched step {k}: {lat[-1] * 1e3:.1f} ms, most PCG iterations {iters[-1]}",
              file=sys.stderr, flush=True)
    b_elapsed = float(sum(lat))
    for s in sims:
        s.close()
    batch.close()
    print(f"[bench] batched: {count} simulations, {steps} steps in {b_elapsed:.3f}s", file=sys.stderr, flush=True)

    alone = [make(lambda w, h, dx, wp, xy: Apic.with_particles(w, h, dx, wp, xy), m) for m in mix]
    for s in alone:
        s.update(opts, dt)
        s.get_max_particle_velocity()
    print(f"[bench] one at a time: created and warmed up", file=sys.stderr, flush=True)
    s_elapsed, s_ps = 0.0, 0
    for r in range(seq_rounds):
        for s, n in zip(alone, n_p):
            t0 = time.perf_counter()
            s.update(opts, dt)
            s.get_max_particle_velocity()  # this update has finished
            s_elapsed += time.perf_counter() - t0
            s_ps += int(n) * s.stats()["substeps"]
        print(f"[bench] one at a time: round {r} done ({s_elapsed:.2f}s so far)", file=sys.stderr, flush=True)
    for s in alone:
        s.close()
    print(f"[bench] one at a time: {seq_rounds} rounds in {s_elapsed:.3f}s", file=sys.stderr, flush=True)
    lat_ms = np.array(lat) * 1e3
    b_value, s_value = ps / b_elapsed, s_ps / s_elapsed
    sides = [m["width"] for m in mix]
    return {
        "workload": BATCH_WORKLOAD,
        "instances": count,
        "particles": int(n_p.sum()),
        "sides": [int(min(sides)), int(max(sides))],
        "dt": dt,
        "solver": SOLVER_NAMES["relative"],
        "batched": {"value": b_value, "unit": "particle-steps/s", "steps": steps, "warmup": warmup,
                    "ms_per_step": float(lat_ms.mean()), "p50_ms": float(np.percentile(lat_ms, 50)),
                    "p99_ms": float(np.percentile(lat_ms, 99)), "pcg_iterations_max_mean": float(np.mean(iters))},
        "one_at_a_time": {"value": s_value, "unit": "particle-steps/s", "rounds": seq_rounds,
                          "ms_per_round": s_elapsed / seq_rounds * 1e3},
        "speedup": b_value / s_value,
    }


def served_run(pool=2048, n_requests=400, rates=(200.0, 800.0, 3200.0), seed=5, side_min=64, side_max=1024,
               dt=0.002):
    """The served workload: requests of 1-20 updates for simulations of a
    resident pool (sides 64-1024, 4-8 ppc, relative 1e-6 solver) arriving as a
    Poisson process, replayed in real time through the continuous-batching
    server and, for the same trace, through plain Apic.update one request at a
    time.  Per arrival rate: aggregate particle-steps/s (first arrival to last
    completion) and request latency (completion - arrival) percentiles.  At the
    lower rates the arrivals bound both; the highest one arrives faster than
    either path can serve, so it measures the throughput each sustains."""
    import copy

    import numpy as np
    from FLIP import scenes
    from FLIP.apic import Apic, WorldProperties
    from FLIP.serve import Server, SimSpec, poisson_trace, run_one_at_a_time, run_served
    t0 = time.perf_counter()
    mix = scenes.instance_mix(pool, side_min, side_max, seed, progress=lambda i: print(
        f"[bench] served workload: seeded {i}/{pool} ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True))
    specs = [SimSpec(m["width"], m["height"], m["dx"], m["xy"], gravity=m["gravity"], gravity_x=m["gravity_x"],
                     solver=SOLVERS["relative"]) for m in mix]
    n_particles = int(sum(sp.xy.shape[0] for sp in specs))
    print(f"[bench] served workload: pool of {pool} simulations, {n_particles} particles "
          f"({time.perf_counter() - t0:.1f}s to seed)", file=sys.stderr, flush=True)
    out = []
    for rate in rates:
        trace = poisson_trace(n_requests, rate, pool, seed * 1000 + int(rate))
        base_trace = copy.deepcopy(trace)
        server = Server(specs, frame_dt=dt)
        for k in range(0, pool, 256):
            server.preload(range(k, min(pool, k + 256)))
            print(f"[bench] served at {rate} req/s: preloaded {len(server.resident)} of {min(pool, k + 256)} "
                  f"({server.used / 2**30:.1f} GiB)", file=sys.stderr, flush=True)
        wall = run_served(server, trace)
        lat = np.array([r.t_done - r.t_arrival for r in trace]) * 1e3
        served = {"value": server.particle_steps / wall, "unit": "particle-steps/s", "wall_s": wall,
                  "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                  "ticks": server.ticks, "mean_sims_per_tick": server.stepped / max(server.ticks, 1),
                  "offloads": server.offloads}
        server.close()
        print(f"[bench] served at {rate} req/s: {wall:.2f}s, p50 {served['p50_ms']:.1f} ms, "
              f"p99 {served['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
        # one at a time: the simulations the trace asks for, created beforehand
        sims = {}
        for sid in sorted({r.sim for r in base_trace}):
            sp = specs[sid]
            s = Apic.with_particles(sp.width, sp.height, sp.dx,
                                    WorldProperties(sp.gravity, sp.border_damping, sp.cfl, sp.solve_error), sp.xy)
            s.set_gravity_x(sp.gravity_x)
            s.set_solver_config(**sp.solver)
            sims[sid] = s
        print(f"[bench] one at a time at {rate} req/s: {len(sims)} simulations created", file=sys.stderr,
              flush=True)
        wall2, ps2 = run_one_at_a_time(sims, specs, base_trace, frame_dt=dt)
        for s in sims.values():
            s.close()
        lat2 = np.array([r.t_done - r.t_arrival for r in base_trace]) * 1e3
        alone = {"value": ps2 / wall2, "unit": "particle-steps/s", "wall_s": wall2,
                 "p50_ms": float(np.percentile(lat2, 50)), "p99_ms": float(np.percentile(lat2, 99))}
        print(f"[bench] one at a time at {rate} req/s: {wall2:.2f}s, p50 {alone['p50_ms']:.1f} ms, "
              f"p99 {alone['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
        out.append({"rate_per_s": rate, "requests": n_requests, "served": served, "one_at_a_time": alone,
                    "throughput_ratio": served["value"] / alone["value"]})
    return {"workload": "cfg5_served_pool", "pool": pool, "particles": n_particles,
            "sides": [int(min(sp.width for sp in specs)), int(max(sp.width for sp in specs))], "dt": dt,
            "steps_per_request": [1, 20], "solver": SOLVER_NAMES["relative"], "rates": out}


def harness_run(sizes=(64, 128, 256, 512, 1024), warmup=15, iters=20, dt=0.016):
    """The reference's own benchmark (benches/fluid_benchmark.rs:10-39):
    Apic::new(size, size, 1.0, props), the constructor's particle seeding,
    under the reference's solver rules; `warmup` updates of dt = 0.016, then
    `iters` timed updates (each followed by a stream synchronisation, as the
    harness times one update call at a time).  Per size: ms per update and
    particle-steps/s.  At cell size 1.0 the pressure solve of these updates
    exits at its first test (r.z below 1e-5), in the oracle as on the GPU."""
    from FLIP.apic import Apic, DebugOptions, WorldProperties
    props = WorldProperties(-9.81, 0.1, 5.0, 1e-5)
    opts = DebugOptions(False)
    out = []
    for size in sizes:
        g = Apic.new(size, size, 1.0, props)
        n = g.dims()["num_particles"]
        for _ in range(warmup):
            g.update(opts, dt)
        g.get_max_particle_velocity()
        t0 = time.perf_counter()
        subs, its = 0, 0
        for _ in range(iters):
            g.update(opts, dt)
            g.get_max_particle_velocity()
            st = g.stats()
            subs += st["substeps"]
            its += st["pcg_iterations_total"]
        el = time.perf_counter() - t0
        g.close()
        out.append({"size": size, "particles": n, "ms_per_update": el / iters * 1e3,
                    "value": n * subs / el, "unit": "particle-steps/s", "substeps_mean": subs / iters,
                    "pcg_iterations_mean": its / iters})
        print(f"[bench] harness {size}^2: {n} particles, {el / iters * 1e3:.3f} ms per update",
              file=sys.stderr, flush=True)
    return {"workload": "reference_harness", "solver": SOLVER_NAMES["reference"], "warmup": warmup,
            "iters": iters, "dt": dt, "sizes": out}


def large_run(steps=50, warmup=3, stage_steps=5, oracle_steps=2):
    """The mixed-precision large case (a 4096 x 2048 falling block, 35.6M
    particles): FP32 transfers and SpMV (FP64 row sums), FP64 PCG reductions,
    residual recompute every 25 iterations, the relative 1e-6 test in FP64.
    `steps` timed updates from the seeded scene after `warmup`, per-stage
    algorithmic GB/s from a pass with every stage timed, and the error
    against the FP64 oracle (pressure, grid velocity, particle positions)
    after the first `oracle_steps` updates from the seeded scene, with the
    FP32 oracle's own distance from the FP64 oracle beside it."""
    import numpy as np
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions, WorldProperties
    name = LARGE_WORKLOAD
    c = scenes.CONFIGS[name]
    W, H, dx, dt = c["width"], c["height"], c["dx"], c["dt"]
    t0 = time.perf_counter()
    xy = scenes.config_scene(name)
    n_p = xy.shape[0]
    print(f"[bench] large workload: {n_p} particles ({time.perf_counter() - t0:.1f}s to seed)",
          file=sys.stderr, flush=True)
    props = WorldProperties(-9.81, 0.1, 5.0, 1e-5)
    r = timed_run(Apic, DebugOptions, props, c, xy, "relative", steps, warmup, stage_steps)
    print(f"[bench] large workload: {steps} steps in {r['elapsed']:.3f}s", file=sys.stderr, flush=True)
    n_f = float(np.mean(r["fluid"]))
    stage_ms = {k: v[0] / max(r["stage_substeps"], 1) for k, v in r["stage_timers"].items()
                if not k.startswith("kernel_")}
    iters = float(np.mean(r["iters"]))
    rates = stage_rates(stage_ms, n_p, W * H, n_f, r["levels"], iters, W, H)
    tr64 = (r["true_residual"] or {}).get("fp64_pressure")
    out = {"workload": name, "grid": [W, H], "dx": dx, "dt": dt, "particles": n_p, "ppc": c["ppc"],
           "seed": c["seed"], "solver": SOLVER_NAMES["relative"], "steps": steps, "warmup": warmup,
           "value": n_p * r["substeps"] / r["elapsed"], "unit": "particle-steps/s",
           "ms_per_step": r["elapsed"] / steps * 1e3, "pcg_iterations_mean": iters,
           "pcg_rel_residual_max": max(x for x in r["rel"] if x >= 0),
           "pcg_exit_reasons": {str(k): int(sum(1 for x in r["reasons"] if x == k)) for k in sorted(set(r["reasons"]))},
           "pcg_true_rel_residual_last": r["true_residual"],
           "valid": all(x == 3 for x in r["reasons"]) and tr64 is not None and tr64 <= 1e-6 * (1 + 1e-6),
           "stages": rates, "autotune": r["tuning"]}
    if oracle_steps > 0:
        out["fp64_oracle_error"] = fp64_oracle_error(c, xy, props, oracle_steps)
    return out


def fp64_oracle_error(c, xy, props, nsteps):
    """Relative L2 distance from the FP64 oracle after each of the first
    `nsteps` updates of the seeded scene (the GPU, and the FP32 oracle, which
    is the reference's own precision), under the relative solver."""
    import numpy as np
    from FLIP.apic import Apic, DebugOptions
    from oracle import pyoracle as po
    W, H, dx, dt = c["width"], c["height"], c["dx"], c["dt"]
    g = Apic.with_particles(W, H, dx, props, xy)
    g.set_solver_config(**SOLVERS["relative"])
    orc = {}
    for prec in (64, 32):
        o = po.Oracle(W, H, dx, po.bench_props(), xy=xy, precision=prec)
        o.set_solver(1, 1e-6, 500, True, 25)
        o.set_threads(baseline_threads())  # (threaded: the reference's own summation freedom)
        orc[prec] = o

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        d = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))

    per_step = []
    for k in range(nsteps):
        t0 = time.perf_counter()
        g.update(DebugOptions(False), dt)
        for o in orc.values():
            o.update(dt)
        e = {"step": k + 1}
        for who, src in (("gpu", g), ("fp32_oracle", orc[32])):
            e[who] = {"pressure": rel(src.read("PRESSURE"), orc[64].read("PRESSURE")),
                      "grid_velocity_u": rel(src.read("MAC_U"), orc[64].read("MAC_U")),
                      "grid_velocity_v": rel(src.read("MAC_V"), orc[64].read("MAC_V")),
                      "positions": max(rel(src.read(f), orc[64].read(f)) for f in ("POS_X", "POS_Y"))}
        e["pcg_iterations"] = {"gpu": g.stats()["pcg_iterations"], "fp64_oracle": orc[64].stats()["pcg_iterations"]}
        per_step.append(e)
        print(f"[bench] large workload, FP64 oracle step {k + 1}/{nsteps}: GPU error {e['gpu']} "
              f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    g.close()
    return {"steps_compared": nsteps, "of_steps": c["steps"], "per_step": per_step}


def copy_calibration(gib=1.0, repeats=3):
    """A plain device-to-device copy of known size (the runtime's copy kernel,
    through torch), for scaling the FETCH_SIZE and WRITE_SIZE counters: run
    under rocprofv3 --pmc, its dispatches read and write `gib` GiB each, far
    more than the L2 and Infinity Cache hold."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / repeats
    print(json.dumps({"copy_bytes": 4 * n, "repeats": repeats, "ms": dt * 1e3,
                      "gbs": 2 * 4 * n / dt / 1e9}))


def traffic_summary(unit):
    """HBM bytes per launch of `unit` (a kernel name, or the PCG iteration)
    from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, each
    divided by what the counter reads per byte of a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_dam_break_1024.json")
    try:
        with open(path) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        return None
    per = tr.get("hbm_bytes_per_launch")
    if not isinstance(per, dict):
        return None
    if unit in UNIT_KERNELS:
        ks = UNIT_KERNELS[unit]
        if not all(k in per for k in ks):
            return None
        return int(sum(per[k] * n for k, n in ks.items()))
    v = per.get(KERNEL_NAMES[unit])
    return int(v) if v is not None else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=15)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solver", choices=sorted(SOLVERS), default="relative")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched mixed-size workload")
    ap.add_argument("--batch-only", action="store_true", help="print only the batched workload's result")
    ap.add_argument("--batch-count", type=int, default=256, help="simulations in the batched workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the served workload")
    ap.add_argument("--serve-only", action="store_true", help="print only the served workload's result")
    ap.add_argument("--serve-pool", type=int, default=2048, help="simulations in the served pool")
    ap.add_argument("--serve-rates", type=lambda v: tuple(float(x) for x in v.split(",")),
                    default=(200.0, 800.0, 3200.0), help="arrival rates of the served workload (req/s)")
    ap.add_argument("--no-large", action="store_true", help="skip the 4096 x 2048 mixed-precision workload")
    ap.add_argument("--large-only", action="store_true", help="print only the large workload's result")
    ap.add_argument("--large-oracle-steps", type=int, default=2,
                    help="updates of the large workload compared with the FP64 oracle")
    ap.add_argument("--no-harness", action="store_true", help="skip the reference harness's workload")
    ap.add_argument("--harness-only", action="store_true", help="print only the reference harness's result")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 1 GiB device copy (for scaling the HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one simulation per GPU and no "
              "multi-GPU path exists (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions, WorldProperties

    if args.copy_calibration:
        copy_calibration()
        return
    if args.batch_only:
        print(json.dumps(batched_run(count=args.batch_count)))
        return
    if args.large_only:
        print(json.dumps(large_run(oracle_steps=args.large_oracle_steps)))
        return
    if args.serve_only:
        print(json.dumps(served_run(pool=args.serve_pool, rates=args.serve_rates)))
        return
    if args.harness_only:
        print(json.dumps(harness_run()))
        return
    cfg = scenes.CONFIGS[WORKLOAD]
    xy = scenes.config_scene(WORKLOAD)
    n_p = xy.shape[0]
    props = WorldProperties(-9.81, 0.1, 5.0, 1e-5)
    r = timed_run(Apic, DebugOptions, props, cfg, xy, args.solver, args.steps, args.warmup, min(args.steps, 20))
    print(f"[bench] timed {args.steps} steps in {r['elapsed']:.3f}s ({args.solver} solver)", file=sys.stderr,
          flush=True)
    other = "reference" if args.solver == "relative" else "relative"
    o_steps = min(args.steps, 50)
    ro = timed_run(Apic, DebugOptions, props, cfg, xy, other, o_steps, args.warmup, 0)
    print(f"[bench] {other} solver: {o_steps} steps in {ro['elapsed']:.3f}s", file=sys.stderr, flush=True)

    value = n_p * r["substeps"] / r["elapsed"]
    n_f = float(np.mean(r["fluid"]))
    algo = algorithmic_bytes(n_p, r["n_alloc"], n_f, cfg["width"] * cfg["height"], r["levels"])
    kernels = {}
    # the SpMV and PCG-iteration timers sample one iteration per solve
    # (iteration 1): their share of the step is their average times the
    # iterations per step
    per_step_iters = float(np.sum(r["iters"])) / args.steps
    for k in ("kernel_g2p", "kernel_p2g", "kernel_spmv", "kernel_pcg_iteration"):
        ms, cnt = r["timers"][k]
        avg_ms = ms / max(cnt, 1)
        per_step = avg_ms * per_step_iters if k in ("kernel_spmv", "kernel_pcg_iteration") else ms / args.steps
        kernels[k] = {"avg_ms": avg_ms, "launches": cnt, "ms_per_step": per_step, "bytes": algo[k],
                      "gbs": algo[k] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None}
    # the dominant unit: the largest share of the step among the timed ones
    # (the SpMV is part of the PCG iteration and is not a candidate)
    dom = max(("kernel_g2p", "kernel_p2g", "kernel_pcg_iteration"), key=lambda k: kernels[k]["ms_per_step"])
    d = kernels[dom]
    stages = {k: round(v[0] / max(r["stage_substeps"], 1), 4) for k, v in r["stage_timers"].items()
              if not k.startswith("kernel_")}
    stage_gbs = stage_rates(stages, n_p, cfg["width"] * cfg["height"], n_f, r["levels"],
                            float(np.mean(r["iters"])), cfg["width"], cfg["height"])
    rels = [x for x in r["rel"] if x >= 0]
    tol = SOLVERS[args.solver].get("rel_tol")
    # every timed solve ended by the tolerance, and the last one's FP64
    # pressure meets it in the true residual evaluated on the host (up to
    # the FP64 rounding of that evaluation)
    tr64 = (r["true_residual"] or {}).get("fp64_pressure")
    valid = all(x == 3 for x in r["reasons"]) and (tol is None or (
        all(x <= tol for x in rels) and tr64 is not None and tr64 <= tol * (1 + 1e-6)))
    traffic = traffic_summary(dom)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded dam break (SplitMix64 jitter), no dataset",
        "config": {"workload": WORKLOAD, "grid": [cfg["width"], cfg["height"]], "dx": cfg["dx"],
                   "particles": n_p, "ppc": cfg["ppc"], "dt": cfg["dt"], "seed": cfg["seed"],
                   "solver": SOLVER_NAMES[args.solver], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": d["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (d["gbs"] / HBM_PEAK_GBS) if d["gbs"] else None, "traffic": traffic,
                     "bytes_per_launch": d["bytes"], "avg_launch_ms": d["avg_ms"],
                     "levels_fluid_cells": [f for f, _ in r["levels"]],
                     "levels_cells": [c for _, c in r["levels"]]},
        "kernels": kernels,
        "stage_ms_per_step": stages,
        "stage_algorithmic_gbs": stage_gbs,
        "pcg_iterations_mean": float(np.mean(r["iters"])),
        "pcg_rel_residual_max": max(rels) if rels else None,
        "pcg_exit_reasons": {str(k): int(sum(1 for x in r["reasons"] if x == k)) for k in sorted(set(r["reasons"]))},
        "pcg_true_rel_residual_last": r["true_residual"],
        "valid": valid,
        "substeps": r["substeps"],
        "autotune": r["tuning"],
        "other_solver": {"solver": SOLVER_NAMES[other], "steps": o_steps,
                         "value": n_p * ro["substeps"] / ro["elapsed"],
                         "ms_per_step": ro["elapsed"] / o_steps * 1e3,
                         "pcg_iterations_mean": float(np.mean(ro["iters"]))},
    }
    if not args.no_harness:
        line["harness"] = harness_run()
    if not args.no_large:
        line["large"] = large_run(oracle_steps=0 if args.no_cpu else args.large_oracle_steps)
    if not args.no_batch:
        line["batched"] = batched_run(count=args.batch_count)
    if not args.no_serve:
        line["served"] = served_run(pool=args.serve_pool, rates=args.serve_rates)
    if not args.no_cpu:
        threads = baseline_threads()
        line["cpu_baseline"] = cpu_baseline(xy, cfg, args.cpu_seconds, args.solver, threads)
        print(f"[bench] CPU baseline on {threads} threads: {line['cpu_baseline']['value']:.3e} particle-steps/s",
              file=sys.stderr, flush=True)
        if threads > 1:
            line["cpu_baseline_one_core"] = cpu_baseline(xy, cfg, args.cpu_seconds, args.solver, 1)
            print(f"[bench] CPU baseline on 1 thread: {line['cpu_baseline_one_core']['value']:.3e} particle-steps/s",
                  file=sys.stderr, flush=True)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
