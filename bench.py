#!/usr/bin/env python3
"""Benchmark of the smoke time step on one MI355X.

Default workload (grid_256): 256^3 cells, FP32, floor + seeded boxes covering
5 % of cells, 4 seeded density sources, velocity noise +-0.01, net buoyancy
+9.8/density with a 0.001 density gate, clamp M = 4, dt = 0.05, and a
pressure projection of 100 Jacobi sweeps.  Inputs are resident in HBM before
the timed region; the timed region is K full steps (sources, mask, buoyancy,
clamp, divergence, 100 sweeps, gradient, velocity and density advection)
enqueued back to back and bracketed by stream synchronisations.

Prints ONE JSON line on stdout (metric, value, roofline, cpu_baseline, ...);
progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "smoke-simulation_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E peak, GB/s
# Algorithmic bytes per cell (unfused streaming model: every stage or sweep
# reads its inputs once and writes its outputs once, FP32 fields, 1-byte flags).
BYTES_FORCE = 29               # rho, u, v, w, flags in; u, v, w out
BYTES_DIVERGENCE = 17          # u, v, w, flags in; d out
BYTES_SWEEP = 13               # P, d, flags in; P out
BYTES_GRADIENT = 29            # P, u, v, w, flags in; u, v, w out
BYTES_ADVECT_VEL = 25          # u, v, w, flags in; u, v, w out
BYTES_ADVECT_DENS = 21         # rho, u, v, w, flags in; rho out
BYTES_REF_HALF = 25            # in-place half-sweep on faces: u, v, w, flags in; u, v, w out
BYTES_FIXED = (BYTES_FORCE + BYTES_DIVERGENCE + BYTES_GRADIENT + BYTES_ADVECT_VEL
               + BYTES_ADVECT_DENS)     # 121


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from smokesim import scenes
    if name == "grid_256":
        return scenes.grid_256()
    if name == "grid_512":
        return scenes.grid_512()
    if name == "reference_80":
        return scenes.reference_80()
    raise SystemExit(f"unknown workload {name}")


def describe(sc):
    p = sc["params"]
    nx, ny, nz = sc["dims"]
    proj = {0: "red-black SOR (reference, on faces)", 1: "red-black SOR (pressure)",
            2: "Jacobi"}[p["projection"]]
    solid = "floor + seeded boxes 5%" if sc.get("solid") is not None else "floor"
    return (f"{sc['name']}: {nx}x{ny}x{nz} FP32, {solid}, {len(sc['obstacles'])} obstacle(s), "
            f"{len(sc['sources'])} source(s), {proj} x{p['iterations']}, dt={sc['dt']}")


def sweep_bytes_per_cell(params):
    """Algorithmic bytes per cell of one full step for the chosen projection."""
    k = params["iterations"]
    if params["projection"] == 0:
        # In-place red-black half-sweep on faces: every u, v, w face is read
        # and written once per half-sweep, plus the flags byte.
        return BYTES_FORCE + BYTES_ADVECT_VEL + BYTES_ADVECT_DENS + 2 * k * BYTES_REF_HALF
    sweeps = k if params["projection"] == 2 else 2 * k
    return BYTES_FIXED + BYTES_SWEEP * sweeps


def cpu_baseline(sc, budget_s, max_steps):
    """The oracle's FP32 step on this host's cores, on the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from smokesim import scenes
    threads = min(16, os.cpu_count() or 1)
    orc.set_threads(threads)
    nx, ny, nz = sc["dims"]
    o = orc.Oracle(nx, ny, nz, 32)
    scenes.apply(o, sc)
    o.step(sc["dt"], 1)  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        o.step(sc["dt"], 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= max_steps:
            break
    cells = nx * ny * nz
    return {
        "value": cells * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
        "sample": f"{steps} full steps of the same {nx}x{ny}x{nz} workload after 1 warm-up step, "
                  f"oracle FP32 restatement, OpenMP over z ({el:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="grid_256")
    ap.add_argument("--projection", choices=["jacobi", "rbsor", "reference"], default=None)
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default Jacobi plan")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: only single-GPU runs are supported (asked for --gpus {args.gpus}); "
              "the step shards by z-slabs across GPUs only in design, not in this build",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import smokesim

    log(smokesim.device_info())
    sc = workload(args.workload)
    if args.projection:
        sc["params"]["projection"] = {"reference": 0, "rbsor": 1, "jacobi": 2}[args.projection]
    if args.iterations is not None:
        sc["params"]["iterations"] = args.iterations
    nx, ny, nz = sc["dims"]
    cells = nx * ny * nz
    log("workload:", describe(sc))

    sim = smokesim.Simulation(nx, ny, nz)
    smokesim.scenes.apply(sim, sc)
    dt = sc["dt"]
    tuned = None
    if sc["params"]["projection"] == 2 and not args.no_autotune:
        plan, ms = sim.autotune()
        tuned = {"plan": plan, "ms_per_sweep": round(ms, 5)}
        log(f"autotune: {plan} at {ms * 1e3:.2f} us/sweep")
    for _ in range(args.warmup):
        sim.step_async(dt)
    sim.synchronize()
    log(f"warm-up: {args.warmup} steps done")

    sim.timing_enable(True)
    sim.timing_reset()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step_async(dt)
    sim.synchronize()
    wall = time.perf_counter() - t0
    timing = sim.timing_read()
    step_ms = sim.timing_step_ms()
    sweep_ms, sweep_launches = sim.timing_sweeps()
    sim.timing_enable(False)
    dens = sim.read_density()
    if not np.isfinite(dens).all():
        raise SystemExit("non-finite density after the timed steps")
    log(f"timed: {args.steps} steps in {wall:.3f} s; density sum {float(dens.sum()):.3f}")

    ms_per_step = 1000.0 * wall / args.steps
    value = cells * args.steps / wall
    p = sc["params"]
    bpc = sweep_bytes_per_cell(p)

    # Dominant kernel: the pressure sweep.  Algorithmic bytes per launch =
    # 13 B x interior cells (the sweep's domain); duration from HIP events
    # around the sweep loops on the simulation's stream.
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    roofline = None
    if sweep_launches > 0:
        avg_ms = sweep_ms / sweep_launches
        if p["projection"] == 2:
            # fused launches: sweeps per launch = all timed sweeps / launches
            per_launch = p["iterations"] * timing["steps"] / sweep_launches
            alg = BYTES_SWEEP * interior * per_launch
            kernel = "k_jacobi_fused"
        elif p["projection"] == 1:
            alg = BYTES_SWEEP * interior
            kernel = "k_rbsor_half"
        else:
            alg = BYTES_REF_HALF * interior
            kernel = "k_project_ref_half"
        achieved = alg / (avg_ms * 1e-3) / 1e9
        roofline = {
            "bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "avg_launch_us": round(avg_ms * 1e3, 2),
            "alg_bytes_per_launch": int(alg), "launches_timed": sweep_launches,
            "jacobi_plan": sim.get_jacobi_plan() if p["projection"] == 2 else None,
            "autotune": tuned,
        }
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if roofline and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        key = f"{sc['name']}:{roofline['kernel']}"
        if key in tr:
            roofline["traffic"] = tr[key]["bytes_per_launch"]

    step_eff = cells * bpc / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": "grid cell-updates/s per full step",
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": describe(sc), "cells": cells, "projection": p["projection"],
            "iterations": p["iterations"], "parallelism": "single-gpu",
        },
        "latency_ms": {
            "p50": round(float(statistics.median(step_ms)), 4),
            "p99": round(float(np.percentile(step_ms, 99)), 4),
        },
        "stage_ms_per_step": {k: round(v / max(timing["steps"], 1), 4)
                              for k, v in timing["stage_ms"].items()},
        "step_roofline": {"bytes_per_cell": bpc, "achieved": round(step_eff, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_eff / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_budget, 50)
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
