#!/usr/bin/env python
"""Benchmark of the MI355X coupled Keller-Miksis ensemble solver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: 16384 systems of N=32 coupled bubbles (128 x 128 grid of forcing amplitude
0.5-2 bar x frequency 20-200 kHz, bubbles on a line at 150 um +-10 % spacing, radii
log-uniform 2-8 um), RKCK45 at rtol = atol = 1e-10 on both state components, one event
(dR/dtau = 0, maxima) with collapse recording. One step = one acoustic period (one Solve())
of every system. Inputs are resident in device memory before the timed region.

Prints one JSON line (metric, value = whole-job system-periods/s, roofline of the solver
kernel on the FP64 pipeline, CPU baseline from the host restatement on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) peak, dense


def flops_fm1(stats, n):
    """Algorithmic FP64 FLOPs (model FM-1): per stage 4N^2 + 47N, per BiCG iteration
    4N^2 + 21N, per step attempt 140N."""
    stages = float(stats[:, 2].sum())
    iters = float(stats[:, 3].sum())
    attempts = float(stats[:, 10].sum())
    return stages * (4 * n * n + 47 * n) + iters * (4 * n * n + 21 * n) + attempts * 140 * n


def cpu_baseline(seconds_budget=20.0, sample=48, seed=7):
    """Host restatement (literal mode, glibc, reference association), OpenMP over systems on
    the host cores, one period from rest for a seeded random sample of the workload's
    systems."""
    from bubbleensemble_implicitextension_amd import workloads as wl
    from oracle import pyoracle

    pyoracle.build()
    ens = wl.config1()
    rng = np.random.default_rng(seed)
    sids = np.sort(rng.choice(ens.NS, size=sample, replace=False))
    sub = ens.subset(sids)
    p = pyoracle.OracleProblem.from_ensemble(sub)
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    p.solve(mode=pyoracle.LITERAL, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "system-periods/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample} seeded-random systems of the same 16384-system workload, "
                      f"first period from rest, literal host restatement, {threads} OpenMP "
                      f"threads, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--end-time", type=float, default=1.0,
                    help="fraction of a period per step (profiling only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(systems are independent: shard them across GPUs by cost, see DESIGN.md)",
              file=sys.stderr)
        return 2

    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    ens = wl.config1()
    N = ens.N
    solver = collapse.make_solver(ens)
    collapse.fill(solver, ens)
    solver.host_view(V.TimeDomain)[ens.NS:] = args.end_time
    solver.SynchroniseFromHostToDevice(V.All)
    for _ in range(args.warmup):
        solver.Solve()
    solver.InsertSynchronisationPoint()
    solver.SynchroniseSolver()
    print(f"# warmup done ({args.warmup} periods)", file=sys.stderr, flush=True)
    solver.reset_stats()
    solver.SynchroniseDevice()

    t0 = time.perf_counter()
    solver.timer_start()
    for k in range(args.steps):
        solver.Solve()
    kernel_ms = solver.timer_stop()
    solver.SynchroniseDevice()
    wall = time.perf_counter() - t0

    stats = solver.stats()
    flops = flops_fm1(stats, N)
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    iters_per_rhs = float(stats[:, 3].sum()) / max(1.0, float(stats[:, 2].sum()))
    value = ens.NS * args.steps / wall

    line = {
        "metric": "system-periods/s (plain FP64 RKCK45 + BiCG, rtol=atol=1e-10)",
        "value": value,
        "unit": "system-periods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry, radii and forcing grid)",
        "config": {"workload": "16384 systems x N=32 KM bubbles, 128x128 PA x f grid, "
                               "RKCK45 rtol=atol=1e-10, 1 period per step",
                   "systems": ens.NS, "units_per_system": N, "end_time": args.end_time,
                   "parallelism": "1 GPU"},
        "bicg_iterations_per_rhs": iters_per_rhs,
        "steps_accepted_per_system_period": float(stats[:, 0].sum()) / (ens.NS * args.steps),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "bexk::solve_kernel<32, RKCK45>",
                     "kernel_ms_per_launch": kernel_ms / args.steps,
                     "flops_per_launch": flops / args.steps},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
