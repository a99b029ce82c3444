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


def cpu_baseline(sample=2048, seed=7):
    """Host restatement (literal mode: glibc, reference association), OpenMP over systems on
    the host cores (dynamic schedule), one period from rest for a seeded random sample of the
    workload's systems (about 13 s with 16 threads, then a few seconds on one thread)."""
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
    # One thread on every 32nd system of the same sample (64 systems over the whole grid; the
    # sample is sorted by system index, i.e. by frequency, so a prefix would be the slowest).
    one_sids = sids[::max(1, sample // 64)][:64]
    one = pyoracle.OracleProblem.from_ensemble(ens.subset(one_sids))
    t1 = time.perf_counter()
    one.solve(mode=pyoracle.LITERAL, nthreads=1)
    dt1 = time.perf_counter() - t1
    return {"value": sample / dt, "unit": "system-periods/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample} seeded-random systems (seed {seed}) of the same 16384-system "
                      f"workload, first period from rest, literal host restatement, {threads} "
                      f"OpenMP threads, {dt:.1f} s",
            "single_thread_value": len(one_sids) / dt1,
            "single_thread_sample": f"one in {max(1, sample // 64)} systems of that sample "
                                    f"({len(one_sids)} systems), 1 thread, {dt1:.1f} s"}


def measured_traffic(kernel_tag):
    """HBM bytes per launch of the solver kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's reading for a plain
    device copy of known size). None when the summary is absent."""
    path = os.path.join(REPO, "profiles", "solve_kernel_n32_16384sys_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch[" + kernel_tag + "]:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration():
    """Plain 1 GiB device-to-device copy (2^27 doubles), for scaling the FETCH_SIZE and
    WRITE_SIZE counters: run under rocprofv3 --pmc and compare with 1 GiB."""
    import torch

    n = 1 << 27
    x = torch.empty(n, dtype=torch.float64, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8}), flush=True)
    return 0


def run_leg(ens, steps, warmup, end_time, network=None):
    """Times `steps` Solve() calls of the whole ensemble (inputs resident on the device)."""
    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    solver = collapse.make_solver(ens)
    collapse.fill(solver, ens)
    solver.host_view(V.TimeDomain)[ens.NS:] = end_time
    if network is not None:
        solver.set_warm_start_network(network)
        solver.SolverOption(O.WarmStart, 1)
    solver.SynchroniseFromHostToDevice(V.All)
    for _ in range(warmup):
        solver.Solve()
    solver.InsertSynchronisationPoint()
    solver.SynchroniseSolver()
    solver.reset_stats()
    solver.SynchroniseDevice()

    t0 = time.perf_counter()
    solver.timer_start()
    for _ in range(steps):
        solver.Solve()
    kernel_ms = solver.timer_stop()
    solver.SynchroniseDevice()
    wall = time.perf_counter() - t0

    solver.SynchroniseFromDeviceToHost(V.All)
    solver.InsertSynchronisationPoint()  # the copies are asynchronous on the solver's stream
    solver.SynchroniseSolver()
    out = {"wall": wall, "kernel_ms": kernel_ms, "stats": solver.stats(),
           "state": solver.state().copy(), "acc": solver.accessories().copy(),
           "events": solver.event_counts().copy()}
    solver.close()
    return out


def run_learned_batched_leg(ens, steps, warmup, end_time, network, heavy):
    """Times `steps` periods of the ensemble split by cost: the `heavy` costliest systems
    (forcing amplitude over frequency) with the learned start, the rest with the plain start,
    both solver objects launched together (inputs resident on the device)."""
    from bubbleensemble_implicitextension_amd.warm_start import LearnedBatchedSolver

    lb = LearnedBatchedSolver(ens, network, heavy=heavy, end_time=end_time)
    for _ in range(warmup):
        lb.Solve()
    lb.synchronise()
    lb.reset_stats()
    lb.synchronise()
    t0 = time.perf_counter()
    for _ in range(steps):
        lb.Solve()
    lb.synchronise()
    wall = time.perf_counter() - t0
    out = lb.results()
    out["wall"] = wall
    out["heavy"] = lb.heavy
    lb.close()
    return out


def error_vs(ref, other):
    """Relative differences of final states and recorded collapse extrema, and event-count
    agreement, of one leg against another on the same systems."""
    ok = np.isfinite(ref["state"]) & np.isfinite(other["state"])
    rs = np.abs(other["state"] - ref["state"]) / np.maximum(np.abs(ref["state"]), 1e-3)
    same = ref["events"] == other["events"]  # [NS, N] units with identical event counts
    ra = np.abs(other["acc"] - ref["acc"]) / np.maximum(np.abs(ref["acc"]), 1e-3)
    ra = ra[same]  # collapse records are comparable where the same events were recorded
    return {"state_max_rel": float(np.max(rs[ok])), "state_median_rel": float(np.median(rs[ok])),
            "state_frac_within_1e-6": float(np.mean(rs[ok] <= 1e-6)),
            "collapse_max_rel_same_events": float(np.max(ra)) if ra.size else None,
            "collapse_frac_within_1e-6": float(np.mean(ra <= 1e-6)) if ra.size else None,
            "units_event_counts_identical_frac": float(np.mean(same))}


def run_pool_leg(end_time, probe_fraction=0.02):
    """Config 2 (20480 systems, N in {8, 16, 32, 64, 128}, per-system coupling) through a
    ProblemPool: autotuned variants, one graph launch for the first `end_time` of a period,
    then the same classes launched one after another for comparison."""
    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    classes = wl.config2()
    pool = collapse.make_pool(classes)

    def reset(pl, cls):
        for c, e in enumerate(cls):
            sv = pl.solver(c)
            collapse.fill(sv, e)
            sv.host_view(V.TimeDomain)[e.NS:] = end_time
            sv.SynchroniseFromHostToDevice(V.All)
            sv.reset_stats()
        pl.synchronise()

    reset(pool, classes)
    tuned = pool.autotune(probe_fraction)
    chosen = {classes[c].N: pool.variants(c)[pool.variant(c)] for c in range(len(classes))}
    reset(pool, classes)
    t0 = time.perf_counter()
    pool.Solve()
    pool.synchronise()
    wall = time.perf_counter() - t0
    stats = [pool.solver(c).stats() for c in range(len(classes))]
    flops = sum(flops_fm1(st, e.N) for st, e in zip(stats, classes))
    seq = {}
    for c, e in enumerate(classes):
        one = collapse.make_pool([e])
        one.set_variant(0, pool.variant(c))
        one.set_order(0, pool.get_order(c))
        reset(one, [e])
        t1 = time.perf_counter()
        one.Solve()
        one.synchronise()
        seq[e.N] = time.perf_counter() - t1
        one.close()
    pool.close()
    ns = sum(e.NS for e in classes)
    return {
        "workload": f"config 2: {ns} systems, N in {{8,16,32,64,128}} x 4096, random 3D "
                    f"positions (50 um min separation), per-system coupling, first {end_time} "
                    "of a period from rest",
        "value": ns * end_time / wall, "unit": "system-periods/s (period fraction counted)",
        "ms": wall * 1e3,
        "classes_one_after_another_ms": 1e3 * sum(seq.values()),
        "class_alone_ms": {str(k): v * 1e3 for k, v in seq.items()},
        "variants": {str(k): v for k, v in chosen.items()},
        "autotune_probe_ms": {str(classes[c].N): tuned[c] for c in tuned},
        "setup": "the autotuner (variant choice and longest-probe-first queue order, probes of "
                 f"{probe_fraction} of the time domain) runs before the timed region: "
                 f"{sum(sum(t.values()) for t in tuned.values()):.0f} ms of probes",
        "bicg_iterations_per_rhs": {str(e.N): float(st[:, 3].sum()) / max(1.0, float(st[:, 2].sum()))
                                    for st, e in zip(stats, classes)},
        "fm1_tflops": flops / wall / 1e12,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm-start leg")
    ap.add_argument("--heavy", type=int, default=64,
                    help="systems on the learned start in the learned+batched leg")
    ap.add_argument("--pool-end-time", type=float, default=0.1,
                    help="period fraction of the config-2 pool leg (0 skips it)")
    ap.add_argument("--end-time", type=float, default=1.0,
                    help="fraction of a period per step (profiling only)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(systems are independent: a multi-GPU run shards them by estimated cost, one "
              "process and one solver per GPU)",
              file=sys.stderr)
        return 2

    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd import warm_start

    ens = wl.config1()
    N = ens.N
    plain = run_leg(ens, args.steps, args.warmup, args.end_time)
    print(f"# plain leg done ({plain['kernel_ms']:.1f} ms per {args.steps} launches)",
          file=sys.stderr, flush=True)
    stats = plain["stats"]
    kernel_ms = plain["kernel_ms"]
    wall = plain["wall"]
    flops = flops_fm1(stats, N)
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    iters_per_rhs = float(stats[:, 3].sum()) / max(1.0, float(stats[:, 2].sum()))
    value = ens.NS * args.steps / wall

    line = {
        "metric": "systems x acoustic periods/s at rtol=atol=1e-10, plain FP64 RKCK45 + BiCG "
                  "(learned warm start vs plain in 'learned', learned+batched in "
                  "'learned_batched'); BiCG iters/RHS",
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
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic("period") if args.end_time == 1.0 else None,
                     "kernel": "bexk::solve_kernel<RolePairLayout32<kPlain>, RKCK45>",
                     "kernel_ms_per_launch": kernel_ms / args.steps,
                     "flops_per_launch": flops / args.steps},
    }
    if not args.no_learned and os.path.exists(warm_start.DEFAULT_NETWORK):
        net = np.load(warm_start.DEFAULT_NETWORK)
        learned = run_leg(ens, args.steps, args.warmup, args.end_time, network=net)
        ls = learned["stats"]
        lv = ens.NS * args.steps / learned["wall"]
        line["learned"] = {
            "value": lv, "unit": "system-periods/s",
            "ms_per_step": learned["wall"] * 1e3 / args.steps,
            "kernel_ms_per_launch": learned["kernel_ms"] / args.steps,
            "bicg_iterations_per_rhs": float(ls[:, 3].sum()) / max(1.0, float(ls[:, 2].sum())),
            "warm_fallbacks": int(ls[:, 31].sum()),
            "speedup_vs_plain": lv / value,
            "network": "5-64-64-1 BF16 MFMA, trained on a held-out-seed ensemble",
            "error_vs_plain": error_vs(plain, learned),
        }
    if not args.no_learned and os.path.exists(warm_start.DEFAULT_NETWORK):
        net = np.load(warm_start.DEFAULT_NETWORK)
        lbr = run_learned_batched_leg(ens, args.steps, args.warmup, args.end_time, net,
                                      args.heavy)
        lbs = lbr["stats"]
        lbv = ens.NS * args.steps / lbr["wall"]
        line["learned_batched"] = {
            "value": lbv, "unit": "system-periods/s",
            "ms_per_step": lbr["wall"] * 1e3 / args.steps,
            "speedup_vs_plain": lbv / value,
            "split": f"{len(lbr['heavy'])} systems with the largest forcing amplitude / "
                     "frequency on the learned start, the others on the plain start, two "
                     "solver objects launched together",
            "learned_systems": [int(k) for k in lbr["heavy"][:8]] + (["..."] if len(lbr["heavy"]) > 8 else []),
            "bicg_iterations_per_rhs": float(lbs[:, 3].sum()) / max(1.0, float(lbs[:, 2].sum())),
            "error_vs_plain": error_vs(plain, lbr),
        }
        print(f"# learned+batched leg done ({lbr['wall'] * 1e3:.1f} ms)", file=sys.stderr,
              flush=True)
    if args.pool_end_time > 0:
        line["pool"] = run_pool_leg(args.pool_end_time)
        print(f"# pool leg done ({line['pool']['ms']:.0f} ms)", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
