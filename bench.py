#!/usr/bin/env python
"""Benchmark of the MI355X coupled Keller-Miksis ensemble solver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--budget S] [--no-cpu-baseline]

Headline workload: 16384 systems of N=32 coupled bubbles (128 x 128 grid of forcing amplitude
0.5-2 bar x frequency 20-200 kHz, bubbles on a line at 150 um +-10 % spacing, radii
log-uniform 2-8 um), RKCK45 at rtol = atol = 1e-10 on both state components, one event
(dR/dtau = 0, maxima) with collapse recording, dense output of 64 points per period (cubic
Hermite at tau = k/64). One step = one acoustic period (one Solve()) of every system, started
from rest: before every Solve() the state, time, accessories and dense-output index are
restored from a device-resident checkpoint by device-to-device copies on the solver's stream,
inside the timed region. Inputs are resident in device memory before the timed region; the
dense output stays in device memory (the PCIe-inclusive rate with each period's dense output
copied to the host is reported beside it).

Further legs, each skipped (and said so) when the wall budget would not hold it: the learned
BiCG warm start; learned + batched (two solver objects launched together: the stiffest
systems on the learned start, the next tier by measured cost one per wave, the rest two per
wave) and the same split on the plain start (checked bit for bit against the plain leg); the
BiCG residual 1e-13 workload (plain and learned); the second period of the headline workload
(the steady state, with an attempt cap for the systems that reach the Keller-Miksis Mach-1
singularity); the uneven-N pool (20480 systems, a tenth of a period); and the N = 256 clusters
(4096 systems, a fraction of a period). The CPU baseline times the host restatement on a
bounded sample, in a child process (`--leg cpu`) so that its OpenMP threads are gone before the
GPU legs that follow. `--leg pool|n256|period2|split|cpu` runs one leg alone and prints its
record.

Prints one JSON line (metric, value = whole-job system-periods/s of the plain method, roofline
of the solver kernel on the FP64 pipeline, CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import threading
import time

import numpy as np

T_START = time.perf_counter()
REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 peak, dense (vector and matrix pipes alike)
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E peak bandwidth
DENSE_POINTS = 64        # dense-output points per period of the headline workload
KERNEL_NAME = "bexk::queue_kernel<RolePairLayout32<kPlain>, RKCK45, dense>"


def elapsed():
    return time.perf_counter() - T_START


def log(msg):
    print(f"# [{elapsed():6.1f} s] {msg}", file=sys.stderr, flush=True)


class Output:
    """The one JSON line; a watchdog prints what is complete if the run nears its limit."""

    def __init__(self):
        self.line = None
        self.lock = threading.Lock()
        self.done = False

    def emit(self, note=None):
        with self.lock:
            if self.done:
                return False
            self.done = True
            if self.line is None:
                return False
            if note:
                self.line["truncated"] = note
            print(json.dumps(self.line), flush=True)
            return True


def flops_fm1(stats, n):
    """Algorithmic FP64 FLOPs (model FM-1): per stage 4N^2 + 47N, per BiCG iteration
    4N^2 + 21N, per step attempt 140N."""
    stages = float(stats[:, 2].sum())
    iters = float(stats[:, 3].sum())
    attempts = float(stats[:, 10].sum())
    return stages * (4 * n * n + 47 * n) + iters * (4 * n * n + 21 * n) + attempts * 140 * n


def fp64_rate(stats, n, wall):
    """FM-1 FLOPs of the counted launches over their wall time, against the FP64 peak."""
    achieved = flops_fm1(stats, n) / wall / 1e12
    return {"achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "note": "FM-1 FLOPs of all parts over the wall time of the timed steps"}


def iters_per_rhs(stats, rows=None):
    st = stats if rows is None else stats[rows]
    return float(st[:, 3].sum()) / max(1.0, float(st[:, 2].sum()))


def host_cpu():
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


def cpu_baseline_child(limit):
    """cpu_baseline() in a child process (bench.py --leg cpu, which touches no GPU): its OpenMP
    worker threads end with the child instead of staying in this process, where their idle
    spinning competed with the host thread that launches and synchronises the GPU legs after
    it (the split's periods measured ~10 % slower behind an in-process baseline)."""
    try:
        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--leg", "cpu"],
                             capture_output=True, text=True, timeout=max(10.0, limit))
    except subprocess.TimeoutExpired:
        return {"skipped": f"the CPU baseline did not finish within {limit:.0f} s"}
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if res.returncode != 0 or not lines:
        return {"skipped": f"the CPU baseline failed (exit {res.returncode}): "
                           f"{res.stderr.strip()[-300:]}"}
    return json.loads(lines[-1])


def cpu_baseline(sample=2048, seed=7):
    """Host restatement (literal mode: glibc, reference association), OpenMP over systems with
    a dynamic schedule, one period from rest for a seeded random sample of the headline
    workload's systems; then one thread on 64 of them. The thread count is the host's CPU share
    of one GPU (16, or fewer usable CPUs)."""
    from bubbleensemble_implicitextension_amd import workloads as wl
    from oracle import pyoracle

    pyoracle.build()
    native = pyoracle.build_native()
    build = ("g++ -O3 -march=native -fopenmp, built on this host" if native else
             "g++ -O3 -march=x86-64-v3 -fopenmp (the -march=native build failed)")
    cpu = host_cpu()
    threads = max(1, min(cpu["usable_cpus"], 16))
    ens = wl.config1()
    rng = np.random.default_rng(seed)
    sids = np.sort(rng.choice(ens.NS, size=sample, replace=False))
    p = pyoracle.OracleProblem.from_ensemble(ens.subset(sids))
    # One thread on every 32nd system of the same sample (64 systems over the whole grid; the
    # sample is sorted by system index, i.e. by frequency, so a prefix would be the slowest).
    one_sids = sids[::max(1, sample // 64)][:64]
    one = pyoracle.OracleProblem.from_ensemble(ens.subset(one_sids))
    with pyoracle.using_library(native or pyoracle.LIB_PATH):
        t0 = time.perf_counter()
        p.solve(mode=pyoracle.LITERAL, nthreads=threads)
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        one.solve(mode=pyoracle.LITERAL, nthreads=1)
        dt1 = time.perf_counter() - t1
    return {"value": sample / dt, "unit": "system-periods/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample} seeded-random systems (seed {seed}) of the 16384-system "
                      f"workload, first period from rest without dense output, literal host "
                      f"restatement ({build}), {threads} OpenMP threads, {dt:.1f} s",
            "host_cpu": cpu["model"], "host_nproc": cpu["nproc"],
            "host_usable_cpus": cpu["usable_cpus"],
            "threads_note": f"{threads} threads: every usable CPU, at most 16 (the host CPU "
                            f"share of one GPU)",
            "single_thread_value": len(one_sids) / dt1,
            "single_thread_sample": f"one in {max(1, sample // 64)} systems of that sample "
                                    f"({len(one_sids)} systems), 1 thread, {dt1:.1f} s"}


def measured_traffic(profile="queue_kernel_rolepair32_16384sys_dense64_pmc.md"):
    """HBM bytes per launch of a leg's kernel from the committed rocprofv3 counter summary of
    the same command (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's
    reading for a plain device copy of known size); by default the headline kernel's. None
    when the summary is absent."""
    path = os.path.join(REPO, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
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


def dense_options(end_time):
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O

    return {O.DenseOutputInterpolation.name: 1,
            O.DenseOutputMinimumTimeStep.name: end_time / DENSE_POINTS}


def make_headline_solver(ens, end_time, network=None, options=None, dense=True):
    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    s = collapse.make_solver(ens, NDO=DENSE_POINTS if dense else 0)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = end_time
    opts = dict(dense_options(end_time)) if dense else {}
    opts.update(options or {})
    collapse.apply_options(s, opts)
    if network is not None:
        s.set_warm_start_network(network)
        s.SolverOption(O.WarmStart, 1)
    s.SynchroniseFromHostToDevice(V.All)
    s.checkpoint_save()
    return s


def time_steps(target, steps, warmup, timer=None):
    """Warmup, then `steps` timed (restore, Solve) pairs on `target` (a solver or a split
    solver). Returns wall seconds and, with a timer solver, the HIP-event time of the timed
    launches on its stream minus the same number of restores timed alone. A split solver's
    parts run on streams of their own; every step ends with all of them done (a period of the
    whole ensemble), so no part of one step overlaps the next."""
    split = hasattr(target, "parts")
    for _ in range(warmup):
        target.checkpoint_restore()
        target.Solve()
        if split:
            target.synchronise()
    target.synchronise() if hasattr(target, "synchronise") else target.SynchroniseDevice()
    restore_ms = 0.0
    if timer is not None:
        timer.timer_start()
        for _ in range(steps):
            timer.checkpoint_restore()
        restore_ms = timer.timer_stop()
    target.reset_stats()
    target.synchronise() if hasattr(target, "synchronise") else target.SynchroniseDevice()
    t0 = time.perf_counter()
    if timer is not None:
        timer.timer_start()
    for _ in range(steps):
        target.checkpoint_restore()
        target.Solve()
        if split:
            target.synchronise()
    event_ms = timer.timer_stop() if timer is not None else None
    target.synchronise() if hasattr(target, "synchronise") else target.SynchroniseDevice()
    wall = time.perf_counter() - t0
    kernel_ms = None if event_ms is None else max(0.0, event_ms - restore_ms)
    return wall, kernel_ms, restore_ms


def collect(s, dense=True):
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    s.SynchroniseFromDeviceToHost(V.All)
    if dense:
        s.SynchroniseFromDeviceToHost(V.DenseOutput)
    s.InsertSynchronisationPoint()  # the copies are asynchronous on the solver's stream
    s.SynchroniseSolver()
    out = {"state": s.state().copy(), "acc": s.accessories().copy(),
           "events": s.event_counts().copy(), "stats": s.stats()}
    if dense:
        idx = s.host_view(V.DenseIndex).copy()
        n = int(idx.max()) if idx.size else 0
        out["dense_index"] = idx
        out["dense_state"] = s.host_view(V.DenseState)[: n * 2 * s.NS * s.UPS].reshape(
            n, 2, s.NS, s.UPS).copy()
    return out


def run_leg(ens, steps, warmup, end_time, network=None, options=None, stream_steps=0):
    """Times `steps` from-rest periods of the whole ensemble (inputs resident on the device)."""
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    s = make_headline_solver(ens, end_time, network=network, options=options)
    wall, kernel_ms, restore_ms = time_steps(s, steps, warmup, timer=s)
    out = collect(s)
    out.update(wall=wall, kernel_ms=kernel_ms, restore_ms=restore_ms)
    if stream_steps > 0:
        # PCIe-inclusive: each period's dense output (index, times, states) goes to pinned host
        # memory before the next period starts.
        s.SynchroniseDevice()
        t0 = time.perf_counter()
        for _ in range(stream_steps):
            s.checkpoint_restore()
            s.Solve()
            s.SynchroniseFromDeviceToHost(V.DenseOutput)
            s.InsertSynchronisationPoint()
            s.SynchroniseSolver()
        out["stream_wall"] = time.perf_counter() - t0
        out["stream_steps"] = stream_steps
    s.close()
    return out


def error_vs(ref, other):
    """Relative differences of final states, dense output and recorded collapse extrema, and
    event-count agreement, of one leg against another on the same systems."""
    ok = np.isfinite(ref["state"]) & np.isfinite(other["state"])
    rs = np.abs(other["state"] - ref["state"]) / np.maximum(np.abs(ref["state"]), 1e-3)
    same = ref["events"] == other["events"]  # [NS, N] units with identical event counts
    ra = np.abs(other["acc"] - ref["acc"]) / np.maximum(np.abs(ref["acc"]), 1e-3)
    ra = ra[same]  # collapse records are comparable where the same events were recorded
    out = {"state_max_rel": float(np.max(rs[ok])), "state_median_rel": float(np.median(rs[ok])),
           "state_frac_within_1e-6": float(np.mean(rs[ok] <= 1e-6)),
           "collapse_max_rel_same_events": float(np.max(ra)) if ra.size else None,
           "collapse_frac_within_1e-6": float(np.mean(ra <= 1e-6)) if ra.size else None,
           "units_event_counts_identical_frac": float(np.mean(same)),
           "systems_event_counts_identical_frac": float(np.mean(np.all(same, axis=1)))}
    if "dense_state" in ref and "dense_state" in other and \
            ref["dense_state"].shape == other["dense_state"].shape:
        a, b = ref["dense_state"], other["dense_state"]
        good = np.isfinite(a) & np.isfinite(b)
        rd = np.abs(b - a)[good] / np.maximum(np.abs(a[good]), 1e-3)
        out["dense_max_rel"] = float(np.max(rd)) if rd.size else None
        out["dense_frac_within_1e-6"] = float(np.mean(rd <= 1e-6)) if rd.size else None
    meets = (out["state_max_rel"] <= 1e-6 and out["units_event_counts_identical_frac"] == 1.0
             and (out["collapse_max_rel_same_events"] or 0.0) <= 1e-6)
    out["meets_1e-6_and_identical_events"] = bool(meets)
    return out


def tail_split(stats, k=64):
    """BiCG iterations per RHS of the k systems with the most RK stages (the launch's tail) and
    of the others (the bulk)."""
    order = np.argsort(-stats[:, 2], kind="stable")
    return {"tail_systems": int(k), "tail": iters_per_rhs(stats, order[:k]),
            "bulk": iters_per_rhs(stats, order[k:])}


def wait_with_progress(pool, solvers, label, every=20.0, limit=None):
    """Waits for a launch (pool or solver); every `every` seconds it logs how many systems have
    finished, read from the per-system timeline the kernels write (enabled before the launch).
    After `limit` seconds it asks the remaining systems to stop (bex_request_stop) and returns
    the number that had finished by then (None if all finished within the limit)."""
    total = sum(s.NS for s in solvers)
    t0 = last = time.perf_counter()
    finished_at_stop = None
    while True:
        done = sum(s.finished_systems() for s in solvers)
        if done >= total:
            break
        now = time.perf_counter()
        if limit is not None and finished_at_stop is None and now - t0 >= limit:
            finished_at_stop = done
            log(f"{label}: {done} of {total} systems finished after {limit:.0f} s; stopping "
                "the others")
            for s in solvers:
                s.request_stop()
        if now - last >= every:
            log(f"{label}: {done} of {total} systems finished")
            last = now
        time.sleep(0.05)
    pool.synchronise()
    for s in solvers:
        s.clear_stop()
    return finished_at_stop


def run_split(ens, network, heavy, medium, steps, warmup, end_time, cost):
    """Times `steps` from-rest periods of the split (batched.LearnedBatchedSolver)
    with dense output; returns wall seconds, the results and the parts."""
    from bubbleensemble_implicitextension_amd import batched

    lb = batched.LearnedBatchedSolver(ens, network, heavy=heavy, medium=medium,
                                      end_time=end_time, ndo=DENSE_POINTS,
                                      options=dense_options(end_time), cost=cost)
    lb.checkpoint_save()
    wall, _, _ = time_steps(lb, steps, warmup)
    res = lb.results(dense=True)
    parts = {"heavy": lb.heavy, "medium": lb.medium, "light": lb.light}
    lb.close()
    return wall, res, parts


def split_text(parts, learned):
    start = "learned" if learned else "plain"
    return (f"{len(parts['heavy'])} systems with the largest forcing amplitude / frequency on the "
            f"{start} start (one per wave), the next {len(parts['medium'])} by the cost measured "
            "in the plain leg's solve on the plain start (one per wave), the other "
            f"{len(parts['light'])} on the plain start two per wave (system-pair layout); two "
            "solver objects launched together (the one-per-wave tiers share one launch), each "
            "step ends when both are done")


def run_split_leg(steps, warmup, heavy, medium):
    """The learned+batched split alone (profiling): one untimed plain solve measures the
    systems' costs, then `steps` timed periods of the split."""
    from bubbleensemble_implicitextension_amd import batched
    from bubbleensemble_implicitextension_amd import workloads as wl

    ens = wl.config1()
    s = make_headline_solver(ens, 1.0)
    s.checkpoint_restore()
    s.Solve()
    s.SynchroniseDevice()
    cost = batched.measured_cost(s.stats(), ens.N)
    s.close()
    net = np.load(batched.DEFAULT_NETWORK)
    wall, res, parts = run_split(ens, net, heavy, medium, steps, warmup, 1.0, cost)
    return {"value": ens.NS * steps / wall, "unit": "system-periods/s",
            "ms_per_step": wall * 1e3 / steps, "split": split_text(parts, True),
            "bicg_iterations_per_rhs": iters_per_rhs(res["stats"])}


def run_pool_leg(end_time, probe_fraction=0.02, sequential=True, limit=None, cap=0):
    """Config 2 (20480 systems, N in {8, 16, 32, 64, 128}, per-system coupling) through a
    ProblemPool: autotuned variants and queue order, one graph launch over [0, end_time] from
    rest, then (optionally) the same classes launched one after another. cap > 0 ends every
    system after that many step attempts (MaximumStepAttempts)."""
    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    classes = wl.config2()
    pool = collapse.make_pool(classes)

    def reset(pl, cls):
        for c, e in enumerate(cls):
            sv = pl.solver(c)
            collapse.fill(sv, e)
            if cap:
                sv.SolverOption(O.MaximumStepAttempts, cap)
            sv.host_view(V.TimeDomain)[e.NS:] = end_time
            sv.SynchroniseFromHostToDevice(V.All)
            sv.reset_stats()
        pl.synchronise()

    reset(pool, classes)
    tuned = pool.autotune(probe_fraction)
    chosen = {classes[c].N: pool.variants(c)[pool.variant(c)] for c in range(len(classes))}
    reset(pool, classes)
    solvers = [pool.solver(c) for c in range(len(classes))]
    for sv in solvers:
        sv.enable_timeline(True)
    pool.synchronise()
    t0 = time.perf_counter()
    pool.Solve()
    stopped_after = wait_with_progress(pool, solvers, "pool", limit=limit)
    wall = time.perf_counter() - t0
    stats = [sv.stats() for sv in solvers]
    slowest = {str(e.N): float(sv.timeline()[:, 1].max()) for sv, e in zip(solvers, classes)}
    completed = {}
    for sv, e in zip(solvers, classes):
        sv.SynchroniseFromDeviceToHost(V.ActualTime)
        sv.InsertSynchronisationPoint()
        sv.SynchroniseSolver()
        completed[str(e.N)] = int(np.count_nonzero(sv.host_view(V.ActualTime) >= end_time - 1e-9))
    flops = sum(flops_fm1(st, e.N) for st, e in zip(stats, classes))
    seq = {}
    for sv in solvers:
        sv.enable_timeline(False)
    if sequential:
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
    achieved = flops / wall / 1e12
    out = {
        "workload": f"{ns} systems, N in {{8,16,32,64,128}} x 4096, random 3D positions (50 um "
                    f"min separation), per-system coupling, RKCK45 rtol=atol=1e-10, "
                    f"tau in [0, {end_time}] from rest",
        "value": ns * end_time / wall, "unit": "system-periods/s",
        "ms": wall * 1e3,
        "variants": {str(k): v for k, v in chosen.items()},
        "autotune_probe_ms": {str(classes[c].N): tuned[c] for c in tuned},
        "setup": "the autotuner (variant choice and longest-probe-first queue order, probes of "
                 f"{probe_fraction} of the time domain) runs before the timed region: "
                 f"{sum(sum(t.values()) for t in tuned.values()):.0f} ms of probes",
        "bicg_iterations_per_rhs": {str(e.N): iters_per_rhs(st) for st, e in zip(stats, classes)},
        "class_last_system_end_s": slowest,
        "systems_completed": completed,
        "attempt_cap": cap or None,
        "systems_capped_or_stopped": {str(e.N): e.NS - completed[str(e.N)] for e in classes},
        "stopped_after_s": None if stopped_after is None else limit,
        "forced_accepts_per_system_max": {str(e.N): int(st[:, 7].max())
                                          for st, e in zip(stats, classes)},
        "max_accepted_steps_per_system": {str(e.N): int(st[:, 0].max())
                                          for st, e in zip(stats, classes)},
        "forced_accepts": {str(e.N): int(st[:, 7].sum()) for st, e in zip(stats, classes)},
        "roofline": {"bound": "mfma", "compute_pipe": "fp64 vector ALU", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "note": "FM-1 FLOPs of all classes over the graph launch's wall time"},
    }
    if seq:
        out["classes_one_after_another_ms"] = 1e3 * sum(seq.values())
        out["class_alone_ms"] = {str(k): v * 1e3 for k, v in seq.items()}
    return out


def run_second_period_leg(limit=None, cap=0):
    """Headline workload's second period (continuing from the end of the first, no dense
    output): per-system finish times from the kernel timeline, so the launch's tail (the systems
    near the Keller-Miksis Mach-1 singularity) and its bulk are reported apart. cap > 0 ends a
    system after that many step attempts of the second period (MaximumStepAttempts)."""
    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    ens = wl.config1()
    s = make_headline_solver(ens, 1.0, dense=False)
    s.Solve()
    s.SynchroniseDevice()
    first = s.stats()
    s.reset_stats()
    if cap:
        s.SolverOption(O.MaximumStepAttempts, cap)
    s.enable_timeline(True)
    s.SynchroniseDevice()
    t0 = time.perf_counter()
    s.Solve()

    class _Launch:  # wait_with_progress synchronises through this
        def synchronise(self):
            s.SynchroniseDevice()
    stopped_after = wait_with_progress(_Launch(), [s], "second period", limit=limit)
    wall = time.perf_counter() - t0
    st = s.stats()
    end = s.timeline()[:, 1]
    s.SynchroniseFromDeviceToHost(V.ActualTime)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    completed = int(np.count_nonzero(s.host_view(V.ActualTime) >= 1.0 - 1e-9))
    s.close()
    order = np.argsort(-end)
    achieved = flops_fm1(st, ens.N) / wall / 1e12
    return {"workload": "16384 systems x N=32, second period (continuing from the first)",
            "value": ens.NS / wall, "unit": "system-periods/s",
            "wall_s": wall, "systems_completed": completed, "attempt_cap": cap or None,
            "fp64": {"achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS},
            "stopped_after_s": None if stopped_after is None else limit, "finish_s_quantiles": {q: float(np.quantile(end, float(q)))
                                                   for q in ("0.5", "0.9", "0.99", "0.999")},
            "last_systems": [{"system": int(k), "finish_s": float(end[k]),
                              "accepted": int(st[k, 0]), "rejected": int(st[k, 1]),
                              "forced_accepts": int(st[k, 7]),
                              "iterations_per_rhs": iters_per_rhs(st, [k])} for k in order[:8]],
            "accepted_per_system_first_period_mean": float(first[:, 0].mean()),
            "accepted_per_system_second_period_mean": float(st[:, 0].mean()),
            "accepted_per_system_second_period_median": float(np.median(st[:, 0])),
            "fm1_flops_ratio_second_to_first": flops_fm1(st, ens.N) / flops_fm1(first, ens.N)}


def run_config4_leg(end_time, ns=4096):
    """Config 4 (ns systems of N = 256 in a 1 mm sphere, per-system coupling) over
    [0, end_time] from rest, one persistent launch through a one-class pool."""
    from bubbleensemble_implicitextension_amd import collapse
    from bubbleensemble_implicitextension_amd import workloads as wl
    from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

    ens = wl.config4(ns=ns)
    pool = collapse.make_pool([ens])
    sv = pool.solver(0)
    sv.host_view(V.TimeDomain)[ens.NS:] = end_time
    sv.SynchroniseFromHostToDevice(V.All)
    sv.checkpoint_save()
    sv.enable_timeline(True)
    runs = {}
    for v, name in enumerate(pool.variants(0)):
        pool.set_variant(0, v)
        sv.checkpoint_restore()
        sv.reset_stats()
        pool.synchronise()
        t0 = time.perf_counter()
        pool.Solve()
        wait_with_progress(pool, [sv], f"N = 256, {name}")
        runs[name] = (time.perf_counter() - t0, sv.stats().copy())
    pool.close()
    best = min(runs, key=lambda k: runs[k][0])
    wall, st = runs[best]
    # Matrix traffic of the coupling solve: one pass over M (N^2 doubles) per BiCG iteration
    # (both products) and one per stage (the right-hand side and the initial residual).
    m_bytes = 8.0 * ens.N * ens.N * (st[:, 3].sum() + st[:, 2].sum())
    achieved = flops_fm1(st, ens.N) / wall / 1e12
    return {"workload": f"{ns} systems x N=256 at random 3D positions in a 1 mm sphere (30 um "
                        f"min separation), per-system coupling, 2-3 bar at 25-50 kHz, RKCK45 "
                        f"rtol=atol=1e-10, tau in [0, {end_time}] from rest",
            "value": ns * end_time / wall, "unit": "system-periods/s (period fraction counted)",
            "ms": wall * 1e3, "variant": best,
            "ms_per_variant": {k: v[0] * 1e3 for k, v in runs.items()},
            "bicg_iterations_per_rhs": iters_per_rhs(st),
            "accepted_steps_per_system": float(st[:, 0].mean()),
            "roofline": {"bound": "hbm", "achieved": m_bytes / wall / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": m_bytes / wall / 1e9 / HBM_PEAK_GBS,
                         "traffic": (measured_traffic(
                             "queue_kernel_wide256_config4_4096sys_trace_pmc.md")
                             if ns == 4096 and end_time == 0.02 else None),
                         "note": "coupling-matrix bytes of the coupling solve (one pass over the "
                                 "system's M per BiCG iteration and one per stage) over the "
                                 "launch; traffic: HBM bytes per launch measured by rocprofv3 "
                                 "for this workload"},
            "fp64": {"achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget", type=float, default=420.0,
                    help="seconds from start after which optional legs are skipped")
    ap.add_argument("--hard-limit", type=float, default=560.0,
                    help="seconds from start at which the completed legs are printed and the "
                         "run ends")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned legs")
    ap.add_argument("--heavy", type=int, default=64,
                    help="systems on the learned start in the learned+batched leg")
    ap.add_argument("--medium", type=int, default=192,
                    help="next costliest systems (measured) on the plain start, one per wave, "
                         "in the learned+batched leg")
    ap.add_argument("--pool-end-time", type=float, default=0.1,
                    help="period fraction of the config-2 pool leg (0 skips it); a full period "
                         "is dominated by the few systems that reach the Mach-1 singularity "
                         "and takes minutes")
    ap.add_argument("--leg-time-limit", type=float, default=None,
                    help="with --leg: seconds after which the leg's remaining systems are asked "
                         "to stop (their state is written back as it stands)")
    ap.add_argument("--attempt-cap", type=int, default=0,
                    help="with --leg pool or period2: MaximumStepAttempts (0: no cap)")
    ap.add_argument("--config4-end-time", type=float, default=0.02,
                    help="period fraction of the N = 256 leg (0 skips it)")
    ap.add_argument("--end-time", type=float, default=1.0,
                    help="fraction of a period per step (profiling only)")
    ap.add_argument("--plain-only", action="store_true",
                    help="headline leg only (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--leg", choices=("pool", "n256", "period2", "split", "cpu"),
                    help="run only this leg (with --pool-end-time / --config4-end-time) and "
                         "print its record instead of the benchmark line")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.leg == "cpu":
        print(json.dumps(cpu_baseline()), flush=True)
        return 0
    if args.leg:
        log(f"{args.leg} leg alone")
        rec = (run_pool_leg(args.pool_end_time, sequential=False, limit=args.leg_time_limit,
                            cap=args.attempt_cap)
               if args.leg == "pool"
               else run_config4_leg(args.config4_end_time) if args.leg == "n256"
               else run_split_leg(args.steps, args.warmup, args.heavy, args.medium)
               if args.leg == "split"
               else run_second_period_leg(limit=args.leg_time_limit, cap=args.attempt_cap))
        print(json.dumps({"leg": args.leg, **rec}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (systems are "
              "independent: a multi-GPU run shards them by estimated cost, one process and one "
              "solver per GPU)", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    from bubbleensemble_implicitextension_amd import batched
    from bubbleensemble_implicitextension_amd import workloads as wl

    out = Output()

    def watchdog():
        if out.emit(note=f"hard limit of {args.hard_limit:.0f} s reached; later legs missing"):
            os._exit(0)
        os._exit(3)

    timer = threading.Timer(max(1.0, args.hard_limit - elapsed()), watchdog)
    timer.daemon = True
    timer.start()

    ens = wl.config1()
    N = ens.N
    K, W = args.steps, args.warmup
    log(f"plain leg: {ens.NS} systems, {W} warmup + {K} timed from-rest periods")
    plain = run_leg(ens, K, W, args.end_time,
                    stream_steps=0 if args.plain_only else min(2, K))
    stats = plain["stats"]
    kernel_ms = plain["kernel_ms"]
    wall = plain["wall"]
    flops = flops_fm1(stats, N)
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    value = ens.NS * K / wall
    log(f"plain leg done: {wall * 1e3 / K:.1f} ms per period")

    line = {
        "metric": "systems\u00d7acoustic periods/sec, learned vs plain at rtol 1e-10; BiCG "
                  "iters/RHS",
        "metric_note": "value: the plain FP64 RKCK45 + BiCG method at rtol=atol=1e-10; the "
                       "learned warm start in 'learned', learned+batched in 'learned_batched', "
                       "BiCG iterations per RHS in 'bicg_iterations_per_rhs'",
        "value": value,
        "unit": "system-periods/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry, radii and forcing grid)",
        "config": {"workload": "16384 systems x N=32 KM bubbles, 128x128 PA x f grid, RKCK45 "
                               "rtol=atol=1e-10, one from-rest period per step (state restored "
                               "on the device before each step), dense output 64 Hermite "
                               "points per period kept in HBM",
                   "systems": ens.NS, "units_per_system": N, "end_time": args.end_time,
                   "dense_points_per_period": DENSE_POINTS, "parallelism": "1 GPU"},
        "bicg_iterations_per_rhs": iters_per_rhs(stats),
        "bicg_iterations_per_rhs_split": tail_split(stats),
        "steps_accepted_per_system_period": float(stats[:, 0].sum()) / (ens.NS * K),
        "roofline": {"bound": "mfma", "compute_pipe": "fp64 vector ALU (v_fma_f64; MI355X's "
                                                      "FP64 vector and matrix peaks are equal)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic() if args.end_time == 1.0 else None,
                     "kernel": KERNEL_NAME,
                     "kernel_ms_per_launch": kernel_ms / K,
                     "restore_ms_per_step": plain["restore_ms"] / K,
                     "flops_per_launch": flops / K},
    }
    if "stream_wall" in plain:
        line["pcie_inclusive"] = {
            "value": ens.NS * plain["stream_steps"] / plain["stream_wall"],
            "unit": "system-periods/s",
            "note": f"{plain['stream_steps']} periods, each followed by a copy of its dense "
                    "output (index, times, 16384x32x2x64 states) to pinned host memory"}
    out.line = line
    plain_s = wall + plain["kernel_ms"] * 1e-3 * W / K + 5.0  # a leg's rough cost

    def fits(cost, name):
        if elapsed() + cost <= args.budget:
            return True
        line.setdefault("skipped", {})[name] = (
            f"wall budget: {elapsed():.0f} s used, ~{cost:.0f} s needed, budget {args.budget:.0f} s")
        log(f"{name} skipped (budget)")
        return False

    if not args.no_cpu_baseline and not args.plain_only and fits(40.0, "cpu_baseline"):
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline_child(args.hard_limit - elapsed() - 5.0)
        log("cpu baseline done")

    have_net = (not args.no_learned and not args.plain_only
                and os.path.exists(batched.DEFAULT_NETWORK))
    net = np.load(batched.DEFAULT_NETWORK) if have_net else None
    if have_net and fits(plain_s, "learned"):
        log("learned leg")
        learned = run_leg(ens, K, W, args.end_time, network=net)
        ls = learned["stats"]
        lv = ens.NS * K / learned["wall"]
        line["learned"] = {
            "value": lv, "unit": "system-periods/s",
            "ms_per_step": learned["wall"] * 1e3 / K,
            "kernel_ms_per_launch": learned["kernel_ms"] / K,
            "bicg_iterations_per_rhs": iters_per_rhs(ls),
            "bicg_iterations_per_rhs_split": tail_split(ls),
            "warm_fallbacks": int(ls[:, 31].sum()),
            "speedup_vs_plain": lv / value,
            "network": "5-64-64-1 BF16 MFMA, trained on a held-out-seed ensemble",
            "bicg_residual": 1e-6,
            "error_vs_plain": error_vs(plain, learned),
        }
        del learned
        log("learned leg done")
    # Learned + batched: three tiers in two launches (batched.LearnedBatchedSolver). The
    # medium part is ranked by the cost each system measured in the plain leg's solve of this
    # period (its counters); in a run of consecutive periods, the previous period's counters.
    measured = batched.measured_cost(stats, N)

    def split_leg(network):
        return run_split(ens, network, args.heavy, args.medium, K, W, args.end_time, measured)

    if have_net and fits(plain_s, "learned_batched"):
        log("learned+batched leg")
        lwall, lbr, parts = split_leg(net)
        heavy = parts["heavy"]
        lbs = lbr["stats"]
        lbv = ens.NS * K / lwall
        line["learned_batched"] = {
            "value": lbv, "unit": "system-periods/s",
            "ms_per_step": lwall * 1e3 / K,
            "speedup_vs_plain": lbv / value,
            "split": split_text(parts, True),
            "learned_systems": [int(k) for k in heavy[:8]] + (["..."] if len(heavy) > 8 else []),
            "bicg_iterations_per_rhs": iters_per_rhs(lbs),
            "bicg_iterations_per_rhs_learned_systems": iters_per_rhs(lbs, heavy),
            "bicg_iterations_per_rhs_learned_systems_plain_start": iters_per_rhs(stats, heavy),
            "fp64": fp64_rate(lbs, N, lwall),
            "error_vs_plain": error_vs(plain, lbr),
        }
        log("learned+batched leg done")
    if not args.plain_only and fits(plain_s, "plain_batched"):
        # The same split with the plain start everywhere: bit-identical to the plain leg, and
        # bound by the same stiffest system, which only the learned start shortens.
        log("plain batched leg")
        pwall, pbr, parts = split_leg(None)
        same = all(np.array_equal(pbr[k].view(np.int64) if pbr[k].dtype.kind == "f" else pbr[k],
                                  plain[k].view(np.int64) if plain[k].dtype.kind == "f" else plain[k])
                   for k in ("state", "acc", "events"))
        pbv = ens.NS * K / pwall
        line["plain_batched"] = {
            "value": pbv, "unit": "system-periods/s", "ms_per_step": pwall * 1e3 / K,
            "speedup_vs_plain": pbv / value, "split": split_text(parts, False),
            "fp64": fp64_rate(pbr["stats"], N, pwall),
            "bit_identical_to_plain": bool(same)}
        log("plain batched leg done")

    # Config 3: the same shapes with BiCG run to an absolute residual of 1e-13.
    k3 = min(K, 3)
    c3_opts = {"BiCGResidualTolerance": 1e-13, "BiCGMaxIterations": 60}
    if not args.plain_only and fits(2.0 * plain_s * k3 / K, "bicg_residual_1e-13"):
        log("residual 1e-13 legs")
        p3 = run_leg(ens, k3, 0, args.end_time, options=c3_opts)
        v3 = ens.NS * k3 / p3["wall"]
        c3 = {"workload": "headline workload with BiCG to max|r| < 1e-13 (at most 60 "
                          "iterations)",
              "steps": k3,
              "plain": {"value": v3, "unit": "system-periods/s",
                        "ms_per_step": p3["wall"] * 1e3 / k3,
                        "bicg_iterations_per_rhs": iters_per_rhs(p3["stats"]),
                        "bicg_iterations_per_rhs_split": tail_split(p3["stats"]),
                        "bicg_maxit_stages": int(p3["stats"][:, 5].sum())}}
        if have_net and fits(2.0 * plain_s * k3 / K, "bicg_residual_1e-13_learned"):
            l3 = run_leg(ens, k3, 0, args.end_time, network=net, options=c3_opts)
            lv3 = ens.NS * k3 / l3["wall"]
            c3["learned"] = {"value": lv3, "unit": "system-periods/s",
                             "ms_per_step": l3["wall"] * 1e3 / k3,
                             "speedup_vs_plain": lv3 / v3,
                             "bicg_iterations_per_rhs": iters_per_rhs(l3["stats"]),
                             "bicg_iterations_per_rhs_split": tail_split(l3["stats"]),
                             "warm_fallbacks": int(l3["stats"][:, 31].sum()),
                             "error_vs_plain": error_vs(p3, l3)}
            del l3
        del p3
        line["bicg_residual_1e-13"] = c3
        log("residual 1e-13 legs done")

    if not args.plain_only and fits(40.0, "second_period_capped"):
        # The steady state: period 2 costs ~5x period 1, and a few systems of the stiff corner
        # reach the Mach-1 singularity there; they end at an attempt cap of 1e5.
        log("second period leg (attempt cap 1e5)")
        line["second_period_capped"] = run_second_period_leg(limit=60.0, cap=100000)
        log("second period leg done")
    if args.pool_end_time > 0 and not args.plain_only and fits(60.0, "pool"):
        log("pool leg")
        line["pool"] = run_pool_leg(args.pool_end_time,
                                    sequential=elapsed() < 0.5 * args.budget)
        log(f"pool leg done ({line['pool']['ms']:.0f} ms)")
    if args.config4_end_time > 0 and not args.plain_only and fits(60.0, "config4"):
        log("N = 256 leg")
        line["clusters_n256"] = run_config4_leg(args.config4_end_time)
        log(f"N = 256 leg done ({line['clusters_n256']['ms']:.0f} ms)")

    timer.cancel()
    out.emit()
    return 0


if __name__ == "__main__":
    sys.exit(main())
