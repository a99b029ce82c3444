"""Collapse-strength harness on top of ProblemSolver.

Restates the reference's collapse-strength driver (CollapseStrength.cu:68-204): fill the
solver from a workload, run transient periods, reset the collapse accessories, run recording
periods, and write one line per (system, collapse) with the per-bubble maxima and minima of
R/R_E. Each period is one Solve(); the Initialization hook restarts every system at TD[0].
"""
from __future__ import annotations

import numpy as np

from .problem_solver import ListOfSolverOptions as O
from .problem_solver import ListOfVariables as V
from .problem_solver import ProblemSolver
from . import workloads as wl


def make_solver(ens, NDO=0, device=0, algorithm=1, **extra):
    return ProblemSolver(ens.NS, ens.N, NUP=wl.NUP, NUA=wl.NUA, NiUA=wl.NIUA, NE=1, NDO=NDO,
                         Algorithm=algorithm, device=device, **extra)


def fill(solver, ens):
    """Copies a workload into the solver's host buffers (FillSolverObject and
    FillCouplingMatrix, CollapseStrength.cu:279-371, 398-408) and applies its options."""
    NS = ens.NS
    td = solver.host_view(V.TimeDomain)
    td[:NS] = 0.0
    td[NS:] = 1.0
    solver.host_view(V.ActualTime)[:] = 0.0
    solver.host_view(V.ActualState)[:] = ens.unit_soa(ens.state0).reshape(-1)
    solver.host_view(V.UnitParameters)[:] = ens.unit_soa(ens.params).reshape(-1)
    solver.host_view(V.CouplingStrength)[:] = 1.0
    solver.host_view(V.CouplingIndex)[:] = 1
    reset_accessories(solver)
    if ens.coupling.ndim != 2:
        apply_options(solver, ens.options)
        return  # per-system matrices: ProblemPool.set_couplings
    N = ens.N
    for row in range(N):
        for col in range(N):
            solver.SetHost(0, V.CouplingMatrix, row, col, float(ens.coupling[row, col]))
    apply_options(solver, ens.options)


def apply_options(solver, options):
    for name, value in options.items():
        opt = O[name]
        if isinstance(value, (list, tuple, np.ndarray)):
            for i, v in enumerate(value):
                solver.SolverOption(opt, i, v)
        else:
            solver.SolverOption(opt, value)


def reset_accessories(solver):
    """ResetAccessories (CollapseStrength.cu:373-396)."""
    solver.host_view(V.UnitAccessories)[:] = 0.0
    solver.host_view(V.IntegerUnitAccessories)[:] = -1


def run_periods(solver, periods):
    for _ in range(periods):
        solver.Solve()
    solver.InsertSynchronisationPoint()
    solver.SynchroniseSolver()


def collapse_protocol(solver, transient, record):
    """Transient periods, accessory reset, recording periods (CollapseStrength.cu:120-151)."""
    solver.SynchroniseFromHostToDevice(V.All)
    run_periods(solver, transient)
    reset_accessories(solver)
    solver.SynchroniseFromHostToDevice(V.UnitAccessories)
    solver.SynchroniseFromHostToDevice(V.IntegerUnitAccessories)
    run_periods(solver, record)
    solver.SynchroniseFromDeviceToHost(V.All)
    solver.InsertSynchronisationPoint()
    solver.SynchroniseSolver()


def write_collapse_file(solver, path, ncoll=wl.NCOLL):
    """Result file of CollapseStrength.cu:173-198: per system and collapse index, PA, f, then
    the N maxima and the N minima of R/R_E."""
    acc = solver.accessories()  # [NS, N, NUA]
    par = solver.host_view(V.UnitParameters).reshape(-1, solver.NS, solver.UPS)
    with open(path, "w") as f:
        for sid in range(solver.NS):
            for c in range(ncoll):
                fields = [f"{sid:6d}", f"{par[15, sid, 0]:17.10e}", f"{par[16, sid, 0]:17.10e}"]
                fields += [f"{v:10.3e}" for v in acc[sid, :, c]]
                fields += [f"{v:10.3e}" for v in acc[sid, :, c + ncoll]]
                f.write(",".join(fields) + ",\n")


def make_pool(ensembles, device=0):
    """ProblemPool with one size class per Ensemble (per-system coupling matrices), filled."""
    from .problem_solver import ProblemPool

    pool = ProblemPool([e.N for e in ensembles], [e.NS for e in ensembles], NUP=wl.NUP,
                       NUA=wl.NUA, NiUA=wl.NIUA, NE=1, device=device)
    for c, e in enumerate(ensembles):
        s = pool.solver(c)
        fill(s, e)
        M = e.coupling if e.coupling.ndim == 3 else np.broadcast_to(e.coupling, (e.NS, e.N, e.N))
        pool.set_couplings(c, M)
    return pool


def stream_dense_output(solver, periods, sink):
    """Runs `periods` Solve() calls and hands each period's dense output to
    sink(period, times [n, NS], states [n, UD, NS*UPS]) before the next period reuses the
    buffers (the dense index is reset to 0 after each period). With interpolation on
    (DenseOutputInterpolation = 1) every period yields the same uniform instants."""
    NS, UPS, UD = solver.NS, solver.UPS, solver.UD
    TLT = NS * UPS
    for k in range(periods):
        solver.Solve()
        solver.SynchroniseFromDeviceToHost(V.DenseOutput)
        solver.InsertSynchronisationPoint()
        solver.SynchroniseSolver()
        idx = solver.host_view(V.DenseIndex)
        n = int(idx.max()) if idx.size else 0
        times = solver.host_view(V.DenseTime)[: n * NS].reshape(n, NS).copy()
        states = solver.host_view(V.DenseState)[: n * UD * TLT].reshape(n, UD, TLT).copy()
        sink(k, times, states)
        idx[:] = 0
        solver.SynchroniseFromHostToDevice(V.DenseOutput)  # index, times and states
