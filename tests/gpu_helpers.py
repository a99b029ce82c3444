"""Shared helpers for tests that run the gfx950 library next to the CPU oracle."""
import numpy as np

from bubbleensemble_implicitextension_amd import collapse
from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V


def run_gpu(ens, periods, algorithm=1, options=None, ndo=0, reset_after=None, end_time=1.0,
            layout=0, setup=None):
    s = collapse.make_solver(ens, algorithm=algorithm, NDO=ndo)
    collapse.fill(s, ens)
    s.SolverOption(O.KernelLayout, layout)
    s.host_view(V.TimeDomain)[ens.NS:] = end_time
    if options:
        collapse.apply_options(s, options)
    if setup:
        setup(s)
    s.SynchroniseFromHostToDevice(V.All)
    for k in range(periods):
        if reset_after is not None and k == reset_after:
            s.InsertSynchronisationPoint()
            s.SynchroniseSolver()
            collapse.reset_accessories(s)
            s.SynchroniseFromHostToDevice(V.UnitAccessories)
            s.SynchroniseFromHostToDevice(V.IntegerUnitAccessories)
        s.Solve()
    s.SynchroniseFromDeviceToHost(V.All)
    if ndo:
        s.SynchroniseFromDeviceToHost(V.DenseOutput)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    return s


def run_oracle(oracle, ens, periods, mode=0, algorithm=1, options=None, ndo=0, reset_after=None,
               nthreads=0, end_time=1.0):
    p = oracle.OracleProblem.from_ensemble(ens, NDO=ndo)
    p.TimeDomain[ens.NS:] = end_time
    p.opts["Algorithm"] = algorithm
    if options:
        p.opts.update(options)
    for k in range(periods):
        if reset_after is not None and k == reset_after:
            p.reset_accessories()
        p.solve(mode=mode, nthreads=nthreads)
    return p


def assert_bitwise(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    if got.dtype.kind == "f":
        same = (got.view(np.int64) == want.view(np.int64)) | (np.isnan(got) & np.isnan(want))
    else:
        same = got == want
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        raise AssertionError(f"{name}: {len(idx)} of {got.size} differ; first at {first}: "
                             f"gpu={got[first]!r} oracle={want[first]!r}")


def compare_solver_to_oracle(s, p, stats=True):
    assert_bitwise("ActualState", s.host_view(V.ActualState), p.ActualState)
    assert_bitwise("ActualTime", s.host_view(V.ActualTime), p.ActualTime)
    assert_bitwise("UnitAccessories", s.host_view(V.UnitAccessories), p.UnitAccessories)
    assert_bitwise("IntegerUnitAccessories", s.host_view(V.IntegerUnitAccessories),
                   p.IntegerUnitAccessories)
    if stats:
        assert_bitwise("stats", s.stats(), p.stats_table())
