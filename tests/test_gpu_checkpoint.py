"""GPU tests of the device checkpoint (bex_checkpoint_save/restore) and of the dense-output
index bounds on the host side of the boundary.

A period solved again from a restored checkpoint must reproduce the first solve bit for bit
(state, accessories, event counters, dense output) and both must equal the CPU oracle; this is
what the benchmark relies on when it times repeated from-rest periods.
"""
import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import collapse
from bubbleensemble_implicitextension_amd import workloads as wl
from bubbleensemble_implicitextension_amd._native import BexError
from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

from gpu_helpers import assert_bitwise, run_oracle

pytestmark = pytest.mark.gpu

BEX_E_ARGUMENT, BEX_E_BOUND = -1, -2


def _solver(ens, ndo, end):
    s = collapse.make_solver(ens, NDO=ndo)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = end
    s.SolverOption(O.DenseOutputInterpolation, 1)
    s.SolverOption(O.DenseOutputMinimumTimeStep, end / ndo)
    s.SynchroniseFromHostToDevice(V.All)
    return s


def _snapshot(s):
    s.SynchroniseFromDeviceToHost(V.All)
    s.SynchroniseFromDeviceToHost(V.DenseOutput)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    return {v: s.host_view(v).copy() for v in (V.ActualState, V.ActualTime, V.UnitAccessories,
                                               V.IntegerUnitAccessories, V.DenseIndex,
                                               V.DenseTime, V.DenseState)}


def test_restored_period_repeats_bit_for_bit(oracle):
    ens = wl.config0().subset([9, 18, 27, 36])
    ndo, end = 16, 0.25
    print(f"case: checkpoint restore, {ens.NS} systems, NDO={ndo}, tau in [0, {end}]", flush=True)
    s = _solver(ens, ndo, end)
    s.checkpoint_save()
    s.Solve()
    first = _snapshot(s)
    st1 = s.stats()
    assert first[V.DenseIndex].min() == ndo
    # A second period continues from the first one's end (different results) ...
    s.Solve()
    second = _snapshot(s)
    assert not np.array_equal(second[V.ActualState], first[V.ActualState])
    # ... and a restore brings back the from-rest period exactly, twice in a row.
    for rep in range(2):
        s.reset_stats()
        s.checkpoint_restore()
        s.Solve()
        again = _snapshot(s)
        for v, want in first.items():
            assert_bitwise(f"{v.name} after restore {rep}", again[v], want)
        assert_bitwise("stats after restore", s.stats(), st1)
    s.close()

    p = run_oracle(oracle, ens, periods=1, ndo=ndo, end_time=end,
                   options={"DenseOutputInterpolation": 1, "DenseOutputMinimumTimeStep": end / ndo})
    assert_bitwise("ActualState vs oracle", first[V.ActualState], p.ActualState)
    assert_bitwise("DenseState vs oracle", first[V.DenseState], p.DenseOutputStates)
    assert_bitwise("DenseTime vs oracle", first[V.DenseTime], p.DenseOutputTimeInstances)


def test_restore_without_checkpoint_fails():
    ens = wl.config0().subset([9])
    s = _solver(ens, 4, 0.1)
    with pytest.raises(BexError) as e:
        s.checkpoint_restore()
    assert e.value.code == BEX_E_ARGUMENT
    s.close()


def test_dense_index_bounds_are_checked_on_the_host():
    ens = wl.config0().subset([9, 18])
    ndo = 4
    s = _solver(ens, ndo, 0.1)
    for bad in (-1, ndo + 1, float("nan")):
        with pytest.raises(BexError) as e:
            s.SetHost(0, V.DenseIndex, bad)
        assert e.value.code == BEX_E_BOUND
    s.SetHost(1, V.DenseIndex, ndo)  # a full buffer is a valid index
    assert s.GetHost(1, V.DenseIndex) == ndo
    # Values written through the host view are checked before they reach the device.
    idx = s.host_view(V.DenseIndex)
    idx[0] = -3
    for var in (V.DenseOutput, V.All):
        with pytest.raises(BexError) as e:
            s.SynchroniseFromHostToDevice(var)
        assert e.value.code == BEX_E_BOUND
    idx[0] = 0
    s.SynchroniseFromHostToDevice(V.All)
    s.Solve()
    s.SynchroniseFromDeviceToHost(V.DenseOutput)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    # System 1 started full: it stores nothing and keeps its index; system 0 fills its buffer.
    assert list(s.host_view(V.DenseIndex)) == [ndo, ndo]
    s.close()


def test_timeline_and_progress_peek():
    """The per-system timeline sits in host-mapped memory: finished systems can be counted
    without a GPU command; enabling clears it, disabling frees it."""
    ens = wl.config0().subset([9, 18, 27, 36])
    print(f"case: timeline, {ens.NS} systems, tau in [0, 0.1]", flush=True)
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = 0.1
    s.SynchroniseFromHostToDevice(V.All)
    s.enable_timeline(True)
    assert s.finished_systems() == 0
    s.Solve()
    s.SynchroniseDevice()
    assert s.finished_systems() == ens.NS
    t = s.timeline()
    assert np.all(t[:, 1] > t[:, 0]) and np.all(t[:, 0] >= 0.0)
    s.enable_timeline(True)
    assert s.finished_systems() == 0
    s.enable_timeline(False)
    with pytest.raises(BexError):
        s.finished_systems()
    s.close()


def test_stop_request_ends_systems_at_their_next_check():
    """A stop requested before the launch ends every system at its first check, after exactly
    255 step attempts, with ActualTime short of the time domain's end; systems that finish in
    fewer attempts are not affected. After clear_stop the same problem runs to the end."""
    ens = wl.config0().subset([9, 18, 27, 36, 0])
    print(f"case: stop request, {ens.NS} systems, tau in [0, 0.3]", flush=True)
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = 0.3
    s.SynchroniseFromHostToDevice(V.All)
    s.checkpoint_save()
    s.request_stop()
    s.Solve()
    s.SynchroniseFromDeviceToHost(V.ActualTime)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    st = s.stats()
    t_end = s.host_view(V.ActualTime).copy()
    stopped = t_end < 0.3 - 1e-9
    assert stopped.any()
    assert np.all(st[stopped, 10] == 255), st[:, 10]
    assert np.all(st[~stopped, 10] <= 255)
    s.clear_stop()
    s.reset_stats()
    s.checkpoint_restore()
    s.Solve()
    s.SynchroniseFromDeviceToHost(V.ActualTime)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    assert np.all(np.abs(s.host_view(V.ActualTime) - 0.3) < 1e-12)
    assert np.all(s.stats()[stopped, 10] > 255)
    s.close()


def test_attempt_cap_ends_long_systems_only():
    """MaximumStepAttempts = 1100: a system still running at its first check past the cap
    (the fifth, after 1279 attempts) ends there with its state and time as they stand; systems
    that finish sooner are bit-identical to an uncapped solve. Negative caps are rejected."""
    ens = wl.config0().subset([9, 18, 27, 36, 0])
    print(f"case: attempt cap 1100, {ens.NS} systems, tau in [0, 0.3]", flush=True)
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = 0.3
    s.SynchroniseFromHostToDevice(V.All)
    s.checkpoint_save()
    s.Solve()
    free = _snapshot(s)
    free_stats = s.stats().copy()
    attempts = free_stats[:, 10]
    s.reset_stats()
    s.checkpoint_restore()
    s.SolverOption(O.MaximumStepAttempts, 1100)
    s.Solve()
    capped = _snapshot(s)
    st = s.stats()
    long = attempts >= 1280
    assert long.any() and (~long).any(), attempts
    assert np.all(st[long, 10] == 1279), st[:, 10]
    assert np.all(capped[V.ActualTime][long] < 0.3 - 1e-9)
    state_free = free[V.ActualState].reshape(2, ens.NS, ens.N)
    state_capped = capped[V.ActualState].reshape(2, ens.NS, ens.N)
    assert_bitwise("uncapped systems", state_capped[:, ~long], state_free[:, ~long])
    assert_bitwise("uncapped stats", st[~long], free_stats[~long])
    with pytest.raises(BexError) as e:
        s.SolverOption(O.MaximumStepAttempts, -1)
    assert e.value.code == BEX_E_ARGUMENT
    s.close()
