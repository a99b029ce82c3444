"""GPU tests of pools of uneven ensembles (per-system coupling matrices, persistent kernels with
device work queues, one hipGraph per period, autotuned launch variants).

Every system of every class is compared bit for bit with the oracle's canonical mode run on
that system alone with its own coupling matrix: scheduling (queue order, variant, graph replay)
must not change any result.
"""
import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import collapse
from bubbleensemble_implicitextension_amd import workloads as wl
from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V
from gpu_helpers import assert_bitwise

pytestmark = pytest.mark.gpu

UNITS = (8, 16, 32, 64, 128)
SYSTEMS = (5, 4, 3, 3, 2)
END = 0.1  # fraction of a period (keeps the N = 128 oracle runs short)


def _classes(seed=11):
    return [wl.config2_class(n, ns, seed=seed) for n, ns in zip(UNITS, SYSTEMS)]


def _pool(classes, end=END):
    pool = collapse.make_pool(classes)
    for c, e in enumerate(classes):
        s = pool.solver(c)
        s.host_view(V.TimeDomain)[e.NS:] = end
        s.SynchroniseFromHostToDevice(V.All)
    return pool


def _results(pool, classes):
    pool.synchronise()
    out = []
    for c in range(len(classes)):
        s = pool.solver(c)
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out.append({"state": s.host_view(V.ActualState).copy(),
                    "time": s.host_view(V.ActualTime).copy(),
                    "acc": s.host_view(V.UnitAccessories).copy(),
                    "iacc": s.host_view(V.IntegerUnitAccessories).copy(),
                    "stats": s.stats().copy()})
    return out


def _oracle_system(oracle, ens, sid, periods, end):
    sub = ens.subset([sid])
    sub.coupling = sub.coupling[0]
    p = oracle.OracleProblem.from_ensemble(sub)
    p.TimeDomain[1:] = end
    for _ in range(periods):
        p.solve(mode=oracle.CANONICAL)
    return p


def _check_against_oracle(oracle, classes, res, periods, end):
    for c, e in enumerate(classes):
        r = res[c]
        NS, N = e.NS, e.N
        state = r["state"].reshape(2, NS, N)
        acc = r["acc"].reshape(-1, NS, N)
        iacc = r["iacc"].reshape(-1, NS, N)
        for sid in range(NS):
            p = _oracle_system(oracle, e, sid, periods, end)
            tag = f"N={N} sid={sid}"
            assert_bitwise(tag + " state", state[:, sid, :], p.ActualState.reshape(2, 1, N)[:, 0])
            assert_bitwise(tag + " time", r["time"][sid], p.ActualTime[0])
            assert_bitwise(tag + " acc", acc[:, sid, :], p.UnitAccessories.reshape(-1, 1, N)[:, 0])
            assert_bitwise(tag + " iacc", iacc[:, sid, :],
                           p.IntegerUnitAccessories.reshape(-1, 1, N)[:, 0])
            assert_bitwise(tag + " stats", r["stats"][sid], p.stats_table()[0])


def test_pool_matches_oracle_every_variant(oracle):
    classes = _classes()
    probe = _pool(classes)
    nvar = max(len(probe.variants(c)) for c in range(len(classes)))
    probe.close()
    for v in range(nvar):
        print(f"case: pool, variant index {v} (clamped per class), end {END}", flush=True)
        pool = _pool(classes)
        for c in range(len(classes)):
            pool.set_variant(c, min(v, len(pool.variants(c)) - 1))
        pool.Solve()
        res = _results(pool, classes)
        pool.close()
        _check_against_oracle(oracle, classes, res, 1, END)


def test_pool_graph_replays_and_order(oracle):
    """Three periods as one graph-replay call equal three single-period calls, with reversed
    and cost-based hand-out orders."""
    classes = _classes(seed=12)
    print("case: pool, 3 periods by replay", flush=True)
    pool = _pool(classes, end=0.05)
    for c, e in enumerate(classes):
        pool.set_order(c, np.arange(e.NS)[::-1])
    pool.solve_periods(3)
    a = _results(pool, classes)
    pool.close()
    print("case: pool, 3 single periods, cost order", flush=True)
    pool = _pool(classes, end=0.05)
    for _ in range(3):
        pool.Solve()
        pool.synchronise()
        for c in range(len(classes)):
            pool.order_by_cost(c)
    b = _results(pool, classes)
    pool.close()
    for c in range(len(classes)):
        for k in a[c]:
            assert_bitwise(f"class {c} {k}", a[c][k], b[c][k])
    _check_against_oracle(oracle, classes, a, 3, 0.05)


def test_pool_autotune_restores_state():
    classes = _classes(seed=13)
    print("case: pool autotune", flush=True)
    pool = _pool(classes)
    pool.synchronise()
    before = []
    for c in range(len(classes)):
        s = pool.solver(c)
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        before.append(s.host_view(V.ActualState).copy())
    times = pool.autotune(0.05)
    for c in range(len(classes)):
        assert set(times[c]) == set(pool.variants(c))
        assert all(t > 0 for t in times[c].values())
        best = min(times[c], key=times[c].get)
        assert pool.variants(c)[pool.variant(c)] == best
        s = pool.solver(c)
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        assert_bitwise(f"class {c} state after autotune", s.host_view(V.ActualState), before[c])
        assert not s.stats().any()
    # The autotuner leaves the queues in longest-probe-first order: the same results.
    pool.Solve()
    tuned = _results(pool, classes)
    pool.close()
    print("case: pool without autotune", flush=True)
    ref = _pool(classes)
    ref.Solve()
    plain = _results(ref, classes)
    ref.close()
    for c in range(len(classes)):
        for key in tuned[c]:
            assert_bitwise(f"class {c} {key} (autotuned)", tuned[c][key], plain[c][key])


def test_config4_class_matches_oracle(oracle):
    """Config-4 shape: N = 256 clusters in a 1 mm sphere, per-system coupling streamed from
    global memory block by block (one lane per unit: 32 x 32 blocks; two: 64 x 64); first 2 %
    of a period against the oracle, system by system, for every launch variant."""
    classes = [wl.config4(ns=2, seed=5)]
    pool = _pool(classes, end=0.02)
    s = pool.solver(0)
    s.checkpoint_save()
    for v, name in enumerate(pool.variants(0)):
        print(f"case: config 4 class, N=256, NS=2, end 0.02, variant {name}", flush=True)
        pool.set_variant(0, v)
        s.checkpoint_restore()
        s.reset_stats()
        pool.Solve()
        res = _results(pool, classes)
        _check_against_oracle(oracle, classes, res, 1, 0.02)
    pool.close()


def test_config4_sixteen_systems_digests():
    """Sixteen config-4 systems (N = 256, per-system coupling streamed from global memory) over
    the first tenth of a period, one pool class: every system's state, time, accessories and
    counters equal the oracle's, through the per-system digests the oracle wrote
    (tests/golden/make_config4_digests.py)."""
    import json
    import os
    import sys

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    from digests import system_digests

    want = json.load(open(os.path.join(here, "config4_digests.json")))
    ens = wl.config4(ns=want["ns"], seed=want["seed"])
    print(f"case: config 4, {ens.NS} systems x N=256, tau in [0, {want['end_time']}]", flush=True)
    pool = _pool([ens], end=want["end_time"])
    pool.Solve()
    res = _results(pool, [ens])[0]
    pool.close()
    got = system_digests(ens.NS, ens.N, res["state"], res["time"], res["acc"], res["iacc"],
                         res["stats"])
    bad = [k for k in range(ens.NS) if got[k] != want["digests"][k]]
    assert not bad, f"systems differing from the oracle: {bad}"
    assert int((res["iacc"] + 1).sum()) == want["events_total"]


def test_config2_five_classes_digests():
    """Sixteen systems of each config-2 size class (N = 8 ... 128), one pool launch over the
    first quarter of a period (autotuned variants and queue order): every system's state, time,
    accessories and counters equal the oracle's, through the per-system digests the oracle
    wrote (tests/golden/make_config2_digests.py)."""
    import json
    import os
    import sys

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    from digests import system_digests

    want = json.load(open(os.path.join(here, "config2_digests.json")))
    classes = [wl.config2_class(n, want["ns"], want["seed"]) for n in want["units"]]
    print(f"case: config 2, {len(classes)} classes x {want['ns']} systems, "
          f"tau in [0, {want['end_time']}]", flush=True)
    pool = _pool(classes, end=want["end_time"])
    for c, e in enumerate(classes):
        pool.solver(c).checkpoint_save()
    pool.autotune()
    for c, e in enumerate(classes):
        pool.solver(c).checkpoint_restore()
        pool.solver(c).reset_stats()
    pool.Solve()
    res = _results(pool, classes)
    pool.close()
    events = 0
    for c, e in enumerate(classes):
        r = res[c]
        got = system_digests(e.NS, e.N, r["state"], r["time"], r["acc"], r["iacc"], r["stats"])
        bad = [k for k in range(e.NS) if got[k] != want["digests"][str(e.N)][k]]
        assert not bad, f"N={e.N}: systems differing from the oracle: {bad}"
        events += int((r["iacc"] + 1).sum())
    assert events == want["events_total"]


def test_class_solver_solve_uses_its_own_matrices(oracle):
    """Solve() on a pool class's own solver runs that class alone with each system's matrix
    (not the first system's): every system matches the oracle."""
    classes = [wl.config2_class(64, 3, seed=14), wl.config2_class(16, 5, seed=14)]
    print("case: class solver Solve(), N=64 and N=16", flush=True)
    pool = _pool(classes)
    for c in range(len(classes)):
        pool.solver(c).Solve()
    res = _results(pool, classes)
    pool.close()
    _check_against_oracle(oracle, classes, res, 1, END)
