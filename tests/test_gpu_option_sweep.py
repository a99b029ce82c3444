"""GPU parity over seeded random solver options (CPU oracle, canonical mode, bit for bit).

Each case draws, from a fixed seed, a small ensemble of 32-unit systems (an odd or even count,
forcing amplitude and frequency across the grid's range) and a set of solver options the
reference exposes (tolerances per component, step limits, event direction and tolerance,
dense output with its save frequency and minimum spacing, ActiveSystems, the BiCG constants,
RKCK45 or RK4), and solves it for two periods with every 32-unit layout (one system per wave,
two per wave, lane groups) and both launch modes (persistent work queue, one workgroup per
system group). State, time, accessories, event counts, counters and dense output must equal the
oracle's in every combination.
"""
import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import workloads as wl
from bubbleensemble_implicitextension_amd.problem_solver import (LAYOUT_GROUP, LAYOUT_ROLE_PAIR,
                                                                  LAYOUT_SYSTEM_PAIR)
from bubbleensemble_implicitextension_amd.problem_solver import ListOfSolverOptions as O
from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

from gpu_helpers import assert_bitwise, compare_solver_to_oracle, run_gpu, run_oracle

SEEDS = [101, 202, 303, 404, 505, 606]


def draw_case(seed, n=32):
    """A seeded ensemble of n-unit systems and an option set (host-side only)."""
    rng = np.random.default_rng(seed)
    ns = int(rng.integers(3, 10))
    pa = rng.uniform(0.5, 1.6, size=ns)
    f = rng.uniform(40.0, 200.0, size=ns)
    _, re_um, d = wl.line_geometry(n, seed=seed)
    params = wl.unit_parameters(pa[:, None], f[:, None], re_um[None, :])
    state0 = np.zeros((ns, n, 2))
    state0[..., 0] = 1.0
    M = wl.coupling_from_geometry(re_um, d)
    ens = wl.Ensemble(f"sweep{seed}", n, pa, f, re_um, params, M, state0, wl.default_options(), 2)
    alg = int(rng.integers(0, 2))
    tol = float(10.0 ** rng.uniform(-10, -7))
    opts = {
        "InitialTimeStep": float(10.0 ** rng.uniform(-4, -2)) if alg == 1 else 2e-4,
        "MinimumTimeStep": 1e-13,
        "MaximumTimeStep": float(rng.choice([1e6, 0.05])),
        "TimeStepGrowLimit": float(rng.uniform(2.0, 5.0)),
        "TimeStepShrinkLimit": float(rng.uniform(0.1, 0.4)),
        "RelativeTolerance": [tol, tol * float(rng.uniform(0.5, 2.0))],
        "AbsoluteTolerance": [tol, tol * float(rng.uniform(0.5, 2.0))],
        "EventDirection": [int(rng.integers(-1, 2))],
        "EventTolerance": [float(10.0 ** rng.uniform(-6, -3))],
        "ActiveSystems": int(rng.integers(max(1, ns - 2), ns + 1)),
        "BiCGResidualTolerance": float(rng.choice([1e-6, 1e-9])),
        "BiCGMaxIterations": int(rng.choice([19, 40])),
        "DenseOutputSaveFrequency": int(rng.integers(1, 4)),
        "DenseOutputMinimumTimeStep": float(rng.choice([0.0, 0.01, 0.05])),
        "DenseOutputInterpolation": int(rng.integers(0, 2)),
    }
    if opts["DenseOutputInterpolation"] == 1:
        opts["DenseOutputMinimumTimeStep"] = 0.05
    end = 0.5 if alg == 1 else 0.2
    return ens, alg, opts, end


def test_cases_are_varied():
    """The drawn cases cover both steppers, every event direction and both dense modes."""
    cases = [draw_case(s) for s in SEEDS]
    assert {alg for _, alg, _, _ in cases} == {0, 1}
    assert {o["EventDirection"][0] for _, _, o, _ in cases} >= {-1, 1}
    assert {o["DenseOutputInterpolation"] for _, _, o, _ in cases} == {0, 1}
    assert any(o["ActiveSystems"] < e.NS for e, _, o, _ in cases)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_random_options_every_layout_and_launch(oracle, seed):
    ens, alg, opts, end = draw_case(seed)
    ndo = 60
    print(f"case: seed {seed}, NS {ens.NS}, algorithm {alg}, options {opts}", flush=True)
    p = run_oracle(oracle, ens, periods=2, algorithm=alg, options=opts, ndo=ndo, end_time=end)
    n = int(p.DenseOutputIndex.max())
    TLT = ens.NS * ens.N
    for layout in (LAYOUT_ROLE_PAIR, LAYOUT_SYSTEM_PAIR, LAYOUT_GROUP):
        for persistent in (1, 0):
            print(f"  layout {layout}, persistent {persistent}", flush=True)
            s = run_gpu(ens, periods=2, algorithm=alg, options=opts, ndo=ndo, end_time=end,
                        layout=layout,
                        setup=lambda s: s.SolverOption(O.PersistentLaunch, persistent))
            compare_solver_to_oracle(s, p)
            assert_bitwise("DenseIndex", s.host_view(V.DenseIndex), p.DenseOutputIndex)
            assert_bitwise("DenseTime", s.host_view(V.DenseTime)[: n * ens.NS],
                           p.DenseOutputTimeInstances[: n * ens.NS])
            assert_bitwise("DenseState", s.host_view(V.DenseState)[: n * TLT * 2],
                           p.DenseOutputStates[: n * TLT * 2])
            s.close()


# Other unit counts and their layouts: lane groups (M in LDS) for 8, 16 and 64 units, role pairs
# for 64 (two waves), the eight-wave chain layout for 128, the streamed wide layout for 96 and
# 256 (a reference fold for 96). Shorter time domains keep the oracle's part to seconds.
OTHER = [(8, LAYOUT_GROUP, 0.5), (16, LAYOUT_GROUP, 0.5), (64, LAYOUT_GROUP, 0.2),
         (64, LAYOUT_ROLE_PAIR, 0.2), (96, LAYOUT_GROUP, 0.1), (128, LAYOUT_ROLE_PAIR, 0.1),
         (256, LAYOUT_GROUP, 0.02)]


@pytest.mark.gpu
@pytest.mark.parametrize("n,layout,end", OTHER)
@pytest.mark.parametrize("seed", [707, 808])
def test_random_options_other_unit_counts(oracle, seed, n, layout, end):
    ens, alg, opts, _ = draw_case(seed, n)
    ndo = 40
    print(f"case: seed {seed}, N {n}, layout {layout}, NS {ens.NS}, algorithm {alg}, end {end}",
          flush=True)
    p = run_oracle(oracle, ens, periods=2, algorithm=alg, options=opts, ndo=ndo, end_time=end)
    s = run_gpu(ens, periods=2, algorithm=alg, options=opts, ndo=ndo, end_time=end, layout=layout)
    compare_solver_to_oracle(s, p)
    n_out = int(p.DenseOutputIndex.max())
    assert_bitwise("DenseIndex", s.host_view(V.DenseIndex), p.DenseOutputIndex)
    assert_bitwise("DenseTime", s.host_view(V.DenseTime)[: n_out * ens.NS],
                   p.DenseOutputTimeInstances[: n_out * ens.NS])
    s.close()
