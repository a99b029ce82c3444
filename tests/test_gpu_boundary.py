"""GPU tests of the drop-in boundary beyond the solve itself: the reference's four coupling
storage schemes and its error conventions.

Coupling storage (Interface.cuh:1001-1070): the full matrix (CCI = 0, CBW = 0), a band of
2 CBW + 1 shifted diagonals with circular extension (CCI = 0, CBW > 0), a circulant band
(CCI = 1, CBW > 0: one value per shifted diagonal) and a full circulant (CCI = 1, CBW = 0).
SetHost accepts exactly the stored elements; the solve must see the matrix those elements
define (zeros outside the band), which is checked bit for bit against a full-storage solver
given that matrix and against the oracle.
"""
import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import collapse
from bubbleensemble_implicitextension_amd import workloads as wl
from bubbleensemble_implicitextension_amd._native import BexError
from bubbleensemble_implicitextension_amd.problem_solver import ListOfVariables as V

from gpu_helpers import assert_bitwise, compare_solver_to_oracle, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

END = 0.2


def shifted_diagonal(n, cbw, row, col):
    """Index of element (row, col) among the stored shifted diagonals, or None outside them
    (Interface.cuh:1020-1070: col - row + CBW, wrapped once into [0, N))."""
    mod = col - row + cbw
    if mod >= n:
        mod -= n
    if mod < 0:
        mod += n
    lim = 2 * cbw + 1 if cbw > 0 else n
    return mod if 0 <= mod < lim else None


def with_coupling(ens, m):
    return wl.Ensemble(ens.name + "-stored", ens.N, ens.pa_bar, ens.f_khz, ens.re_um, ens.params,
                       m, ens.state0, dict(ens.options), ens.periods)


def solve_stored(ens, cbw, cci, values):
    """A solver with the given storage scheme, coupling 0 set element by element through
    SetHost from `values` (the dense matrix whose stored elements are written); returns the
    solver after one Solve() and the set of (row, col) that SetHost accepted."""
    s = collapse.make_solver(ens, CBW=cbw, CCI=cci)
    n = ens.N
    # Everything but the coupling (fill leaves per-system matrices to the caller).
    collapse.fill(s, with_coupling(ens, np.zeros((ens.NS, n, n))))
    accepted = set()
    for row in range(n):
        for col in range(n):
            try:
                s.SetHost(0, V.CouplingMatrix, row, col, float(values[row, col]))
                accepted.add((row, col))
            except BexError:
                pass
    s.host_view(V.TimeDomain)[ens.NS:] = END
    s.SynchroniseFromHostToDevice(V.All)
    s.Solve()
    s.SynchroniseFromDeviceToHost(V.All)
    s.InsertSynchronisationPoint()
    s.SynchroniseSolver()
    return s, accepted


@pytest.mark.parametrize("cbw,cci", [(3, 0), (2, 1), (0, 1)])
def test_coupling_storage_schemes(oracle, cbw, cci):
    ens = wl.small_ensemble(n=32, ns=4, seed=29)
    n = ens.N
    print(f"coupling storage CBW={cbw} CCI={cci} N={n}", flush=True)
    m = ens.coupling
    if cci == 0:
        values = m  # band of the physical matrix
        expect = np.zeros_like(m)
        for row in range(n):
            for col in range(n):
                if shifted_diagonal(n, cbw, row, col) is not None:
                    expect[row, col] = m[row, col]
    else:
        # Circulant: one coupling value per shifted diagonal, taken from row 0 of the physical
        # matrix (zero on the main diagonal: no self-coupling). Every SetHost of a diagonal
        # writes the same value, so the order of the writes does not matter.
        lim = 2 * cbw + 1 if cbw > 0 else n
        per_diag = np.array([m[0, (k - cbw) % n] for k in range(lim)])
        values = np.zeros_like(m)
        expect = np.zeros_like(m)
        for row in range(n):
            for col in range(n):
                k = shifted_diagonal(n, cbw, row, col)
                if k is not None:
                    values[row, col] = per_diag[k]
                    expect[row, col] = per_diag[k]
    want_accepted = {(r, c) for r in range(n) for c in range(n)
                     if shifted_diagonal(n, cbw, r, c) is not None}

    s, accepted = solve_stored(ens, cbw, cci, values)
    assert accepted == want_accepted
    # GetHost returns what was set, element by element (Interface.cuh:1469), and zero elsewhere.
    for row, col in [(0, 0), (0, 1), (1, 0), (0, n - 1), (n - 1, 0), (5, 9), (17, 16)]:
        want = values[row, col] if (row, col) in accepted else 0.0
        assert s.GetHost(0, V.CouplingMatrix, row, col) == want

    full = run_gpu(with_coupling(ens, expect), periods=1, end_time=END)
    for var in (V.ActualState, V.ActualTime, V.UnitAccessories, V.IntegerUnitAccessories):
        assert_bitwise(var.name, s.host_view(var), full.host_view(var))
    assert_bitwise("stats", s.stats(), full.stats())
    p = run_oracle(oracle, with_coupling(ens, expect), periods=1, end_time=END)
    compare_solver_to_oracle(s, p)
    s.close()
    full.close()


def test_error_conventions():
    """Bound checks and wrong variable/overload combinations raise BexError with the
    reference's message text (Interface.cuh:718-733, 903-908) instead of exiting; the object
    stays usable afterwards."""
    ens = wl.small_ensemble(n=8, ns=4, seed=3)
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    with pytest.raises(BexError, match="SystemNumber"):
        s.SetHost(ens.NS, 0, V.ActualState, 0, 1.0)
    with pytest.raises(BexError, match="UnitNumber"):
        s.SetHost(0, ens.N, V.ActualState, 0, 1.0)
    with pytest.raises(BexError, match="Coupling Matrix Row"):
        s.SetHost(0, V.CouplingMatrix, ens.N, 0, 1.0)
    with pytest.raises(BexError):
        s.SetHost(0, 0, V.TimeDomain, 0, 1.0)  # a system-scope variable through the unit overload
    with pytest.raises(BexError):
        s.GetHost(0, V.ActualState, 0)  # a unit-scope variable through the system overload
    s.SetHost(1, 2, V.ActualState, 1, 0.25)
    assert s.GetHost(1, 2, V.ActualState, 1) == 0.25
    s.close()
