"""GPU tests of the drop-in boundary beyond the solve itself: the reference's four coupling
storage schemes, its error conventions, its device helpers, and its driver's flow written
against the C++ header.

Coupling storage (Interface.cuh:1001-1070): the full matrix (CCI = 0, CBW = 0), a band of
2 CBW + 1 shifted diagonals with circular extension (CCI = 0, CBW > 0), a circulant band
(CCI = 1, CBW > 0: one value per shifted diagonal) and a full circulant (CCI = 1, CBW = 0).
SetHost accepts exactly the stored elements; the solve must see the matrix those elements
define (zeros outside the band), which is checked bit for bit against a full-storage solver
given that matrix and against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import collapse
from bubbleensemble_implicitextension_amd import problem_solver as ps
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


def solve_stored(ens, cbw, cci, values, layout=ps.LAYOUT_AUTO):
    """A solver with the given storage scheme, coupling 0 set element by element through
    SetHost from `values` (the dense matrix whose stored elements are written); returns the
    solver after one Solve() with the given kernel layout and the set of (row, col) that
    SetHost accepted."""
    s = collapse.make_solver(ens, CBW=cbw, CCI=cci)
    s.SolverOption(ps.ListOfSolverOptions.KernelLayout, layout)
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
    # The system-pair layout reads the expanded matrix and its transpose the same way.
    sp, _ = solve_stored(ens, cbw, cci, values, layout=ps.LAYOUT_SYSTEM_PAIR)
    for var in (V.ActualState, V.ActualTime, V.UnitAccessories, V.IntegerUnitAccessories):
        assert_bitwise(var.name, sp.host_view(var), full.host_view(var))
    assert_bitwise("stats", sp.stats(), full.stats())
    sp.close()
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
    # The persistent launch's hand-out order must list every system exactly once.
    for bad in ([0, 1, 2], [0, 1, 2, 2], [0, 1, 2, 4], [-1, 0, 1, 2]):
        with pytest.raises(BexError, match="every system once"):
            s.set_launch_order(bad)
    s.set_launch_order([3, 2, 1, 0])
    s.set_launch_order(None)
    # The learned-start flags need one entry per system.
    for bad in ([1, 0, 1], [1, 0, 1, 0, 1]):
        with pytest.raises(BexError, match="one flag per system"):
            s.set_warm_start_systems(bad)
    s.set_warm_start_systems([1, 0, 0, 1])
    s.set_warm_start_systems(None)
    s.close()


def test_persistent_grid_query_and_cap():
    """The persistent launch's grid: one wave per work item up to what fits on the device (a
    role-pair item is one system, a system-pair item two), at most PersistentWorkgroups when
    set; a negative cap is rejected, and the query fails for the hardware-dispatched launch."""
    O = ps.ListOfSolverOptions
    ens = wl.small_ensemble(n=32, ns=9, seed=4)
    assert ens.NS == 9
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    s.SolverOption(O.KernelLayout, ps.LAYOUT_ROLE_PAIR)
    assert s.persistent_grid() == (9, 1)
    s.SolverOption(O.KernelLayout, ps.LAYOUT_SYSTEM_PAIR)
    assert s.persistent_grid() == (5, 1)       # five pairs, the last one with one system
    s.SolverOption(O.PersistentWorkgroups, 3)
    assert s.persistent_grid() == (3, 1)
    s.SolverOption(O.KernelLayout, ps.LAYOUT_ROLE_PAIR)
    assert s.persistent_grid() == (3, 1)
    s.SolverOption(O.ActiveSystems, 2)
    assert s.persistent_grid() == (2, 1)       # two role-pair systems
    s.SolverOption(O.KernelLayout, ps.LAYOUT_SYSTEM_PAIR)
    assert s.persistent_grid() == (1, 1)       # one system pair
    with pytest.raises(BexError, match="PersistentWorkgroups"):
        s.SolverOption(O.PersistentWorkgroups, -1)
    s.SolverOption(O.PersistentLaunch, 0)
    with pytest.raises(BexError, match="no persistent kernel"):
        s.persistent_grid()
    s.close()


def test_device_helpers():
    """ListCUDADevices / SelectDeviceByClosestRevision / PrintPropertiesOfSpecificDevice
    (Interface.cuh:259-345) report the HIP device."""
    devs = ps.list_devices()
    assert len(devs) >= 1
    d = devs[0]
    print(f"device 0: {d['name']} {d['arch']} CUs {d['multiprocessor_count']}", flush=True)
    assert d["arch"].startswith("gfx950")
    assert d["warp_size"] == 64
    assert d["multiprocessor_count"] >= 1 and d["total_global_mem"] > 0
    assert d["max_threads_per_block"] >= 1024 and d["shared_mem_per_block"] >= 64 * 1024
    assert 0 <= ps.SelectDeviceByClosestRevision(d["major"], d["minor"]) < len(devs)
    with pytest.raises(BexError):
        ps._native.check(ps._native.load().bex_get_device_properties(
            len(devs), ps.ctypes.byref(ps._native.DeviceProperties())))


REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(REPO, "bubbleensemble_implicitextension_amd", "collapse_strength")
RADII = os.path.join(REPO, "tests", "golden", "BubbleSizes_N32_D1_d150_RE2-8.txt")
COUPLING = os.path.join(REPO, "tests", "golden", "CouplingMatrix_N32_D1_d150_RE2-8.txt")


def app_grid(lo, hi, n):
    """The collapse-strength application's linear grid: lo + (i / (n - 1)) (hi - lo), the last
    point exactly hi."""
    v = [lo + (i / (n - 1)) * (hi - lo) for i in range(n)]
    v[-1] = hi
    return np.array(v)


def read_result_file(path):
    rows = [line.rstrip(",\n").split(",") for line in open(path)]
    return np.array([[float(x) for x in r] for r in rows])


def test_cpp_driver_matches_python_protocol(tmp_path):
    """The reference driver's flow (CollapseStrength.cu:68-204: device listing, 8 x 8 PA x f grid
    of 32-bubble systems on the reference's own fixture geometry, 4 transient periods, accessory
    reset, 32 recording periods, result file) compiled against the drop-in C++ header gives the
    result file of the Python protocol (collapse.collapse_protocol, write_collapse_file) on the
    same inputs, to the file's printed precision (the two restate the parameter formulas with
    different association, a last-bit difference)."""
    assert os.path.exists(APP), "collapse_strength is not built: run __graft_entry__.build()"
    pa_lo, pa_hi, f_lo, f_hi = 0.5, 0.8, 100.0, 200.0  # below the bubbles' resonances
    cpp_out = tmp_path / "cpp.txt"
    print("C++ driver: 64 systems x 36 periods", flush=True)
    r = subprocess.run([APP, RADII, COUPLING, str(cpp_out), str(pa_lo), str(pa_hi), str(f_lo),
                        str(f_hi)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Device number: 0" in r.stdout and "Selected device number: 0" in r.stdout

    re_um = np.loadtxt(RADII)
    m = np.loadtxt(COUPLING).reshape(32, 32)  # row-major text, row = target bubble
    f_vals, pa_vals = app_grid(f_lo, f_hi, 8), app_grid(pa_lo, pa_hi, 8)
    F, PA = np.meshgrid(f_vals, pa_vals, indexing="ij")  # frequency outer, amplitude inner
    pa, f = PA.reshape(-1), F.reshape(-1)
    params = wl.unit_parameters(pa[:, None], f[:, None], re_um[None, :])
    state0 = np.zeros((64, 32, 2))
    state0[..., 0] = 1.0
    ens = wl.Ensemble("fixture", 32, pa, f, re_um, params, m, state0, wl.default_options(), 36)
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    print("Python protocol: 64 systems x 36 periods", flush=True)
    collapse.collapse_protocol(s, 4, 32)
    py_out = tmp_path / "py.txt"
    collapse.write_collapse_file(s, str(py_out))
    s.close()

    a, b = read_result_file(cpp_out), read_result_file(py_out)
    assert a.shape == b.shape == (64 * 32, 3 + 64)
    np.testing.assert_array_equal(a[:, :3], b[:, :3])  # system, PA, f
    np.testing.assert_allclose(a[:, 3:], b[:, 3:], rtol=2e-3, atol=0)
    assert (a[:, 3:] > 0).mean() > 0.5  # collapses were recorded

    # The same scan on the C++ split solver (SplitProblemSolver: 4 + 8 systems one per wave in
    # one launch, the other 52 two per wave in another, plain start): the same result file, byte
    # for byte.
    split_out = tmp_path / "split.txt"
    print("C++ driver on the split solver: 64 systems x 36 periods", flush=True)
    r = subprocess.run([APP, RADII, COUPLING, str(split_out), str(pa_lo), str(pa_hi), str(f_lo),
                        str(f_hi), "4", "8"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "split solver: 2 parts" in r.stdout
    assert split_out.read_text() == cpp_out.read_text()


PRINT_APP = r'''
#include "CoupledSystems_PerBlock_Interface.hpp"
// 3 systems of 32 units, two systems per block of 48 threads: 64 logical threads per block,
// padded to 96; the second block holds one system (Interface.cuh:386-391).
typedef ProblemSolver<3, 32, 2, 48, 2, 1, 0, 0, 21, 0, 0, 0, 64, 1, 0, 0, 1, 4, RKCK45, double> S;
int main() {
  S s(0);
  for (int sid = 0; sid < 3; sid++) {
    s.SetHost(sid, TimeDomain, 0, 0.0);
    s.SetHost(sid, TimeDomain, 1, 1.0 + sid);
    s.SetHost(sid, ActualTime, 0, 0.125 * sid);
    s.SetHost(sid, CouplingStrength, 0, 2.0 - sid);
    s.SetHost(sid, DenseIndex, sid + 1);
    for (int uid = 0; uid < 32; uid++) {
      s.SetHost(sid, uid, ActualState, 0, 1.0 + 0.01 * sid + 0.001 * uid);
      s.SetHost(sid, uid, ActualState, 1, -0.5 * uid);
      for (int k = 0; k < 21; k++) s.SetHost(sid, uid, UnitParameters, k, 1e-3 * (k + 1) * (uid + 1) - sid);
      for (int k = 0; k < 64; k++) s.SetHost(sid, uid, UnitAccessories, k, k == uid ? 3.5 : 0.0);
      s.SetHost(sid, uid, IntegerUnitAccessories, 0, uid - sid);
      for (int i = 0; i < 4; i++)
        for (int c = 0; c < 2; c++) s.SetHost(sid, uid, DenseState, c, i, 100.0 * i + 10.0 * c + uid);
    }
    for (int i = 0; i < 4; i++) s.SetHost(sid, DenseTime, i, 0.25 * i + sid);
  }
  s.SetHost(CouplingIndex, 0, 0);
  for (int r = 0; r < 32; r++)
    for (int c = 0; c < 32; c++) s.SetHost(0, CouplingMatrix, r, c, r == c ? 0.0 : 1.0 / (1 + r + 2 * c));
  ListOfVariables vars[] = {TimeDomain, ActualTime, ActualState, UnitParameters, UnitAccessories,
                            IntegerUnitAccessories, CouplingStrength, CouplingIndex};
  for (ListOfVariables v : vars) s.Print(v);
  s.Print(CouplingMatrix, 0);
  s.Print(DenseOutput, 1);
  s.Print(DenseOutput, 2);
  return 0;
}
'''


def test_print_files_match_between_cpp_header_and_python(tmp_path):
    """Print(var), Print(CouplingMatrix, k) and Print(DenseOutput, sid) give the same bytes from
    the C++ header and from the Python mirror, on a padded layout (two systems per block of 48
    threads); the format itself is pinned on the CPU (test_text_dump.py)."""
    import shutil
    from bubbleensemble_implicitextension_amd import _native

    if not shutil.which("g++"):
        pytest.skip("no host compiler")
    print("case: Print on 3 systems x 32 units, SPB 2, TPB 48", flush=True)
    cpp_dir, py_dir = tmp_path / "cpp", tmp_path / "py"
    cpp_dir.mkdir()
    py_dir.mkdir()
    src = tmp_path / "print_app.cpp"
    src.write_text(PRINT_APP)
    exe = tmp_path / "print_app"
    pkg = os.path.dirname(_native.LIB_PATH)
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(repo, "include"), str(src),
                    "-L" + pkg, "-lbex", "-Wl,-rpath," + pkg, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], cwd=cpp_dir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    s = ps.ProblemSolver(3, 32, TPB=48, SPB=2, NDO=4)
    for sid in range(3):
        s.SetHost(sid, V.TimeDomain, 0, 0.0)
        s.SetHost(sid, V.TimeDomain, 1, 1.0 + sid)
        s.SetHost(sid, V.ActualTime, 0, 0.125 * sid)
        s.SetHost(sid, V.CouplingStrength, 0, 2.0 - sid)
        s.SetHost(sid, V.DenseIndex, sid + 1)
        for uid in range(32):
            s.SetHost(sid, uid, V.ActualState, 0, 1.0 + 0.01 * sid + 0.001 * uid)
            s.SetHost(sid, uid, V.ActualState, 1, -0.5 * uid)
            for k in range(21):
                s.SetHost(sid, uid, V.UnitParameters, k, 1e-3 * (k + 1) * (uid + 1) - sid)
            for k in range(64):
                s.SetHost(sid, uid, V.UnitAccessories, k, 3.5 if k == uid else 0.0)
            s.SetHost(sid, uid, V.IntegerUnitAccessories, 0, uid - sid)
            for i in range(4):
                for c in range(2):
                    s.SetHost(sid, uid, V.DenseState, c, i, 100.0 * i + 10.0 * c + uid)
        for i in range(4):
            s.SetHost(sid, V.DenseTime, i, 0.25 * i + sid)
    s.SetHost(V.CouplingIndex, 0, 0)
    for row in range(32):
        for col in range(32):
            s.SetHost(0, V.CouplingMatrix, row, col,
                      0.0 if row == col else 1.0 / (1 + row + 2 * col))
    for v in (V.TimeDomain, V.ActualTime, V.ActualState, V.UnitParameters, V.UnitAccessories,
              V.IntegerUnitAccessories, V.CouplingStrength, V.CouplingIndex):
        s.Print(v, directory=str(py_dir))
    s.Print(V.CouplingMatrix, 0, directory=str(py_dir))
    s.Print(V.DenseOutput, 1, directory=str(py_dir))
    s.Print(V.DenseOutput, 2, directory=str(py_dir))
    with pytest.raises(BexError):
        s.Print(V.CouplingMatrix, 1)  # NC = 1
    with pytest.raises(BexError):
        s.Print(V.DenseOutput, 3)
    s.close()

    names = sorted(os.listdir(cpp_dir))
    assert names == sorted(os.listdir(py_dir))
    assert "ActualStateInSolverObject.txt" in names and "DenseOutput_2.txt" in names
    assert "CouplingMatrixInSolverObject_0.txt" in names
    for n in names:
        assert (cpp_dir / n).read_text() == (py_dir / n).read_text(), n
    state = (cpp_dir / "ActualStateInSolverObject.txt").read_text().splitlines()
    assert len(state) == 192
    assert state[64].split(",")[0].strip() == "PADDING"  # the first block's padded threads
    assert state[128].split(",")[0].strip() == "PADDING"  # the second block's empty system
    assert state[96].split(",")[0].strip() == "1.0200000000e+00"  # system 2, unit 0
