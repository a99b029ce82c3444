"""The reference's text dumps (Print, WriteToFile*, Interface.cuh:1531-1947) in the C++ header and
in the Python mirror, checked on the CPU against files written out by hand from the format
rules: fields right-aligned in 18 characters, floating-point values as C++ std::scientific with
precision 10, every field followed by a comma; unit scope with PADDING rows for padded threads
and for the unused threads of a partly filled last block; the per-system dense-output report
with its section headers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import problem_solver as ps

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# NS = 3 systems of UPS = 2 units, TPB = 3, SPB = 2: 4 logical threads per block, padded to 6,
# 2 blocks, 12 logical threads; threads 0-3 and 6-7 hold units, 4-5 and 8-11 are padding.
LAYOUT = dict(NS=3, UPS=2, TPB=3, SPB=2)
TLT = 12
REAL = [0, 1, 2, 3, 6, 7]


def unit_data():
    d = np.full(2 * TLT, 999.0)  # padding rows must never be printed
    col0 = {0: 1.5, 1: -2.25, 2: 3e-5, 3: 4e10, 6: 0.0, 7: -1.0}
    for i in REAL:
        d[i] = col0[i]
        d[TLT + i] = i + 0.5
    return d


def unit_ints():
    d = np.full(TLT, 77, dtype=np.int32)
    for i in REAL:
        d[i] = -1 if i % 2 == 0 else i
    return d


P = "           PADDING,"
EXPECTED_UNIT = "".join([
    "  1.5000000000e+00,  5.0000000000e-01,\n",
    " -2.2500000000e+00,  1.5000000000e+00,\n",
    "  3.0000000000e-05,  2.5000000000e+00,\n",
    "  4.0000000000e+10,  3.5000000000e+00,\n",
    P + P + "\n", P + P + "\n",
    "  0.0000000000e+00,  6.5000000000e+00,\n",
    " -1.0000000000e+00,  7.5000000000e+00,\n",
    P + P + "\n", P + P + "\n", P + P + "\n", P + P + "\n"])
EXPECTED_INT = "".join([
    "                -1,\n", "                 1,\n", "                -1,\n", "                 3,\n",
    P + "\n", P + "\n", "                -1,\n", "                 7,\n",
    P + "\n", P + "\n", P + "\n", P + "\n"])
SYSTEM = np.array([0.0, 0.25, 1.0, 1.0, 1.0, 2.0])  # TimeDomain [2][NS], column-major
EXPECTED_SYSTEM = ("  0.0000000000e+00,  1.0000000000e+00,\n"
                   "  2.5000000000e-01,  1.0000000000e+00,\n"
                   "  1.0000000000e+00,  2.0000000000e+00,\n")

# Dense-output report of system 2 (logical threads 6, 7) with UD = 2, NUP = 2, NUA = 1,
# NiUA = 1, NSP = 1, NSA = 0, NiSA = 0, NGP = 2, NiGP = 1 and two stored points.
W = 18


def f(v):
    return f"{v:>{W}}" if isinstance(v, str) else (f"{v:>{W}d}" if isinstance(v, int)
                                                   else f"{v:>{W}.10e}")


EXPECTED_DENSE = "".join([
    "UnitParameters:\n", f("Serial No.") + "," + f("U0") + "," + f("U1") + "\n",
    f(0) + "," + f(0.0) + "," + f(-1.0) + "\n",
    f(1) + "," + f(6.5) + "," + f(7.5) + "\n", "\n",
    "UnitAccessories:\n", f("Serial No.") + "," + f("U0") + "," + f("U1") + "\n",
    f(0) + "," + f(0.0) + "," + f(-1.0) + "\n", "\n",
    "IntegerUnitAccessories:\n", f("Serial No.") + "," + f("U0") + "," + f("U1") + "\n",
    f(0) + "," + f(-1) + "," + f(7) + "\n", "\n",
    "SystemParameters:\n", f(12.0) + "\n\n",
    "SystemAccessories:\n", "\n\n",
    "IntegerSystemAccessories:\n", "\n\n",
    "GlobalParameters:\n", f(3.0) + "," + f(-4.0) + "\n\n",
    "IntegerGlobalParameters:\n", f(5) + "\n\n",
    "Time series:\n", f("t") + "," + f("U0/C0") + "," + f("U0/C1") + "," + f("U1/C0") + ","
    + f("U1/C1") + "\n",
    f(0.5) + "," + f(60.0) + "," + f(160.0) + "," + f(70.0) + "," + f(170.0) + "\n",
    f(0.75) + "," + f(1060.0) + "," + f(1160.0) + "," + f(1070.0) + "," + f(1170.0) + "\n"])


def dense_buffers():
    V = ps.ListOfVariables
    ndo, ud = 3, 2
    times = np.zeros(ndo * 3)
    times[2 + 0 * 3], times[2 + 1 * 3] = 0.5, 0.75
    states = np.zeros(ndo * ud * TLT)
    for i in range(ndo):
        for c in range(ud):
            for th in range(TLT):
                states[th + c * TLT + i * ud * TLT] = 1000.0 * i + 100.0 * c + 10.0 * th
    return {V.UnitParameters: unit_data(), V.UnitAccessories: unit_data()[:TLT],
            V.IntegerUnitAccessories: unit_ints(), V.SystemParameters: np.array([10.0, 11.0, 12.0]),
            V.SystemAccessories: np.zeros(1), V.IntegerSystemAccessories: np.zeros(1, np.int32),
            V.GlobalParameters: np.array([3.0, -4.0]),
            V.IntegerGlobalParameters: np.array([5], np.int32),
            V.DenseTime: times, V.DenseState: states}


DENSE_DIMS = dict(UD=2, NUP=2, NUA=1, NiUA=1, NSP=1, NSA=0, NiSA=0, NGP=2, NiGP=1)


def test_python_writers_follow_the_format_rules(tmp_path):
    t = ps.ThreadLayout.make(**LAYOUT)
    assert t.TotalLogicalThreads == TLT and t.ThreadPaddingPerBlock == 2
    ps.write_unit_scope(tmp_path / "u.txt", 2, unit_data(), t)
    ps.write_unit_scope(tmp_path / "i.txt", 1, unit_ints(), t)
    ps.write_system_scope(tmp_path / "s.txt", 3, 2, SYSTEM)
    ps.write_dense_report(tmp_path / "d.txt", 2, 2, t, DENSE_DIMS, dense_buffers())
    assert (tmp_path / "u.txt").read_text() == EXPECTED_UNIT
    assert (tmp_path / "i.txt").read_text() == EXPECTED_INT
    assert (tmp_path / "s.txt").read_text() == EXPECTED_SYSTEM
    assert (tmp_path / "d.txt").read_text() == EXPECTED_DENSE


CPP = r'''
#include "CoupledSystems_PerBlock_Interface.hpp"
#include <cstdio>
int main() {
  using namespace bex_detail;
  const ThreadLayout t = MakeThreadLayout(3, 2, 3, 2);
  double u[24]; int iu[12]; double states[3 * 2 * 12], times[9] = {0};
  for (int k = 0; k < 24; k++) u[k] = 999.0;
  for (int k = 0; k < 12; k++) iu[k] = 77;
  const int real[6] = {0, 1, 2, 3, 6, 7};
  const double c0[6] = {1.5, -2.25, 3e-5, 4e10, 0.0, -1.0};
  for (int r = 0; r < 6; r++) {
    u[real[r]] = c0[r];
    u[12 + real[r]] = real[r] + 0.5;
    iu[real[r]] = real[r] % 2 == 0 ? -1 : real[r];
  }
  for (int i = 0; i < 3; i++)
    for (int c = 0; c < 2; c++)
      for (int th = 0; th < 12; th++) states[th + c * 12 + i * 24] = 1000.0 * i + 100.0 * c + 10.0 * th;
  times[2] = 0.5; times[5] = 0.75;
  const double sys[6] = {0.0, 0.25, 1.0, 1.0, 1.0, 2.0};
  const double sp[3] = {10.0, 11.0, 12.0}, gp[2] = {3.0, -4.0}, sa[1] = {0.0};
  const int igp[1] = {5}, isa[1] = {0};
  WriteUnitScope("u.txt", 2, u, t);
  WriteUnitScope("i.txt", 1, iu, t);
  WriteSystemScope("s.txt", 3, 2, sys);
  DenseReport r{t, 2, 2, 1, 1, 1, 0, 0, 2, 1, u, u, sp, sa, gp, times, states, iu, isa, igp};
  WriteDenseOutputReport("d.txt", 2, 2, r);
  return 0;
}
'''


def test_cpp_header_writers_follow_the_format_rules(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no host compiler")
    src = tmp_path / "dump.cpp"
    src.write_text(CPP)
    exe = tmp_path / "dump"
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(REPO, "include"), str(src), "-o",
                    str(exe)], check=True)
    subprocess.run([str(exe)], cwd=tmp_path, check=True, timeout=60)
    assert (tmp_path / "u.txt").read_text() == EXPECTED_UNIT
    assert (tmp_path / "i.txt").read_text() == EXPECTED_INT
    assert (tmp_path / "s.txt").read_text() == EXPECTED_SYSTEM
    assert (tmp_path / "d.txt").read_text() == EXPECTED_DENSE


FLOAT_INSTANCE = r"""
#include "CoupledSystems_PerBlock_Interface.hpp"
// The reference's float instantiation compiles; Set/Get take and return float.
using Solver = ProblemSolver<4, 8, 2, 8, 1, 1, 0, 0, 21, 0, 0, 0, 64, 1, 0, 0, 1, 0, RKCK45, float>;
float probe(Solver& s) {
  s.SetHost(0, 0, ActualState, 0, 1.0f);
  s.SolverOption(RelativeTolerance, 0, 1e-6f);
  return s.GetHost<float>(0, 0, ActualState, 0);
}
"""


def test_cpp_header_accepts_float_precision(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no host compiler")
    src = tmp_path / "float_instance.cpp"
    src.write_text(FLOAT_INSTANCE)
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(REPO, "include"),
                    str(src)], check=True)
