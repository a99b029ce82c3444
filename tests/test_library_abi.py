"""CPU checks of the C ABI boundary: the built library loads and exports every entry point
declared in include/bex_capi.h (no GPU calls are made)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(REPO, "include", "bex_capi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bex_\w+)\(", text, re.M)))


def test_header_declares_the_reference_surface():
    names = declared_entry_points()
    for n in ("bex_create", "bex_destroy", "bex_solver_option", "bex_solver_option_indexed",
              "bex_set_host_unit", "bex_get_host_unit", "bex_sync_h2d", "bex_sync_d2h",
              "bex_solve", "bex_synchronise_solver", "bex_insert_synchronisation_point",
              "bex_synchronise_device"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from bubbleensemble_implicitextension_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libbex.so not built")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_entry_points()) <= set(_native.EXPORTED_SYMBOLS)


def test_compiled_dims_match_keller_miksis():
    from bubbleensemble_implicitextension_amd import _native, problem_solver
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libbex.so not built")
    d = problem_solver.compiled_dims()
    assert (d["UD"], d["NUP"], d["NUA"], d["NiUA"], d["NE"]) == (2, 21, 64, 1, 1)


def test_cpp_header_free_functions(tmp_path):
    """The drop-in C++ header compiles with a host compiler and its free functions behave as the
    reference's (Interface.cuh:70-75, 2035-2106): enumerator names, the out-of-range strings, and
    a device listing that is empty where there is no device."""
    import shutil
    import subprocess
    from bubbleensemble_implicitextension_amd import _native
    if not os.path.exists(_native.LIB_PATH) or not shutil.which("g++"):
        pytest.skip("libbex.so not built or no host compiler")
    src = tmp_path / "names.cpp"
    src.write_text(r'''
#include "CoupledSystems_PerBlock_Interface.hpp"
int main() {
  std::cout << SolverOptionsToString(InitialTimeStep) << ' '
            << SolverOptionsToString(SharedCouplingMatrices) << ' '
            << SolverOptionsToString(DenseOutputInterpolation) << ' '
            << VariablesToString(All) << ' ' << VariablesToString(DenseState) << ' '
            << VariablesToString(IntegerSystemAccessories) << '\n'
            << SolverOptionsToString(static_cast<ListOfSolverOptions>(-7)) << '\n'
            << VariablesToString(static_cast<ListOfVariables>(99)) << '\n';
  int n = 0;
  if (bex_device_count(&n) != BEX_OK) return 3;
  if (n == 0) ListCUDADevices();
  return 0;
}
''')
    exe = tmp_path / "names"
    pkg = os.path.dirname(_native.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(REPO, "include"), str(src),
                    "-L" + pkg, "-lbex", "-Wl,-rpath," + pkg, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == ("InitialTimeStep SharedCouplingMatrices DenseOutputInterpolation "
                        "All DenseState IntegerSystemAccessories")
    assert lines[1] == "Non-existent solver option!"
    assert lines[2] == "Non-existent variable!"
