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
