"""Loader for libbex.so, the gfx950 solver library (C ABI in include/bex_capi.h).

The library is built in-tree by ``python -m bubbleensemble_implicitextension_amd.build``.
There is no fallback: if the library is missing or cannot be loaded, every solver entry point
raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbex.so")

NSTAT = 32


class ProblemDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("NS", "UPS", "UD", "TPB", "SPB", "NC", "CBW", "CCI", "NUP", "NSP", "NGP", "NiGP",
                 "NUA", "NiUA", "NSA", "NiSA", "NE", "NDO", "Algorithm")]


class DeviceProperties(ctypes.Structure):
    """bex_device_properties (include/bex_capi.h)."""
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 256),
                ("total_global_mem", ctypes.c_longlong),
                ("shared_mem_per_block", ctypes.c_longlong),
                ("total_const_mem", ctypes.c_longlong)] + \
               [(n, ctypes.c_int) for n in
                ("regs_per_block", "multiprocessor_count", "major", "minor", "clock_rate_khz",
                 "memory_clock_rate_khz", "memory_bus_width", "warp_size",
                 "max_threads_per_multiprocessor", "max_threads_per_block")] + \
               [("max_threads_dim", ctypes.c_int * 3), ("max_grid_size", ctypes.c_int * 3)] + \
               [(n, ctypes.c_int) for n in ("concurrent_copy", "concurrent_kernels",
                                            "ecc_enabled")]


_lib = None

_SIGNATURES = {
    "bex_compiled_dims": ([ctypes.POINTER(ProblemDesc)], ctypes.c_int),
    "bex_create": ([ctypes.POINTER(ProblemDesc), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
                   ctypes.c_int),
    "bex_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "bex_solver_option": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "bex_solver_option_indexed": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double],
                                  ctypes.c_int),
    "bex_set_host_unit": ([ctypes.c_void_p] + [ctypes.c_int] * 4 + [ctypes.c_double], ctypes.c_int),
    "bex_set_host_system": ([ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_double],
                            ctypes.c_int),
    "bex_set_host_global": ([ctypes.c_void_p] + [ctypes.c_int] * 2 + [ctypes.c_double],
                            ctypes.c_int),
    "bex_set_host_coupling": ([ctypes.c_void_p] + [ctypes.c_int] * 4 + [ctypes.c_double],
                              ctypes.c_int),
    "bex_set_host_dense_index": ([ctypes.c_void_p] + [ctypes.c_int] * 2 + [ctypes.c_double],
                                 ctypes.c_int),
    "bex_set_host_dense_state": ([ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.c_double],
                                 ctypes.c_int),
    "bex_get_host_unit": ([ctypes.c_void_p] + [ctypes.c_int] * 4 +
                          [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_get_host_system": ([ctypes.c_void_p] + [ctypes.c_int] * 3 +
                            [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_get_host_global": ([ctypes.c_void_p] + [ctypes.c_int] * 2 +
                            [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_get_host_coupling": ([ctypes.c_void_p] + [ctypes.c_int] * 4 +
                              [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_get_host_dense_index": ([ctypes.c_void_p] + [ctypes.c_int] * 2 +
                                 [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_get_host_dense_state": ([ctypes.c_void_p] + [ctypes.c_int] * 5 +
                                 [ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_host_view": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                       ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)],
                      ctypes.c_int),
    "bex_sync_h2d": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "bex_sync_d2h": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "bex_solve": ([ctypes.c_void_p], ctypes.c_int),
    "bex_synchronise_device": ([ctypes.c_void_p], ctypes.c_int),
    "bex_insert_synchronisation_point": ([ctypes.c_void_p], ctypes.c_int),
    "bex_synchronise_solver": ([ctypes.c_void_p], ctypes.c_int),
    "bex_checkpoint_save": ([ctypes.c_void_p], ctypes.c_int),
    "bex_checkpoint_restore": ([ctypes.c_void_p], ctypes.c_int),
    "bex_set_launch_order": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int],
                             ctypes.c_int),
    "bex_set_warm_start_systems": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_ubyte),
                                    ctypes.c_int], ctypes.c_int),
    "bex_persistent_grid": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "bex_request_stop": ([ctypes.c_void_p], ctypes.c_int),
    "bex_clear_stop": ([ctypes.c_void_p], ctypes.c_int),
    "bex_get_stats": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "bex_reset_stats": ([ctypes.c_void_p], ctypes.c_int),
    "bex_timer_start": ([ctypes.c_void_p], ctypes.c_int),
    "bex_timer_stop": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_enable_timeline": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "bex_get_timeline": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_peek_timeline": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "bex_set_warm_start_network": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "bex_enable_sampling": ([ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int], ctypes.c_int),
    "bex_get_samples": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_longlong,
                         ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "bex_mlp_workspace_floats": ([ctypes.c_longlong], ctypes.c_longlong),
    "bex_mlp_loss_grad": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                          ctypes.c_int),
    "bex_mlp_adam": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, ctypes.c_void_p], ctypes.c_int),
    "bex_mlp_predict": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "bex_mlp_predict_mfma": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                              ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p],
                             ctypes.c_int),
    "bex_mlp_predict_mfma_noalloc": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                                      ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                      ctypes.c_void_p], ctypes.c_int),
    "bex_pool_create": ([ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                         ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "bex_pool_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "bex_pool_class_solver": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
                              ctypes.c_int),
    "bex_pool_set_couplings": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "bex_pool_set_order": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                           ctypes.c_int),
    "bex_pool_order_by_cost": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "bex_pool_get_order": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                           ctypes.c_int),
    "bex_pool_solve": ([ctypes.c_void_p], ctypes.c_int),
    "bex_pool_solve_periods": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "bex_pool_synchronise": ([ctypes.c_void_p], ctypes.c_int),
    "bex_pool_variant_count": ([ctypes.c_int], ctypes.c_int),
    "bex_pool_variant_name": ([ctypes.c_int, ctypes.c_int], ctypes.c_char_p),
    "bex_pool_set_variant": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "bex_pool_get_variant": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                             ctypes.c_int),
    "bex_pool_autotune": ([ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)],
                          ctypes.c_int),
    "bex_pool_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "bex_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "bex_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "bex_get_device_properties": ([ctypes.c_int, ctypes.POINTER(DeviceProperties)], ctypes.c_int),
    "bex_select_device_by_closest_revision": ([ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def load():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m bubbleensemble_implicitextension_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


class BexError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def pool_check(code, handle=None):
    if code != 0:
        msg = load().bex_pool_last_error(handle)
        raise BexError(code, msg.decode() if msg else "")
    return code


def check(code, handle=None):
    if code != 0:
        msg = load().bex_last_error(handle)
        raise BexError(code, msg.decode() if msg else "")
    return code
