"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests, the smoke check and the CPU-baseline leg of bench.py import this module, and only
to check or time the host restatement of the reference solver path. It never feeds the
product path.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbex_oracle.so")

CANONICAL = 0
LITERAL = 1

STAT_NAMES = ["accepted", "rejected", "stages", "bicg_iterations", "bicg_rho_exits",
              "bicg_maxit", "events", "forced_accepts", "terminated", "events_omitted",
              "attempts"]
NSTAT = 32
HIST0 = 11


class Dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("NS", "N", "UD", "NUP", "NSP", "NGP", "NiGP", "NUA", "NiUA", "NSA", "NiSA",
                 "NE", "NDO", "NC")]


class Options(ctypes.Structure):
    _fields_ = [
        ("InitialTimeStep", ctypes.c_double),
        ("ActiveSystems", ctypes.c_int),
        ("MaximumTimeStep", ctypes.c_double),
        ("MinimumTimeStep", ctypes.c_double),
        ("TimeStepGrowLimit", ctypes.c_double),
        ("TimeStepShrinkLimit", ctypes.c_double),
        ("DenseOutputSaveFrequency", ctypes.c_int),
        ("DenseOutputMinimumTimeStep", ctypes.c_double),
        ("BiCGRhoTolerance", ctypes.c_double),
        ("BiCGResidualTolerance", ctypes.c_double),
        ("BiCGMaxIterations", ctypes.c_int),
        ("RelativeTolerance", ctypes.POINTER(ctypes.c_double)),
        ("AbsoluteTolerance", ctypes.POINTER(ctypes.c_double)),
        ("EventTolerance", ctypes.POINTER(ctypes.c_double)),
        ("EventDirection", ctypes.POINTER(ctypes.c_int)),
        ("Algorithm", ctypes.c_int),
        ("DenseOutputInterpolation", ctypes.c_int),
    ]


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class Buffers(ctypes.Structure):
    _fields_ = [
        ("TimeDomain", _DP), ("ActualState", _DP), ("ActualTime", _DP),
        ("UnitParameters", _DP), ("SystemParameters", _DP), ("GlobalParameters", _DP),
        ("IntegerGlobalParameters", _IP), ("UnitAccessories", _DP),
        ("IntegerUnitAccessories", _IP), ("SystemAccessories", _DP),
        ("IntegerSystemAccessories", _IP), ("CouplingMatrix", _DP),
        ("DenseOutputTimeInstances", _DP), ("DenseOutputStates", _DP),
        ("DenseOutputIndex", _IP),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_variant(plant, out_path):
    """Builds the oracle with planted mistake `plant` (BEX_ORACLE_PLANT in bex_oracle.cpp) into
    out_path, for tests that show a pin rejects that mistake."""
    subprocess.run(["make", "-s", "-C", HERE, "variant", f"PLANT={int(plant)}",
                    f"OUT={out_path}"], check=True)


NATIVE_LIB_PATH = os.path.join(HERE, "libbex_oracle_native.so")


def build_native():
    """Builds the oracle for this host's own instruction set (-march=native), for timing the
    CPU baseline on the machine it runs on. Returns the library path, or None if the build
    fails."""
    # -B: always rebuilt, so a copy built on another machine is never loaded here.
    r = subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], capture_output=True,
                       timeout=300)
    return NATIVE_LIB_PATH if r.returncode == 0 and os.path.exists(NATIVE_LIB_PATH) else None


_lib = None


@contextlib.contextmanager
def using_library(path):
    """Routes every oracle call inside the block to another build of the oracle."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def _load(path):
    L = ctypes.CDLL(path)
    L.bex_oracle_solve.argtypes = [ctypes.POINTER(Dims), ctypes.POINTER(Options),
                                   ctypes.POINTER(Buffers),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.c_int,
                                   ctypes.c_int]
    L.bex_oracle_solve.restype = ctypes.c_int
    L.bex_oracle_bicg.argtypes = [ctypes.c_int] + [_DP] * 5 + [ctypes.c_double,
                                                               ctypes.c_double,
                                                               ctypes.c_int,
                                                               ctypes.c_int, _DP]
    L.bex_oracle_bicg.restype = ctypes.c_int
    L.bex_oracle_km_rhs.argtypes = [_DP, ctypes.c_double, _DP, ctypes.c_int, _DP, _DP, _DP]
    L.bex_oracle_km_rhs.restype = None
    L.bex_oracle_pow.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.bex_oracle_pow.restype = ctypes.c_double
    L.bex_oracle_sincospi.argtypes = [ctypes.c_double, ctypes.c_int, _DP, _DP]
    L.bex_oracle_sin.argtypes = [ctypes.c_double, ctypes.c_int]
    L.bex_oracle_sin.restype = ctypes.c_double
    L.bex_oracle_cos.argtypes = [ctypes.c_double, ctypes.c_int]
    L.bex_oracle_cos.restype = ctypes.c_double
    L.bex_oracle_rkck45_linear.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _DP, _DP]
    L.bex_oracle_rkck45_linear.restype = ctypes.c_int
    L.bex_oracle_rk4_linear.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _DP]
    L.bex_oracle_rk4_linear.restype = ctypes.c_int
    L.bex_oracle_control.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.POINTER(Options), ctypes.c_int,
                                     _DP, _IP, _IP, _IP]
    L.bex_oracle_control.restype = None
    return L


def _dp(a):
    return a.ctypes.data_as(_DP) if a is not None and a.size else None


def _ip(a):
    return a.ctypes.data_as(_IP) if a is not None and a.size else None


class OracleProblem:
    """Host SoA buffers in the reference layout plus solver options, solved on the CPU."""

    def __init__(self, NS, N, NUP=21, NUA=64, NiUA=1, NE=1, NDO=0, UD=2, NC=1, NSP=0, NSA=0,
                 NiSA=0, NGP=0, NiGP=0):
        self.dims = Dims(NS, N, UD, NUP, NSP, NGP, NiGP, NUA, NiUA, NSA, NiSA, NE, NDO, NC)
        TLT = NS * N
        self.TimeDomain = np.zeros(2 * NS)
        self.ActualState = np.zeros(UD * TLT)
        self.ActualTime = np.zeros(NS)
        self.UnitParameters = np.zeros(NUP * TLT)
        self.SystemParameters = np.zeros(max(NSP * NS, 1))
        self.GlobalParameters = np.zeros(max(NGP, 1))
        self.IntegerGlobalParameters = np.zeros(max(NiGP, 1), dtype=np.int32)
        self.UnitAccessories = np.zeros(NUA * TLT)
        self.IntegerUnitAccessories = np.zeros(NiUA * TLT, dtype=np.int32)
        self.SystemAccessories = np.zeros(max(NSA * NS, 1))
        self.IntegerSystemAccessories = np.zeros(max(NiSA * NS, 1), dtype=np.int32)
        self.CouplingMatrix = np.zeros(NC * N * N)
        self.DenseOutputTimeInstances = np.zeros(max(NDO * NS, 1))
        self.DenseOutputStates = np.zeros(max(NDO * UD * TLT, 1))
        self.DenseOutputIndex = np.zeros(NS, dtype=np.int32)
        self.stats = np.zeros(NS * NSTAT, dtype=np.int64)
        self.opts = {
            "InitialTimeStep": 1e-2, "ActiveSystems": NS, "MaximumTimeStep": 1e6,
            "MinimumTimeStep": 1e-12, "TimeStepGrowLimit": 5.0, "TimeStepShrinkLimit": 0.1,
            "DenseOutputSaveFrequency": 1, "DenseOutputMinimumTimeStep": 0.0,
            "BiCGRhoTolerance": 1e-6, "BiCGResidualTolerance": 1e-6, "BiCGMaxIterations": 19,
            "RelativeTolerance": [1e-8] * UD, "AbsoluteTolerance": [1e-8] * UD,
            "EventTolerance": [1e-6] * max(NE, 1), "EventDirection": [0] * max(NE, 1),
            "Algorithm": 1,
        }

    @classmethod
    def from_ensemble(cls, ens, NDO=0):
        p = cls(ens.NS, ens.N, NDO=NDO)
        p.TimeDomain[: ens.NS] = 0.0
        p.TimeDomain[ens.NS:] = 1.0
        p.ActualState[:] = ens.unit_soa(ens.state0).reshape(-1)
        p.UnitParameters[:] = ens.unit_soa(ens.params).reshape(-1)
        p.IntegerUnitAccessories[:] = -1
        if ens.coupling.ndim != 2:
            raise ValueError("the oracle takes one coupling matrix for all systems")
        p.CouplingMatrix[:] = ens.coupling_colmajor()
        p.opts.update(ens.options)
        return p

    def reset_accessories(self):
        """ResetAccessories of the collapse-strength driver (CollapseStrength.cu:373-396)."""
        self.UnitAccessories[:] = 0.0
        self.IntegerUnitAccessories[:] = -1

    def solve(self, mode=CANONICAL, nthreads=0):
        o = self.opts
        self._rt = np.asarray(o["RelativeTolerance"], dtype=np.float64)
        self._at = np.asarray(o["AbsoluteTolerance"], dtype=np.float64)
        self._et = np.asarray(o["EventTolerance"], dtype=np.float64)
        self._ed = np.asarray(o["EventDirection"], dtype=np.int32)
        opt = Options(o["InitialTimeStep"], int(o.get("ActiveSystems", self.dims.NS)),
                      o["MaximumTimeStep"], o["MinimumTimeStep"], o["TimeStepGrowLimit"],
                      o["TimeStepShrinkLimit"], int(o["DenseOutputSaveFrequency"]),
                      o["DenseOutputMinimumTimeStep"], o["BiCGRhoTolerance"],
                      o["BiCGResidualTolerance"], int(o["BiCGMaxIterations"]),
                      _dp(self._rt), _dp(self._at), _dp(self._et), _ip(self._ed),
                      int(o.get("Algorithm", 1)), int(o.get("DenseOutputInterpolation", 0)))
        buf = Buffers(_dp(self.TimeDomain), _dp(self.ActualState), _dp(self.ActualTime),
                      _dp(self.UnitParameters), _dp(self.SystemParameters),
                      _dp(self.GlobalParameters), _ip(self.IntegerGlobalParameters),
                      _dp(self.UnitAccessories), _ip(self.IntegerUnitAccessories),
                      _dp(self.SystemAccessories), _ip(self.IntegerSystemAccessories),
                      _dp(self.CouplingMatrix), _dp(self.DenseOutputTimeInstances),
                      _dp(self.DenseOutputStates), _ip(self.DenseOutputIndex))
        rc = lib().bex_oracle_solve(ctypes.byref(self.dims), ctypes.byref(opt),
                                    ctypes.byref(buf),
                                    self.stats.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                    mode, nthreads)
        if rc != 0:
            raise RuntimeError(f"bex_oracle_solve failed with {rc}")

    # Views -----------------------------------------------------------------------------
    def state(self):
        NS, N, UD = self.dims.NS, self.dims.N, self.dims.UD
        return self.ActualState.reshape(UD, NS, N).transpose(1, 2, 0)

    def accessories(self):
        NS, N, K = self.dims.NS, self.dims.N, self.dims.NUA
        return self.UnitAccessories.reshape(K, NS, N).transpose(1, 2, 0)

    def event_counts(self):
        NS, N = self.dims.NS, self.dims.N
        return self.IntegerUnitAccessories.reshape(-1, NS, N)[0] + 1

    def stats_table(self):
        return self.stats.reshape(self.dims.NS, NSTAT)


def bicg(M, x1, rD, CPT, xu, rho_tol=1e-6, res_tol=1e-6, max_iter=19, mode=CANONICAL):
    N = len(x1)
    Mc = np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).reshape(-1)  # column-major
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x1, rD, CPT, xu)]
    x = np.zeros(N)
    it = lib().bex_oracle_bicg(N, _dp(Mc), *[_dp(a) for a in arrs], rho_tol, res_tol, max_iter,
                               mode, _dp(x))
    return x, it


def km_rhs(X, T, upar, mode=CANONICAL):
    X = np.ascontiguousarray(X, dtype=np.float64)
    upar = np.ascontiguousarray(upar, dtype=np.float64)
    F = np.zeros(2)
    cpt = np.zeros(1)
    cpf = np.zeros(1)
    lib().bex_oracle_km_rhs(_dp(X), float(T), _dp(upar), mode, _dp(F), _dp(cpt), _dp(cpf))
    return F, cpt[0], cpf[0]


def rkck45_linear(lam, y0, h, mode=CANONICAL):
    y = ctypes.c_double()
    e = ctypes.c_double()
    ok = lib().bex_oracle_rkck45_linear(lam, y0, h, mode, ctypes.byref(y), ctypes.byref(e))
    return y.value, e.value, ok


def rk4_linear(lam, y0, h, mode=CANONICAL):
    y = ctypes.c_double()
    ok = lib().bex_oracle_rk4_linear(lam, y0, h, mode, ctypes.byref(y))
    return y.value, ok


def control(update_step, is_finite, rel_err, dt, mode=CANONICAL, MinimumTimeStep=1e-12,
            MaximumTimeStep=1e6, TimeStepGrowLimit=5.0, TimeStepShrinkLimit=0.1):
    opt = Options()
    opt.MinimumTimeStep = MinimumTimeStep
    opt.MaximumTimeStep = MaximumTimeStep
    opt.TimeStepGrowLimit = TimeStepGrowLimit
    opt.TimeStepShrinkLimit = TimeStepShrinkLimit
    dt_new = ctypes.c_double()
    upd, term, forced = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().bex_oracle_control(update_step, is_finite, rel_err, dt, ctypes.byref(opt), mode,
                             ctypes.byref(dt_new), ctypes.byref(upd), ctypes.byref(term),
                             ctypes.byref(forced))
    return dt_new.value, upd.value, term.value, forced.value


def pow_(x, y, mode=CANONICAL):
    return lib().bex_oracle_pow(x, y, mode)


def sincospi(x, mode=CANONICAL):
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib().bex_oracle_sincospi(x, mode, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sin_(x, mode=CANONICAL):
    return lib().bex_oracle_sin(x, mode)


def cos_(x, mode=CANONICAL):
    return lib().bex_oracle_cos(x, mode)
