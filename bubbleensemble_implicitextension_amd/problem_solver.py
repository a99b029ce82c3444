"""Python mirror of the reference's ProblemSolver host API over libbex.so.

Same enumerations, member names, overloads and argument meaning as
SourceCodes/CoupledSystems_PerBlock_Interface.cuh:33-248. Errors the reference treats as fatal
(bound checks, wrong overload/variable combinations, runtime failures) raise ``BexError``
instead of calling exit(); the C++ header CoupledSystems_PerBlock_Interface.hpp restores the
print-and-exit behaviour for C++ callers.
"""
from __future__ import annotations

import ctypes
import enum

import numpy as np

from . import _native


class Algorithms(enum.IntEnum):  # Interface.cuh:33
    RK4 = 0
    RKCK45 = 1


class ListOfVariables(enum.IntEnum):  # Interface.cuh:35-53
    All = 0
    TimeDomain = 1
    ActualTime = 2
    ActualState = 3
    UnitParameters = 4
    SystemParameters = 5
    GlobalParameters = 6
    IntegerGlobalParameters = 7
    UnitAccessories = 8
    IntegerUnitAccessories = 9
    SystemAccessories = 10
    IntegerSystemAccessories = 11
    CouplingMatrix = 12
    CouplingStrength = 13
    CouplingIndex = 14
    DenseOutput = 15
    DenseIndex = 16
    DenseTime = 17
    DenseState = 18


class ListOfSolverOptions(enum.IntEnum):  # Interface.cuh:55-68, then additions
    InitialTimeStep = 0
    ActiveSystems = 1
    MaximumTimeStep = 2
    MinimumTimeStep = 3
    TimeStepGrowLimit = 4
    TimeStepShrinkLimit = 5
    RelativeTolerance = 6
    AbsoluteTolerance = 7
    EventTolerance = 8
    EventDirection = 9
    DenseOutputSaveFrequency = 10
    DenseOutputMinimumTimeStep = 11
    SharedGlobalVariables = 12
    SharedCouplingMatrices = 13
    BiCGRhoTolerance = 14
    BiCGResidualTolerance = 15
    BiCGMaxIterations = 16
    KernelLayout = 17
    PriorityStepThreshold = 18
    WarmStart = 19
    DenseOutputInterpolation = 20
    MaximumStepAttempts = 21
    PersistentLaunch = 22
    PersistentWorkgroups = 23


LAYOUT_AUTO, LAYOUT_GROUP, LAYOUT_ROLE_PAIR, LAYOUT_SYSTEM_PAIR = 0, 1, 2, 3
MLP_NPARAM = 4609  # warm-start network parameters (csrc/bex_mlp.h)
SAMPLE_FLOATS = 10  # floats per training-sample row


# Reference-style bare names (the reference uses unscoped enums).
globals().update({m.name: m for m in Algorithms})
globals().update({m.name: m for m in ListOfVariables})
globals().update({m.name: m for m in ListOfSolverOptions})

_UNIT_VARS = {ListOfVariables.ActualState, ListOfVariables.UnitParameters,
              ListOfVariables.UnitAccessories, ListOfVariables.IntegerUnitAccessories}


def compiled_dims():
    d = _native.ProblemDesc()
    _native.check(_native.load().bex_compiled_dims(ctypes.byref(d)))
    return {f: getattr(d, f) for f, _ in d._fields_}


# -- device helpers (Interface.cuh:259-345) ----------------------------------------------------
def list_devices():
    """Properties of every visible device, as ListCUDADevices prints them (a list of dicts;
    empty without a device)."""
    lib = _native.load()
    n = ctypes.c_int()
    _native.check(lib.bex_device_count(ctypes.byref(n)))
    out = []
    for i in range(n.value):
        p = _native.DeviceProperties()
        _native.check(lib.bex_get_device_properties(i, ctypes.byref(p)))
        d = {f: getattr(p, f) for f, _ in p._fields_}
        d["name"], d["arch"] = p.name.decode(), p.arch.decode()
        d["max_threads_dim"], d["max_grid_size"] = list(p.max_threads_dim), list(p.max_grid_size)
        out.append(d)
    return out


def SelectDeviceByClosestRevision(major, minor):
    """The device whose architecture revision is closest to major.minor."""
    dev = ctypes.c_int()
    _native.check(_native.load().bex_select_device_by_closest_revision(int(major), int(minor),
                                                                       ctypes.byref(dev)))
    return dev.value


class ProblemSolver:
    """One solver object: pinned host + device buffers for NS systems of UPS units.

    Constructor arguments are the reference's template parameters
    (Interface.cuh:135) plus the device ordinal.
    """

    def __init__(self, NS, UPS, UD=2, TPB=None, SPB=1, NC=1, CBW=0, CCI=0, NUP=21, NSP=0, NGP=0,
                 NiGP=0, NUA=64, NiUA=1, NSA=0, NiSA=0, NE=1, NDO=0,
                 Algorithm=Algorithms.RKCK45, device=0, _handle=None):
        if TPB is None:
            TPB = SPB * UPS
        self.desc = _native.ProblemDesc(NS, UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP, NiGP,
                                        NUA, NiUA, NSA, NiSA, NE, NDO, int(Algorithm))
        self._lib = _native.load()
        self._owned = _handle is None
        if _handle is None:
            h = ctypes.c_void_p()
            _native.check(self._lib.bex_create(ctypes.byref(self.desc), device, ctypes.byref(h)))
            self._h = h
        else:
            self._h = _handle
        self.NS, self.UPS, self.UD, self.NDO = NS, UPS, UD, NDO
        self.dims = {f: getattr(self.desc, f) for f, _ in self.desc._fields_}

    # -- lifetime --------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                self._lib.bex_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        return _native.check(rc, self._h)

    # -- options (Interface.cuh:771-865) ------------------------------------------------
    def SolverOption(self, option, *args):
        option = ListOfSolverOptions(option)
        if len(args) == 1:
            self._check(self._lib.bex_solver_option(self._h, int(option), float(args[0])))
        elif len(args) == 2:
            self._check(self._lib.bex_solver_option_indexed(self._h, int(option), int(args[0]),
                                                             float(args[1])))
        else:
            raise TypeError("SolverOption(option, value) or SolverOption(option, index, value)")

    # -- element access (Interface.cuh:867-1140, 1338-1529) -------------------------------
    def SetHost(self, *args):
        fn, call = self._dispatch(args, setter=True)
        self._check(fn(self._h, *call))

    def GetHost(self, *args):
        out = ctypes.c_double()
        fn, call = self._dispatch(args + (None,), setter=False)
        self._check(fn(self._h, *call[:-1], ctypes.byref(out)))
        return out.value

    def _dispatch(self, args, setter):
        L = self._lib
        pre = "bex_set_host_" if setter else "bex_get_host_"
        isvar = [isinstance(a, ListOfVariables) for a in args]
        n = len(args)
        v = args[-1]
        val = float(v) if setter else None
        if n == 6 and isvar[2]:  # (sid, uid, DenseState, comp, k, v)
            return getattr(L, pre + "dense_state"), [int(args[0]), int(args[1]), int(args[2]),
                                                     int(args[3]), int(args[4]), val]
        if n == 5 and isvar[2]:  # (sid, uid, var, k, v)
            return getattr(L, pre + "unit"), [int(args[0]), int(args[1]), int(args[2]),
                                              int(args[3]), val]
        if n == 5 and isvar[1]:  # (c, CouplingMatrix, row, col, v)
            return getattr(L, pre + "coupling"), [int(args[0]), int(args[1]), int(args[2]),
                                                  int(args[3]), val]
        if n == 4 and isvar[1]:  # (sid, var, k, v)
            return getattr(L, pre + "system"), [int(args[0]), int(args[1]), int(args[2]), val]
        if n == 3 and isvar[0]:  # (var, k, v)
            return getattr(L, pre + "global"), [int(args[0]), int(args[1]), val]
        if n == 3 and isvar[1]:  # (sid, DenseIndex, v)
            return getattr(L, pre + "dense_index"), [int(args[0]), int(args[1]), val]
        raise TypeError("no SetHost/GetHost overload takes these arguments")

    def host_view(self, var):
        """Zero-copy numpy view of a pinned host buffer in the reference layout."""
        ptr = ctypes.c_void_p()
        count = ctypes.c_longlong()
        esz = ctypes.c_int()
        self._check(self._lib.bex_host_view(self._h, int(ListOfVariables(var)), ctypes.byref(ptr),
                                            ctypes.byref(count), ctypes.byref(esz)))
        if count.value == 0:
            return np.zeros(0, dtype=np.float64 if esz.value == 8 else np.int32)
        ctype = ctypes.c_double if esz.value == 8 else ctypes.c_int
        arr = (ctype * count.value).from_address(ptr.value)
        view = np.ctypeslib.as_array(arr)
        view.flags.writeable = True
        return view

    # -- synchronisation (Interface.cuh:1142-1336, 1949-1983) -----------------------------
    def SynchroniseFromHostToDevice(self, var):
        self._check(self._lib.bex_sync_h2d(self._h, int(ListOfVariables(var))))

    def SynchroniseFromDeviceToHost(self, var):
        self._check(self._lib.bex_sync_d2h(self._h, int(ListOfVariables(var))))

    def Solve(self):
        self._check(self._lib.bex_solve(self._h))

    def SynchroniseDevice(self):
        self._check(self._lib.bex_synchronise_device(self._h))

    def InsertSynchronisationPoint(self):
        self._check(self._lib.bex_insert_synchronisation_point(self._h))

    def SynchroniseSolver(self):
        self._check(self._lib.bex_synchronise_solver(self._h))

    # -- device checkpoint (no reference counterpart: the reference checkpoints through the
    #    host with SynchroniseFromDeviceToHost(All), Interface.cuh:1310-1328) --------------
    def checkpoint_save(self):
        """Snapshot, in device memory, of everything a Solve() writes back (time domain, state,
        time, accessories, dense-output index); asynchronous on the solver's stream."""
        self._check(self._lib.bex_checkpoint_save(self._h))

    def checkpoint_restore(self):
        """Copies the last device snapshot back (asynchronous on the solver's stream)."""
        self._check(self._lib.bex_checkpoint_restore(self._h))

    # -- persistent launch (no reference counterpart) ------------------------------------
    def persistent_grid(self):
        """(workgroups, waves per workgroup) of the persistent launch the next Solve() starts."""
        wg, w = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.bex_persistent_grid(self._h, ctypes.byref(wg), ctypes.byref(w)))
        return wg.value, w.value

    def set_warm_start_systems(self, flags):
        """Which systems take the learned BiCG start when WarmStart is on: one flag per system
        (False: the plain start, results bit-identical to the plain solve); None: every
        system."""
        if flags is None:
            self._check(self._lib.bex_set_warm_start_systems(self._h, None, 0))
            return
        f = np.ascontiguousarray(np.asarray(flags, dtype=bool), dtype=np.uint8)
        self._check(self._lib.bex_set_warm_start_systems(
            self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), int(f.size)))

    def set_launch_order(self, order):
        """Order in which Solve()'s persistent launch hands out the systems: every system
        index once, longest expected run first; None returns to index order. Results do not
        depend on it."""
        if order is None:
            self._check(self._lib.bex_set_launch_order(self._h, None, 0))
            return
        o = np.ascontiguousarray(order, dtype=np.int32)
        self._check(self._lib.bex_set_launch_order(
            self._h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(o.size)))

    # -- counters ----------------------------------------------------------------------
    def stats(self):
        out = np.zeros(self.NS * _native.NSTAT, dtype=np.int64)
        self._check(self._lib.bex_get_stats(self._h,
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        return out.reshape(self.NS, _native.NSTAT)

    def timer_start(self):
        self._check(self._lib.bex_timer_start(self._h))

    def timer_stop(self):
        ms = ctypes.c_double()
        self._check(self._lib.bex_timer_stop(self._h, ctypes.byref(ms)))
        return ms.value

    def enable_timeline(self, on=True):
        self._check(self._lib.bex_enable_timeline(self._h, 1 if on else 0))

    def timeline(self):
        """[NS, 2] start/end seconds of each system in the last launch (relative to the
        earliest start)."""
        out = np.zeros(2 * self.NS, dtype=np.int64)
        tps = ctypes.c_double()
        self._check(self._lib.bex_get_timeline(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), ctypes.byref(tps)))
        t = out.reshape(self.NS, 2).astype(np.float64)
        valid = t[:, 0] > 0
        if valid.any():
            t -= t[valid, 0].min()
        return t / tps.value

    def request_stop(self):
        """Asks every running system to end at its next check (every 256 step attempts), state
        written back as it stands; stays in force until clear_stop()."""
        self._check(self._lib.bex_request_stop(self._h))

    def clear_stop(self):
        self._check(self._lib.bex_clear_stop(self._h))

    def finished_systems(self):
        """Systems of the current (or last) launch that have finished, read while the launch
        may still run (needs enable_timeline, which clears the count)."""
        out = np.zeros(2 * self.NS, dtype=np.int64)
        self._check(self._lib.bex_peek_timeline(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        return int(np.count_nonzero(out[1::2]))

    def reset_stats(self):
        self._check(self._lib.bex_reset_stats(self._h))

    # -- BiCG warm start (no reference counterpart) ---------------------------------------
    def set_warm_start_network(self, params):
        """Uploads a trained network (4609 FP32 parameters, layout of csrc/bex_mlp.h)."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32).reshape(-1))
        if p.size != MLP_NPARAM:
            raise ValueError(f"expected {MLP_NPARAM} parameters, got {p.size}")
        self._check(self._lib.bex_set_warm_start_network(
            self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def enable_sampling(self, capacity, every):
        self._check(self._lib.bex_enable_sampling(self._h, int(capacity), int(every)))

    def samples(self):
        """Collected training rows [n, 10]: 5 features, x/b (clamped), x, b, uncoupled
        acceleration, 0."""
        n = ctypes.c_longlong()
        self._check(self._lib.bex_get_samples(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, SAMPLE_FLOATS), dtype=np.float32)
        if n.value:
            self._check(self._lib.bex_get_samples(
                self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n.value,
                ctypes.byref(n)))
        return out

    # -- text dumps (Interface.cuh:1531-1947) ---------------------------------------------
    def thread_layout(self):
        d = self.dims
        return ThreadLayout.make(d["NS"], d["UPS"], d["TPB"], d["SPB"])

    def Print(self, var, serial=None, directory="."):
        """Print(var): the host copy of `var` to <var>InSolverObject.txt; Print(CouplingMatrix,
        k): the original coupling matrix k to CouplingMatrixInSolverObject_k.txt;
        Print(DenseOutput, sid): system sid's report and stored time series to
        DenseOutput_sid.txt. Same file names and text as the reference; files go to
        `directory` (the working directory by default)."""
        import os

        var = ListOfVariables(var)
        d = self.dims
        t = self.thread_layout()
        V = ListOfVariables
        if serial is None:
            path = os.path.join(directory, f"{var.name}InSolverObject.txt")
            unit = {V.ActualState: d["UD"], V.UnitParameters: d["NUP"],
                    V.UnitAccessories: d["NUA"], V.IntegerUnitAccessories: d["NiUA"]}
            system = {V.TimeDomain: (d["NS"], 2), V.ActualTime: (d["NS"], 1),
                      V.SystemParameters: (d["NS"], d["NSP"]),
                      V.GlobalParameters: (d["NGP"], 1),
                      V.IntegerGlobalParameters: (d["NiGP"], 1),
                      V.SystemAccessories: (d["NS"], d["NSA"]),
                      V.IntegerSystemAccessories: (d["NS"], d["NiSA"]),
                      V.CouplingStrength: (d["NS"], d["NC"]), V.CouplingIndex: (d["NC"], 1)}
            if var in unit:
                write_unit_scope(path, unit[var], self.host_view(var), t)
            elif var in system:
                rows, cols = system[var]
                write_system_scope(path, rows, cols, self.host_view(var))
            else:
                raise _native.BexError(-3, f"Print: option {var.name}: this option needs "
                                           "different argument configuration or not applicable")
            return path
        k = int(serial)
        if var == V.CouplingMatrix:
            _print_bound("CouplingMatrix", k, d["NC"] - 1)
            N = d["UPS"]
            m = np.array([[self.GetHost(k, V.CouplingMatrix, r, c) for r in range(N)]
                          for c in range(N)]).reshape(-1)  # column-major
            path = os.path.join(directory, f"CouplingMatrixInSolverObject_{k}.txt")
            write_system_scope(path, N, N, m)
            return path
        if var == V.DenseOutput:
            _print_bound("DenseOutput", k, d["NS"] - 1)
            path = os.path.join(directory, f"DenseOutput_{k}.txt")
            write_dense_report(path, int(self.host_view(V.DenseIndex)[k]), k, t, d,
                               {v: self.host_view(v) for v in (
                                   V.UnitParameters, V.UnitAccessories, V.IntegerUnitAccessories,
                                   V.SystemParameters, V.SystemAccessories,
                                   V.IntegerSystemAccessories, V.GlobalParameters,
                                   V.IntegerGlobalParameters, V.DenseTime, V.DenseState)})
            return path
        raise _native.BexError(-3, f"Print: option {var.name}: this option needs different "
                                   "argument configuration or not applicable")

    # -- convenience views (reference layout with TPB == SPB*UPS) --------------------------
    def state(self):
        return self.host_view(ListOfVariables.ActualState).reshape(
            self.UD, self.NS, self.UPS).transpose(1, 2, 0)

    def accessories(self):
        v = self.host_view(ListOfVariables.UnitAccessories)
        return v.reshape(-1, self.NS, self.UPS).transpose(1, 2, 0)

    def event_counts(self):
        v = self.host_view(ListOfVariables.IntegerUnitAccessories)
        return v.reshape(-1, self.NS, self.UPS)[0] + 1


# -- text dumps in the reference's format (Interface.cuh:1531-1805) ------------------------
# Fields right-aligned in 18 characters, floating-point values in scientific notation with 10
# digits (C++ iostream std::scientific, precision 10), each followed by a comma.

class ThreadLayout:
    """Logical-thread layout of the reference (Interface.cuh:386-391, 875-879)."""

    def __init__(self, NS, UPS, SPB, ltpb, pad, tlt):
        self.NS, self.UPS, self.SPB = NS, UPS, SPB
        self.LogicalThreadsPerBlock, self.ThreadPaddingPerBlock = ltpb, pad
        self.TotalLogicalThreads = tlt

    @classmethod
    def make(cls, NS, UPS, TPB, SPB):
        ltpb = SPB * UPS
        launches = ltpb // TPB + (1 if ltpb % TPB else 0)
        pad = launches * TPB - ltpb
        grid = NS // SPB + (1 if NS % SPB else 0)
        return cls(NS, UPS, SPB, ltpb, pad, (ltpb + pad) * grid)

    def thread_of(self, sid, uid):
        return ((sid // self.SPB) * (self.LogicalThreadsPerBlock + self.ThreadPaddingPerBlock)
                + (sid % self.SPB) * self.UPS + uid)


def _field(v):
    if isinstance(v, str):
        return f"{v:>18}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>18d}"
    return f"{float(v):>18.10e}"


def _fields(values):
    return [_field(v) for v in values]


def write_unit_scope(path, columns, data, t):
    """One row per logical thread, one column per component (data[k * TLT + thread]); padded
    threads and the unused threads of a partly filled last block print PADDING."""
    rows = t.TotalLogicalThreads
    per_block = t.LogicalThreadsPerBlock + t.ThreadPaddingPerBlock
    limit = (t.NS // t.SPB) * per_block + (t.NS % t.SPB) * t.UPS
    data = np.asarray(data)
    cast = int if data.dtype.kind in "iu" else float
    with open(path, "w") as f:
        for i in range(rows):
            real = (i % per_block) < t.LogicalThreadsPerBlock and i < limit
            row = [(_field(cast(data[i + j * rows])) if real else _field("PADDING")) + ","
                   for j in range(columns)]
            f.write("".join(row) + "\n")


def write_system_scope(path, rows, columns, data):
    """rows x columns, column-major data[i + j * rows]."""
    data = np.asarray(data)
    cast = int if data.dtype.kind in "iu" else float
    with open(path, "w") as f:
        for i in range(rows):
            f.write("".join(_field(cast(data[i + j * rows])) + "," for j in range(columns)) + "\n")


def write_dense_report(path, points, sid, t, dims, bufs):
    """Per-system report: unit tables, system and global values, then the stored time series
    (one row per point: t, then U<j>/C<k> for every unit and component)."""
    V = ListOfVariables
    UPS, UD, tlt, NS = t.UPS, dims["UD"], t.TotalLogicalThreads, t.NS
    out = []

    def unit_table(title, count, data):
        data = np.asarray(data)
        cast = int if data.dtype.kind in "iu" else float
        out.append(f"{title}:\n")
        out.append(_field("Serial No.") + "," + ",".join(_field(f"U{u}") for u in range(UPS)) + "\n")
        for k in range(count):
            vals = [cast(data[t.thread_of(sid, u) + k * tlt]) for u in range(UPS)]
            out.append(_field(k) + "," + ",".join(_fields(vals)) + "\n")
        out.append("\n")

    def row(title, count, data, first, stride):
        data = np.asarray(data)
        cast = int if data.dtype.kind in "iu" else float
        out.append(f"{title}:\n")
        out.append(",".join(_fields(cast(data[first + k * stride]) for k in range(count))))
        out.append("\n\n")

    unit_table("UnitParameters", dims["NUP"], bufs[V.UnitParameters])
    unit_table("UnitAccessories", dims["NUA"], bufs[V.UnitAccessories])
    unit_table("IntegerUnitAccessories", dims["NiUA"], bufs[V.IntegerUnitAccessories])
    row("SystemParameters", dims["NSP"], bufs[V.SystemParameters], sid, NS)
    row("SystemAccessories", dims["NSA"], bufs[V.SystemAccessories], sid, NS)
    row("IntegerSystemAccessories", dims["NiSA"], bufs[V.IntegerSystemAccessories], sid, NS)
    row("GlobalParameters", dims["NGP"], bufs[V.GlobalParameters], 0, 1)
    row("IntegerGlobalParameters", dims["NiGP"], bufs[V.IntegerGlobalParameters], 0, 1)
    out.append("Time series:\n")
    out.append(_field("t") + "," + ",".join(_field(f"U{u}/C{c}") for u in range(UPS)
                                            for c in range(UD)) + "\n")
    times, states = np.asarray(bufs[V.DenseTime]), np.asarray(bufs[V.DenseState])
    for i in range(points):
        vals = [float(states[t.thread_of(sid, u) + c * tlt + i * tlt * UD])
                for u in range(UPS) for c in range(UD)]
        out.append(_field(float(times[sid + i * NS])) + "," + ",".join(_fields(vals)) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))


def _print_bound(what, index, last):
    """BoundCheck of Print (Interface.cuh:718-733)."""
    if 0 <= index <= last:
        return
    acceptable = "none" if last < 0 else f"0-{last}"
    raise _native.BexError(-2, f"Print: option {what}: acceptable index {acceptable}, current "
                               f"index {index}")


class ProblemPool:
    """Uneven ensemble: size classes of systems with their own coupling matrices, solved by
    persistent kernels with device work queues, one hipGraph launch per period (C ABI
    bex_pool_*; no reference counterpart).

    solver(c) returns class c's ProblemSolver (options, SetHost/GetHost, host views and
    synchronisation work as for a stand-alone solver; its Solve() runs that class alone)."""

    def __init__(self, units, systems, NUP=21, NUA=64, NiUA=1, NE=1, device=0):
        self._lib = _native.load()
        units = [int(u) for u in units]
        systems = [int(n) for n in systems]
        if len(units) != len(systems):
            raise ValueError("units and systems must have the same length")
        n = len(units)
        h = ctypes.c_void_p()
        _native.pool_check(self._lib.bex_pool_create(
            n, (ctypes.c_int * n)(*units), (ctypes.c_int * n)(*systems), device, ctypes.byref(h)))
        self._h = h
        self.units, self.systems = units, systems
        self._solvers = []
        for c in range(n):
            sh = ctypes.c_void_p()
            _native.pool_check(self._lib.bex_pool_class_solver(self._h, c, ctypes.byref(sh)), self._h)
            self._solvers.append(ProblemSolver(systems[c], units[c], NUP=NUP, NUA=NUA, NiUA=NiUA,
                                               NE=NE, _handle=sh))

    def _check(self, rc):
        return _native.pool_check(rc, self._h)

    def close(self):
        if getattr(self, "_h", None):
            for s in self._solvers:
                s.close()
            self._lib.bex_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def solver(self, c):
        return self._solvers[c]

    def set_couplings(self, c, M, first=0):
        """M: [count, N, N] with M[s, i, j] the coupling of unit i to unit j of system
        first + s."""
        M = np.asarray(M, dtype=np.float64)
        colmajor = np.ascontiguousarray(np.swapaxes(M, -1, -2))
        self._check(self._lib.bex_pool_set_couplings(
            self._h, c, first, M.shape[0], colmajor.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def set_order(self, c, order):
        o = np.ascontiguousarray(np.asarray(order, dtype=np.int32))
        self._check(self._lib.bex_pool_set_order(self._h, c,
                                                 o.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def order_by_cost(self, c):
        self._check(self._lib.bex_pool_order_by_cost(self._h, c))

    def get_order(self, c):
        """The order in which class c's systems are handed out."""
        o = np.zeros(self.systems[c], dtype=np.int32)
        self._check(self._lib.bex_pool_get_order(self._h, c,
                                                 o.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return o

    def Solve(self):
        self._check(self._lib.bex_pool_solve(self._h))

    def solve_periods(self, periods):
        self._check(self._lib.bex_pool_solve_periods(self._h, int(periods)))

    def synchronise(self):
        self._check(self._lib.bex_pool_synchronise(self._h))

    def variants(self, c):
        n = self.units[c]
        return [self._lib.bex_pool_variant_name(n, v).decode()
                for v in range(self._lib.bex_pool_variant_count(n))]

    def set_variant(self, c, v):
        self._check(self._lib.bex_pool_set_variant(self._h, c, int(v)))

    def variant(self, c):
        v = ctypes.c_int()
        self._check(self._lib.bex_pool_get_variant(self._h, c, ctypes.byref(v)))
        return v.value

    def autotune(self, probe_fraction=0.05):
        """Returns {class: {variant name: probe ms}}, leaves the fastest variant selected and
        each class's work queue ordered by the probe's per-system work, longest first."""
        counts = [self._lib.bex_pool_variant_count(n) for n in self.units]
        times = (ctypes.c_double * sum(counts))()
        self._check(self._lib.bex_pool_autotune(self._h, float(probe_fraction), times))
        out, k = {}, 0
        for c, nv in enumerate(counts):
            names = self.variants(c)
            out[c] = {names[v]: times[k + v] for v in range(nv)}
            k += nv
        return out
