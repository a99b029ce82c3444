// Drop-in C++ host interface of the MI355X coupled-systems per-block solver.
//
// Same class template, enumerations, member functions, overloads, argument meaning and error
// behaviour as the reference's host interface
// (SourceCodes/CoupledSystems_PerBlock_Interface.cuh:33-248): an application written against
// the reference's ProblemSolver compiles against this header unchanged and links against
// libbex.so, whose kernel is built for the system definition the application would otherwise
// #include before the interface. The integration is FP64 for Precision = float too.
//
// Errors the reference treats as fatal (bound checks, wrong overload/variable combination,
// HIP failures, memory shortfall) print a message to std::cerr and exit(EXIT_FAILURE), as in
// Interface.cuh:13-22 and :718-733.
#ifndef COUPLEDSYSTEMS_PERBLOCK_INTERFACE_HPP
#define COUPLEDSYSTEMS_PERBLOCK_INTERFACE_HPP

#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <utility>
#include <vector>
#include <type_traits>

#include "bex_capi.h"

enum Algorithms { RK4 = BEX_RK4, RKCK45 = BEX_RKCK45 };  // Interface.cuh:33

enum ListOfVariables {  // Interface.cuh:35-53
  All = BEX_All,
  TimeDomain,
  ActualTime,
  ActualState,
  UnitParameters,
  SystemParameters,
  GlobalParameters,
  IntegerGlobalParameters,
  UnitAccessories,
  IntegerUnitAccessories,
  SystemAccessories,
  IntegerSystemAccessories,
  CouplingMatrix,
  CouplingStrength,
  CouplingIndex,
  DenseOutput,
  DenseIndex,
  DenseTime,
  DenseState
};

enum ListOfSolverOptions {  // Interface.cuh:55-68, then the BiCG constants
  InitialTimeStep = BEX_InitialTimeStep,
  ActiveSystems,
  MaximumTimeStep,
  MinimumTimeStep,
  TimeStepGrowLimit,
  TimeStepShrinkLimit,
  RelativeTolerance,
  AbsoluteTolerance,
  EventTolerance,
  EventDirection,
  DenseOutputSaveFrequency,
  DenseOutputMinimumTimeStep,
  SharedGlobalVariables,
  SharedCouplingMatrices,
  BiCGRhoTolerance,
  BiCGResidualTolerance,
  BiCGMaxIterations,
  KernelLayout,
  PriorityStepThreshold,
  WarmStart,
  DenseOutputInterpolation,
  MaximumStepAttempts,
  PersistentLaunch,
  PersistentWorkgroups
};

// --- Names of the enumerators (Interface.cuh:70-71, 2035-2106) ----------------------------

inline std::string SolverOptionsToString(ListOfSolverOptions Option) {
  static const char* const names[] = {
      "InitialTimeStep",       "ActiveSystems",          "MaximumTimeStep",
      "MinimumTimeStep",       "TimeStepGrowLimit",      "TimeStepShrinkLimit",
      "RelativeTolerance",     "AbsoluteTolerance",      "EventTolerance",
      "EventDirection",        "DenseOutputSaveFrequency", "DenseOutputMinimumTimeStep",
      "SharedGlobalVariables", "SharedCouplingMatrices", "BiCGRhoTolerance",
      "BiCGResidualTolerance", "BiCGMaxIterations",      "KernelLayout",
      "PriorityStepThreshold", "WarmStart",              "DenseOutputInterpolation",
      "MaximumStepAttempts",   "PersistentLaunch",       "PersistentWorkgroups"};
  const int i = static_cast<int>(Option) - static_cast<int>(InitialTimeStep);
  const int n = static_cast<int>(sizeof(names) / sizeof(names[0]));
  return (i >= 0 && i < n) ? names[i] : "Non-existent solver option!";
}

inline std::string VariablesToString(ListOfVariables Option) {
  static const char* const names[] = {
      "All",            "TimeDomain",        "ActualTime",
      "ActualState",    "UnitParameters",    "SystemParameters",
      "GlobalParameters", "IntegerGlobalParameters", "UnitAccessories",
      "IntegerUnitAccessories", "SystemAccessories", "IntegerSystemAccessories",
      "CouplingMatrix", "CouplingStrength",  "CouplingIndex",
      "DenseOutput",    "DenseIndex",        "DenseTime",
      "DenseState"};
  const int i = static_cast<int>(Option) - static_cast<int>(All);
  const int n = static_cast<int>(sizeof(names) / sizeof(names[0]));
  return (i >= 0 && i < n) ? names[i] : "Non-existent variable!";
}

// --- Text dumps of host buffers (Interface.cuh:1531-1805) ---------------------------------
// One implementation, written in this header's own words, of the reference's three file
// writers: fields right-aligned in 18 characters, floating-point values in scientific notation
// with 10 digits, every field followed by a comma.

namespace bex_detail {

constexpr int kDumpWidth = 18;  // Interface.cuh:1540-1542

inline void DumpFormat(std::ofstream& f) {
  f.precision(10);
  f.flags(std::ios::scientific);
}

// Reference layout of the logical threads (Interface.cuh:386-391, 875-879).
struct ThreadLayout {
  int NS, UPS, SPB, LogicalThreadsPerBlock, ThreadPaddingPerBlock;
  long long TotalLogicalThreads;
  long long ThreadOf(int sid, int uid) const {
    return (long long)(sid / SPB) * (LogicalThreadsPerBlock + ThreadPaddingPerBlock) +
           (long long)(sid % SPB) * UPS + uid;
  }
};

inline ThreadLayout MakeThreadLayout(int NS, int UPS, int TPB, int SPB) {
  ThreadLayout t{NS, UPS, SPB, SPB * UPS, 0, 0};
  const int launches = t.LogicalThreadsPerBlock / TPB + (t.LogicalThreadsPerBlock % TPB ? 1 : 0);
  t.ThreadPaddingPerBlock = launches * TPB - t.LogicalThreadsPerBlock;
  const int grid = NS / SPB + (NS % SPB ? 1 : 0);
  t.TotalLogicalThreads = (long long)(t.LogicalThreadsPerBlock + t.ThreadPaddingPerBlock) * grid;
  return t;
}

// Unit scope: one row per logical thread, one column per component; padded threads and the
// unused threads of a partly filled last block print PADDING (Interface.cuh:1531-1575).
template <typename T>
void WriteUnitScope(const std::string& name, int columns, const T* data, const ThreadLayout& t) {
  std::ofstream f(name.c_str());
  DumpFormat(f);
  const long long rows = t.TotalLogicalThreads;
  const long long perBlock = t.LogicalThreadsPerBlock + t.ThreadPaddingPerBlock;
  const long long limit = (long long)(t.NS / t.SPB) * perBlock + (long long)(t.NS % t.SPB) * t.UPS;
  for (long long i = 0; i < rows; i++) {
    const bool real = (i % perBlock) < t.LogicalThreadsPerBlock && i < limit;
    for (int j = 0; j < columns; j++) {
      f.width(kDumpWidth);
      if (real)
        f << data[i + (long long)j * rows];
      else
        f << "PADDING";
      f << ',';
    }
    f << '\n';
  }
}

// System and global scope: rows x columns, column-major (Interface.cuh:1577-1601).
template <typename T>
void WriteSystemScope(const std::string& name, long long rows, int columns, const T* data) {
  std::ofstream f(name.c_str());
  DumpFormat(f);
  for (long long i = 0; i < rows; i++) {
    for (int j = 0; j < columns; j++) {
      f.width(kDumpWidth);
      f << data[i + (long long)j * rows] << ',';
    }
    f << '\n';
  }
}

// Host buffers and dimensions the per-system dense-output report reads.
struct DenseReport {
  ThreadLayout t;
  int UD, NUP, NUA, NiUA, NSP, NSA, NiSA, NGP, NiGP;
  const double *UnitParameters, *UnitAccessories, *SystemParameters, *SystemAccessories,
      *GlobalParameters, *TimeInstances, *States;
  const int *IntegerUnitAccessories, *IntegerSystemAccessories, *IntegerGlobalParameters;
};

template <typename T>
void ReportUnitTable(std::ofstream& f, const char* title, int count, const T* data,
                     const ThreadLayout& t, int sid) {
  f << title << ":\n";
  f.width(kDumpWidth);
  f << "Serial No." << ',';
  for (int u = 0; u < t.UPS; u++) {
    f.width(kDumpWidth);
    f << "U" + std::to_string(u);
    if (u < t.UPS - 1) f << ',';
  }
  f << '\n';
  for (int k = 0; k < count; k++) {
    f.width(kDumpWidth);
    f << k << ',';
    for (int u = 0; u < t.UPS; u++) {
      f.width(kDumpWidth);
      f << data[t.ThreadOf(sid, u) + (long long)k * t.TotalLogicalThreads];
      if (u < t.UPS - 1) f << ',';
    }
    f << '\n';
  }
  f << '\n';
}

template <typename T>
void ReportRow(std::ofstream& f, const char* title, int count, const T* data, long long first,
               long long stride) {
  f << title << ":\n";
  for (int k = 0; k < count; k++) {
    f.width(kDumpWidth);
    f << data[first + (long long)k * stride];
    if (k < count - 1) f << ',';
  }
  f << '\n' << '\n';
}

// Per-system report: its unit tables, system and global values, then its dense-output time
// series, one row per stored point (Interface.cuh:1603-1805).
inline void WriteDenseOutputReport(const std::string& name, int points, int sid,
                                   const DenseReport& r) {
  std::ofstream f(name.c_str());
  DumpFormat(f);
  const ThreadLayout& t = r.t;
  ReportUnitTable(f, "UnitParameters", r.NUP, r.UnitParameters, t, sid);
  ReportUnitTable(f, "UnitAccessories", r.NUA, r.UnitAccessories, t, sid);
  ReportUnitTable(f, "IntegerUnitAccessories", r.NiUA, r.IntegerUnitAccessories, t, sid);
  ReportRow(f, "SystemParameters", r.NSP, r.SystemParameters, sid, t.NS);
  ReportRow(f, "SystemAccessories", r.NSA, r.SystemAccessories, sid, t.NS);
  ReportRow(f, "IntegerSystemAccessories", r.NiSA, r.IntegerSystemAccessories, sid, t.NS);
  ReportRow(f, "GlobalParameters", r.NGP, r.GlobalParameters, 0, 1);
  ReportRow(f, "IntegerGlobalParameters", r.NiGP, r.IntegerGlobalParameters, 0, 1);
  f << "Time series:\n";
  f.width(kDumpWidth);
  f << "t" << ',';
  for (int u = 0; u < t.UPS; u++)
    for (int c = 0; c < r.UD; c++) {
      f.width(kDumpWidth);
      f << "U" + std::to_string(u) + "/C" + std::to_string(c);
      if (u < t.UPS - 1 || c < r.UD - 1) f << ',';
    }
  f << '\n';
  const long long stateSize = t.TotalLogicalThreads * r.UD;
  for (int i = 0; i < points; i++) {
    f.width(kDumpWidth);
    f << r.TimeInstances[sid + (long long)i * t.NS] << ',';
    for (int u = 0; u < t.UPS; u++)
      for (int c = 0; c < r.UD; c++) {
        f.width(kDumpWidth);
        f << r.States[t.ThreadOf(sid, u) + (long long)c * t.TotalLogicalThreads + i * stateSize];
        if (u < t.UPS - 1 || c < r.UD - 1) f << ',';
      }
    f << '\n';
  }
}

}  // namespace bex_detail

// --- Device helpers (Interface.cuh:73-75, 259-345) --------------------------------------
// Same names and output as the reference's free functions, for the HIP devices. The
// "compute capability" is the architecture revision HIP reports (9.5 for gfx950), followed by
// the target id.

namespace bex_detail {
inline void DeviceCheck(int rc, const char* function) {
  if (rc == BEX_OK) return;
  std::cerr << "ERROR: In function " << function << "!" << std::endl;
  std::cerr << "       " << bex_last_error(nullptr) << std::endl;
  std::exit(EXIT_FAILURE);
}

inline void PrintDeviceBody(const bex_device_properties& p) {
  std::cout << "--------------------------------------" << std::endl;
  std::cout << "Total global memory:        " << p.total_global_mem / 1024 / 1024 << " Mb" << std::endl;
  std::cout << "Total shared memory:        " << p.shared_mem_per_block / 1024 << " Kb" << std::endl;
  std::cout << "Number of 32-bit registers: " << p.regs_per_block << std::endl;
  std::cout << "Total constant memory:      " << p.total_const_mem / 1024 << " Kb" << std::endl;
  std::cout << std::endl;
  std::cout << "Number of multiprocessors:  " << p.multiprocessor_count << std::endl;
  std::cout << "Compute capability:         " << p.major << "." << p.minor << " (" << p.arch << ")"
            << std::endl;
  std::cout << "Core clock rate:            " << p.clock_rate_khz / 1000 << " MHz" << std::endl;
  std::cout << "Memory clock rate:          " << p.memory_clock_rate_khz / 1000 << " MHz" << std::endl;
  std::cout << "Memory bus width:           " << p.memory_bus_width << " bits" << std::endl;
  std::cout << "Peak memory bandwidth:      "
            << 2.0 * p.memory_clock_rate_khz * (p.memory_bus_width / 8) / 1.0e6 << " GB/s"
            << std::endl;
  std::cout << std::endl;
  std::cout << "Warp size:                  " << p.warp_size << std::endl;
  std::cout << "Max. warps per multiproc:   " << p.max_threads_per_multiprocessor / p.warp_size
            << std::endl;
  std::cout << "Max. threads per multiproc: " << p.max_threads_per_multiprocessor << std::endl;
  std::cout << "Max. threads per block:     " << p.max_threads_per_block << std::endl;
  std::cout << "Max. block dimensions:      " << p.max_threads_dim[0] << " * "
            << p.max_threads_dim[1] << " * " << p.max_threads_dim[2] << std::endl;
  std::cout << "Max. grid dimensions:       " << p.max_grid_size[0] << " * " << p.max_grid_size[1]
            << " * " << p.max_grid_size[2] << std::endl;
  std::cout << std::endl;
  std::cout << "Concurrent memory copy:     " << p.concurrent_copy << std::endl;
  std::cout << "Execution multiple kernels: " << p.concurrent_kernels << std::endl;
  std::cout << "ECC support turned on:      " << p.ecc_enabled << std::endl << std::endl;
}
}  // namespace bex_detail

inline void ListCUDADevices() {  // Interface.cuh:259-294
  int n = 0;
  bex_detail::DeviceCheck(bex_device_count(&n), "ListCUDADevices");
  for (int i = 0; i < n; i++) {
    bex_device_properties p;
    bex_detail::DeviceCheck(bex_get_device_properties(i, &p), "ListCUDADevices");
    std::cout << std::endl;
    std::cout << "Device number: " << i << std::endl;
    std::cout << "Device name:   " << p.name << std::endl;
    bex_detail::PrintDeviceBody(p);
  }
  std::cout << std::endl;
}

inline int SelectDeviceByClosestRevision(int MajorRevision, int MinorRevision) {  // :296-310
  int device = 0;
  bex_detail::DeviceCheck(bex_select_device_by_closest_revision(MajorRevision, MinorRevision, &device),
                          "SelectDeviceByClosestRevision");
  std::cout << "HIP Device Number Closest to Revision " << MajorRevision << "." << MinorRevision
            << ": " << device << std::endl << std::endl << std::endl;
  return device;
}

inline void PrintPropertiesOfSpecificDevice(int SelectedDevice) {  // :312-345
  bex_device_properties p;
  bex_detail::DeviceCheck(bex_get_device_properties(SelectedDevice, &p),
                          "PrintPropertiesOfSpecificDevice");
  std::cout << "Selected device number: " << SelectedDevice << std::endl;
  std::cout << "Selected device name:   " << p.name << std::endl;
  bex_detail::PrintDeviceBody(p);
  std::cout << std::endl;
}

template <int NS, int UPS, int UD, int TPB, int SPB, int NC, int CBW, int CCI, int NUP, int NSP,
          int NGP, int NiGP, int NUA, int NiUA, int NSA, int NiSA, int NE, int NDO,
          Algorithms Algorithm, class Precision>
class ProblemSolver {
  // The reference instantiates Precision = float or double (Interface.cuh:135). Both are
  // accepted here; the solver integrates in FP64 either way (collapse stiffness needs it), so
  // with float the values are converted at SetHost/GetHost and the results are the FP64 solve
  // of the float inputs rather than a single-precision solve.
  static_assert(std::is_same<Precision, double>::value || std::is_same<Precision, float>::value,
                "Precision must be float or double");

 private:
  bex_solver* handle_ = nullptr;

  void Check(int rc, const char* function) const {
    if (rc == BEX_OK) return;
    std::cerr << "ERROR: In solver member function " << function << "!" << std::endl;
    std::cerr << "       " << bex_last_error(handle_) << std::endl;
    std::exit(EXIT_FAILURE);
  }

  // Host copy of a variable (the pinned buffer the Set/Get calls and synchronisation use).
  template <typename T>
  const T* HostBuffer(ListOfVariables var) {
    void* ptr = nullptr;
    long long count = 0;
    int esz = 8;
    Check(bex_host_view(handle_, var, &ptr, &count, &esz), "Print");
    return static_cast<const T*>(ptr);
  }

  static bex_detail::ThreadLayout Layout() { return bex_detail::MakeThreadLayout(NS, UPS, TPB, SPB); }

  // The original (uncompacted) coupling matrix k as GetHost returns it (Interface.cuh:1076).
  std::vector<double> OriginalCoupling(int k) {
    std::vector<double> m((size_t)UPS * UPS);
    for (int col = 0; col < UPS; col++)
      for (int row = 0; row < UPS; row++)
        Check(bex_get_host_coupling(handle_, k, CouplingMatrix, row, col,
                                    &m[row + (size_t)col * UPS]),
              "Print");
    return m;
  }

 public:
  // Interface.cuh:352-657
  explicit ProblemSolver(int device) {
    bex_problem_desc d{NS,  UPS, UD,   TPB, SPB,  NC,  CBW, CCI, NUP,      NSP,
                       NGP, NiGP, NUA, NiUA, NSA, NiSA, NE,  NDO, Algorithm};
    int rc = bex_create(&d, device, &handle_);
    if (rc != BEX_OK) {
      std::cerr << "ERROR: creating the solver object failed!" << std::endl;
      std::cerr << "       " << bex_last_error(nullptr) << std::endl;
      std::exit(EXIT_FAILURE);
    }
  }
  // Interface.cuh:660-714
  ~ProblemSolver() {
    if (handle_) bex_destroy(handle_);
  }
  ProblemSolver(const ProblemSolver&) = delete;
  ProblemSolver& operator=(const ProblemSolver&) = delete;

  // Interface.cuh:771-865
  template <typename T>
  void SolverOption(ListOfSolverOptions option, T value) {
    Check(bex_solver_option(handle_, option, (double)value), "SolverOption");
  }
  template <typename T>
  void SolverOption(ListOfSolverOptions option, int index, T value) {
    Check(bex_solver_option_indexed(handle_, option, index, (double)value), "SolverOption");
  }

  // Interface.cuh:867-1140
  template <typename T>
  void SetHost(int sid, int uid, ListOfVariables var, int k, T v) {
    Check(bex_set_host_unit(handle_, sid, uid, var, k, (double)v), "SetHost");
  }
  template <typename T>
  void SetHost(int sid, ListOfVariables var, int k, T v) {
    Check(bex_set_host_system(handle_, sid, var, k, (double)v), "SetHost");
  }
  template <typename T>
  void SetHost(ListOfVariables var, int k, T v) {
    Check(bex_set_host_global(handle_, var, k, (double)v), "SetHost");
  }
  template <typename T>
  void SetHost(int c, ListOfVariables var, int row, int col, T v) {
    Check(bex_set_host_coupling(handle_, c, var, row, col, (double)v), "SetHost");
  }
  template <typename T>
  void SetHost(int sid, ListOfVariables var, T v) {
    Check(bex_set_host_dense_index(handle_, sid, var, (double)v), "SetHost");
  }
  template <typename T>
  void SetHost(int sid, int uid, ListOfVariables var, int comp, int k, T v) {
    Check(bex_set_host_dense_state(handle_, sid, uid, var, comp, k, (double)v), "SetHost");
  }

  // Interface.cuh:1142-1336
  void SynchroniseFromHostToDevice(ListOfVariables var) {
    Check(bex_sync_h2d(handle_, var), "SynchroniseFromHostToDevice");
  }
  void SynchroniseFromDeviceToHost(ListOfVariables var) {
    Check(bex_sync_d2h(handle_, var), "SynchroniseFromDeviceToHost");
  }

  // Interface.cuh:1338-1529
  template <typename T>
  T GetHost(int sid, int uid, ListOfVariables var, int k) {
    double v;
    Check(bex_get_host_unit(handle_, sid, uid, var, k, &v), "GetHost");
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, ListOfVariables var, int k) {
    double v;
    Check(bex_get_host_system(handle_, sid, var, k, &v), "GetHost");
    return (T)v;
  }
  template <typename T>
  T GetHost(ListOfVariables var, int k) {
    double v;
    Check(bex_get_host_global(handle_, var, k, &v), "GetHost");
    return (T)v;
  }
  template <typename T>
  T GetHost(int c, ListOfVariables var, int row, int col) {
    double v;
    Check(bex_get_host_coupling(handle_, c, var, row, col, &v), "GetHost");
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, ListOfVariables var) {
    double v;
    Check(bex_get_host_dense_index(handle_, sid, var, &v), "GetHost");
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, int uid, ListOfVariables var, int comp, int k) {
    double v;
    Check(bex_get_host_dense_state(handle_, sid, uid, var, comp, k, &v), "GetHost");
    return (T)v;
  }

  // Interface.cuh:1807-1920: the host copy of a variable to <Variable>InSolverObject.txt in the
  // working directory (unit scope: one row per logical thread; system scope: one row per
  // system; global scope: one row per value).
  void Print(ListOfVariables var) {
    const bex_detail::ThreadLayout t = Layout();
    const std::string file = VariablesToString(var) + "InSolverObject.txt";
    switch (var) {
      case TimeDomain: bex_detail::WriteSystemScope(file, NS, 2, HostBuffer<double>(var)); break;
      case ActualTime: bex_detail::WriteSystemScope(file, NS, 1, HostBuffer<double>(var)); break;
      case ActualState: bex_detail::WriteUnitScope(file, UD, HostBuffer<double>(var), t); break;
      case UnitParameters: bex_detail::WriteUnitScope(file, NUP, HostBuffer<double>(var), t); break;
      case SystemParameters:
        bex_detail::WriteSystemScope(file, NS, NSP, HostBuffer<double>(var));
        break;
      case GlobalParameters:
        bex_detail::WriteSystemScope(file, NGP, 1, HostBuffer<double>(var));
        break;
      case IntegerGlobalParameters:
        bex_detail::WriteSystemScope(file, NiGP, 1, HostBuffer<int>(var));
        break;
      case UnitAccessories: bex_detail::WriteUnitScope(file, NUA, HostBuffer<double>(var), t); break;
      case IntegerUnitAccessories:
        bex_detail::WriteUnitScope(file, NiUA, HostBuffer<int>(var), t);
        break;
      case SystemAccessories:
        bex_detail::WriteSystemScope(file, NS, NSA, HostBuffer<double>(var));
        break;
      case IntegerSystemAccessories:
        bex_detail::WriteSystemScope(file, NS, NiSA, HostBuffer<int>(var));
        break;
      case CouplingStrength:
        bex_detail::WriteSystemScope(file, NS, NC, HostBuffer<double>(var));
        break;
      case CouplingIndex: bex_detail::WriteSystemScope(file, NC, 1, HostBuffer<int>(var)); break;
      default: PrintOptionError(var);
    }
  }

  // Interface.cuh:1922-1947: Print(CouplingMatrix, k) writes the original matrix k to
  // CouplingMatrixInSolverObject_k.txt; Print(DenseOutput, sid) writes system sid's report and
  // stored time series to DenseOutput_sid.txt. Both check the index first.
  void Print(ListOfVariables var, int serial) {
    if (var == CouplingMatrix) {
      PrintBoundCheck("CouplingMatrix", serial, NC - 1);
      const std::vector<double> m = OriginalCoupling(serial);
      bex_detail::WriteSystemScope("CouplingMatrixInSolverObject_" + std::to_string(serial) + ".txt",
                                   UPS, UPS, m.data());
    } else if (var == DenseOutput) {
      PrintBoundCheck("DenseOutput", serial, NS - 1);
      bex_detail::DenseReport r{Layout(), UD, NUP, NUA, NiUA, NSP, NSA, NiSA, NGP, NiGP,
                                HostBuffer<double>(UnitParameters),
                                HostBuffer<double>(UnitAccessories),
                                HostBuffer<double>(SystemParameters),
                                HostBuffer<double>(SystemAccessories),
                                HostBuffer<double>(GlobalParameters),
                                HostBuffer<double>(DenseTime),
                                HostBuffer<double>(DenseState),
                                HostBuffer<int>(IntegerUnitAccessories),
                                HostBuffer<int>(IntegerSystemAccessories),
                                HostBuffer<int>(IntegerGlobalParameters)};
      const int points = HostBuffer<int>(DenseIndex)[serial];
      bex_detail::WriteDenseOutputReport("DenseOutput_" + std::to_string(serial) + ".txt", points,
                                         serial, r);
    } else {
      PrintOptionError(var);
    }
  }

 private:
  // Interface.cuh:718-733 (BoundCheck) and :1942-1945.
  static void PrintBoundCheck(const char* what, int index, int last) {
    if (index >= 0 && index <= last) return;
    std::cerr << "ERROR: In solver member function Print!" << std::endl;
    std::cerr << "       Option: " << what << std::endl;
    if (last < 0)
      std::cerr << "       Acceptable index: none" << std::endl;
    else
      std::cerr << "       Acceptable index: 0-" << last << std::endl;
    std::cerr << "       Current index:    " << index << std::endl;
    std::exit(EXIT_FAILURE);
  }
  static void PrintOptionError(ListOfVariables var) {
    std::cerr << "ERROR: In solver member function Print!" << std::endl;
    std::cerr << "       Option: " << VariablesToString(var) << std::endl;
    std::cerr << "       This option needs different argument configuration or not applicable!"
              << std::endl;
    std::exit(EXIT_FAILURE);
  }

 public:
  // Interface.cuh:1949-1983
  void Solve() { Check(bex_solve(handle_), "Solve"); }
  void SynchroniseDevice() { Check(bex_synchronise_device(handle_), "SynchroniseDevice"); }
  void InsertSynchronisationPoint() {
    Check(bex_insert_synchronisation_point(handle_), "InsertSynchronisationPoint");
  }
  void SynchroniseSolver() { Check(bex_synchronise_solver(handle_), "SynchroniseSolver"); }

  // Additions: the underlying handle and the per-system counters.
  bex_solver* Handle() { return handle_; }

  // Learned BiCG start (no reference counterpart): 4609 FP32 parameters of the network in
  // bex_mlp.h; enable with SolverOption(WarmStart, 1).
  void SetWarmStartNetwork(const float* params) {
    Check(bex_set_warm_start_network(handle_, params), "SetWarmStartNetwork");
  }

  // Device checkpoint of everything a Solve() writes back (no reference counterpart; the
  // reference checkpoints through the host, Interface.cuh:1310-1328).
  void CheckpointSave() { Check(bex_checkpoint_save(handle_), "CheckpointSave"); }
  void CheckpointRestore() { Check(bex_checkpoint_restore(handle_), "CheckpointRestore"); }

  // Hand-out order of the persistent launch (no reference counterpart): every system once,
  // longest expected run first; an empty vector returns to index order.
  void SetLaunchOrder(const std::vector<int>& order) {
    Check(bex_set_launch_order(handle_, order.empty() ? nullptr : order.data(),
                               static_cast<int>(order.size())),
          "SetLaunchOrder");
  }

  // Which systems take the learned BiCG start when WarmStart is on (no reference counterpart):
  // one flag per system, 0 = plain start; an empty vector: every system.
  void SetWarmStartSystems(const std::vector<unsigned char>& flags) {
    Check(bex_set_warm_start_systems(handle_, flags.empty() ? nullptr : flags.data(),
                                     static_cast<int>(flags.size())),
          "SetWarmStartSystems");
  }
};

// Split solver (no reference counterpart): one ensemble held by two solver objects that run
// together, each a persistent launch on its own stream — the `heavy` systems (on the learned
// BiCG start when a network is given) and the `medium` systems (plain start) one per wave in
// one launch, at the top issue priority, and the others two per wave (system-pair layout,
// 32 units; one per wave otherwise). The member functions are ProblemSolver's, with the system
// indices of the whole ensemble; the plain-start systems' results equal ProblemSolver's bit for
// bit. Print is not provided (each part is an ordinary solver object; Part(k) gives its handle
// for the C calls).
template <int NS, int UPS, int UD, int TPB, int SPB, int NC, int CBW, int CCI, int NUP, int NSP,
          int NGP, int NiGP, int NUA, int NiUA, int NSA, int NiSA, int NE, int NDO,
          Algorithms Algorithm, class Precision>
class SplitProblemSolver {
  static_assert(std::is_same<Precision, double>::value || std::is_same<Precision, float>::value,
                "Precision must be float or double");

 private:
  struct Piece {
    bex_solver* h = nullptr;
    std::vector<int> sids;  // whole-ensemble indices of the part's systems, ascending
  };
  std::vector<Piece> parts_;
  std::vector<int> part_of_, local_of_;

  void Check(int rc, const char* function, const bex_solver* h) const {
    if (rc == BEX_OK) return;
    std::cerr << "ERROR: In solver member function " << function << "!" << std::endl;
    std::cerr << "       " << bex_last_error(h) << std::endl;
    std::exit(EXIT_FAILURE);
  }
  // The part and local index of system sid (the reference's bound check on the system index).
  std::pair<bex_solver*, int> Route(int sid, const char* function) const {
    if (sid < 0 || sid >= NS) {
      std::cerr << "ERROR: In solver member function " << function << "!" << std::endl;
      std::cerr << "       Option: SystemNumber, index " << sid
                << " out of range (acceptable index: 0-" << NS - 1 << ")" << std::endl;
      std::exit(EXIT_FAILURE);
    }
    return {parts_[part_of_[sid]].h, local_of_[sid]};
  }

 public:
  SplitProblemSolver(int device, const std::vector<int>& heavy, const std::vector<int>& medium,
                     const float* network = nullptr)
      : part_of_(NS, 2), local_of_(NS, 0) {
    // tier: 0 heavy, 1 medium, 2 light; part: 0 heavy and medium, 1 light
    std::vector<int> tier(NS, 2);
    for (int k = 0; k < 2; k++) {
      for (int sid : (k == 0 ? heavy : medium)) {
        if (sid < 0 || sid >= NS || tier[sid] != 2) {
          std::cerr << "ERROR: SplitProblemSolver: system " << sid
                    << " is out of range or listed twice" << std::endl;
          std::exit(EXIT_FAILURE);
        }
        tier[sid] = k;
      }
    }
    std::vector<std::vector<int>> groups(2);
    for (int sid = 0; sid < NS; sid++) groups[tier[sid] < 2 ? 0 : 1].push_back(sid);
    std::vector<int> slot(2, -1);
    for (int k = 0; k < 2; k++) {
      if (groups[k].empty()) continue;
      Piece part;
      part.sids = groups[k];
      bex_problem_desc d{(int)part.sids.size(), UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP,
                         NiGP, NUA, NiUA, NSA, NiSA, NE, NDO, Algorithm};
      if (bex_create(&d, device, &part.h) != BEX_OK) {
        std::cerr << "ERROR: creating the solver object failed!" << std::endl;
        std::cerr << "       " << bex_last_error(nullptr) << std::endl;
        std::exit(EXIT_FAILURE);
      }
      if (k == 0) {
        Check(bex_solver_option(part.h, BEX_PriorityStepThreshold, 1.0), "SplitProblemSolver", part.h);
        if (network && !heavy.empty()) {
          Check(bex_set_warm_start_network(part.h, network), "SplitProblemSolver", part.h);
          Check(bex_solver_option(part.h, BEX_WarmStart, 1.0), "SplitProblemSolver", part.h);
          std::vector<unsigned char> learned(part.sids.size());
          for (size_t i = 0; i < part.sids.size(); i++) learned[i] = tier[part.sids[i]] == 0;
          Check(bex_set_warm_start_systems(part.h, learned.data(), (int)learned.size()),
                "SplitProblemSolver", part.h);
        }
      } else if (UPS == 32) {
        Check(bex_solver_option(part.h, BEX_KernelLayout, BEX_LAYOUT_SYSTEM_PAIR),
              "SplitProblemSolver", part.h);
      }
      slot[k] = (int)parts_.size();
      for (int i = 0; i < (int)part.sids.size(); i++) local_of_[part.sids[i]] = i;
      parts_.push_back(part);
    }
    for (int sid = 0; sid < NS; sid++) part_of_[sid] = slot[tier[sid] < 2 ? 0 : 1];
  }
  ~SplitProblemSolver() {
    for (Piece& p : parts_)
      if (p.h) bex_destroy(p.h);
  }
  SplitProblemSolver(const SplitProblemSolver&) = delete;
  SplitProblemSolver& operator=(const SplitProblemSolver&) = delete;

  int Parts() const { return (int)parts_.size(); }
  bex_solver* Part(int k) { return parts_.at(k).h; }
  const std::vector<int>& PartSystems(int k) const { return parts_.at(k).sids; }

  // Options apply to every part (a KernelLayout given here replaces the parts' own layouts);
  // ActiveSystems n keeps systems 0 .. n-1 of the ensemble, i.e. each part's systems below n,
  // which are its first local ones.
  template <typename T>
  void SolverOption(ListOfSolverOptions option, T value) {
    for (Piece& p : parts_) {
      double v = (double)value;
      if (option == ActiveSystems) {
        int m = 0;
        while (m < (int)p.sids.size() && p.sids[m] < (int)value) m++;
        v = (double)m;
      }
      Check(bex_solver_option(p.h, option, v), "SolverOption", p.h);
    }
  }
  template <typename T>
  void SolverOption(ListOfSolverOptions option, int index, T value) {
    for (Piece& p : parts_)
      Check(bex_solver_option_indexed(p.h, option, index, (double)value), "SolverOption", p.h);
  }

  template <typename T>
  void SetHost(int sid, int uid, ListOfVariables var, int k, T v) {
    auto r = Route(sid, "SetHost");
    Check(bex_set_host_unit(r.first, r.second, uid, var, k, (double)v), "SetHost", r.first);
  }
  template <typename T>
  void SetHost(int sid, ListOfVariables var, int k, T v) {
    auto r = Route(sid, "SetHost");
    Check(bex_set_host_system(r.first, r.second, var, k, (double)v), "SetHost", r.first);
  }
  template <typename T>
  void SetHost(ListOfVariables var, int k, T v) {
    for (Piece& p : parts_) Check(bex_set_host_global(p.h, var, k, (double)v), "SetHost", p.h);
  }
  template <typename T>
  void SetHost(int c, ListOfVariables var, int row, int col, T v) {
    for (Piece& p : parts_)
      Check(bex_set_host_coupling(p.h, c, var, row, col, (double)v), "SetHost", p.h);
  }
  template <typename T>
  void SetHost(int sid, ListOfVariables var, T v) {
    auto r = Route(sid, "SetHost");
    Check(bex_set_host_dense_index(r.first, r.second, var, (double)v), "SetHost", r.first);
  }
  template <typename T>
  void SetHost(int sid, int uid, ListOfVariables var, int comp, int k, T v) {
    auto r = Route(sid, "SetHost");
    Check(bex_set_host_dense_state(r.first, r.second, uid, var, comp, k, (double)v), "SetHost",
          r.first);
  }

  template <typename T>
  T GetHost(int sid, int uid, ListOfVariables var, int k) {
    auto r = Route(sid, "GetHost");
    double v;
    Check(bex_get_host_unit(r.first, r.second, uid, var, k, &v), "GetHost", r.first);
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, ListOfVariables var, int k) {
    auto r = Route(sid, "GetHost");
    double v;
    Check(bex_get_host_system(r.first, r.second, var, k, &v), "GetHost", r.first);
    return (T)v;
  }
  template <typename T>
  T GetHost(ListOfVariables var, int k) {
    double v;
    Check(bex_get_host_global(parts_[0].h, var, k, &v), "GetHost", parts_[0].h);
    return (T)v;
  }
  template <typename T>
  T GetHost(int c, ListOfVariables var, int row, int col) {
    double v;
    Check(bex_get_host_coupling(parts_[0].h, c, var, row, col, &v), "GetHost", parts_[0].h);
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, ListOfVariables var) {
    auto r = Route(sid, "GetHost");
    double v;
    Check(bex_get_host_dense_index(r.first, r.second, var, &v), "GetHost", r.first);
    return (T)v;
  }
  template <typename T>
  T GetHost(int sid, int uid, ListOfVariables var, int comp, int k) {
    auto r = Route(sid, "GetHost");
    double v;
    Check(bex_get_host_dense_state(r.first, r.second, uid, var, comp, k, &v), "GetHost", r.first);
    return (T)v;
  }

  void SynchroniseFromHostToDevice(ListOfVariables var) {
    for (Piece& p : parts_) Check(bex_sync_h2d(p.h, var), "SynchroniseFromHostToDevice", p.h);
  }
  void SynchroniseFromDeviceToHost(ListOfVariables var) {
    for (Piece& p : parts_) Check(bex_sync_d2h(p.h, var), "SynchroniseFromDeviceToHost", p.h);
  }
  // Every part's launch, heavy first; asynchronous like ProblemSolver::Solve().
  void Solve() {
    for (Piece& p : parts_) Check(bex_solve(p.h), "Solve", p.h);
  }
  void SynchroniseDevice() { Check(bex_synchronise_device(parts_[0].h), "SynchroniseDevice", parts_[0].h); }
  void InsertSynchronisationPoint() {
    for (Piece& p : parts_)
      Check(bex_insert_synchronisation_point(p.h), "InsertSynchronisationPoint", p.h);
  }
  void SynchroniseSolver() {
    for (Piece& p : parts_) Check(bex_synchronise_solver(p.h), "SynchroniseSolver", p.h);
  }
};

// Uneven ensembles (no reference counterpart): size classes of systems with their own
// coupling matrices, one graph launch per period. Class(c) is a solver handle for the C calls
// (bex_set_host_*, bex_get_host_*, bex_solver_option*, bex_sync_*).
class ProblemPool {
 private:
  bex_pool* pool_ = nullptr;
  std::vector<int> ns_;  // systems per class

  void Check(int rc, const char* function) const {
    if (rc == BEX_OK) return;
    std::cerr << "ERROR: In pool member function " << function << "!" << std::endl;
    std::cerr << "       " << bex_pool_last_error(pool_) << std::endl;
    std::exit(EXIT_FAILURE);
  }

 public:
  ProblemPool(const std::vector<int>& units, const std::vector<int>& systems, int device) {
    if (units.size() != systems.size()) {
      std::cerr << "ERROR: ProblemPool: units and systems differ in length" << std::endl;
      std::exit(EXIT_FAILURE);
    }
    ns_ = systems;
    if (bex_pool_create((int)units.size(), units.data(), systems.data(), device, &pool_) !=
        BEX_OK) {
      std::cerr << "ERROR: creating the pool failed!" << std::endl;
      std::cerr << "       " << bex_pool_last_error(nullptr) << std::endl;
      std::exit(EXIT_FAILURE);
    }
  }
  ~ProblemPool() {
    if (pool_) bex_pool_destroy(pool_);
  }
  ProblemPool(const ProblemPool&) = delete;
  ProblemPool& operator=(const ProblemPool&) = delete;

  bex_solver* Class(int c) {
    bex_solver* s = nullptr;
    Check(bex_pool_class_solver(pool_, c, &s), "Class");
    return s;
  }
  void SetCouplings(int c, int first, int count, const double* M) {
    Check(bex_pool_set_couplings(pool_, c, first, count, M), "SetCouplings");
  }
  void SetOrder(int c, const std::vector<int>& order) {
    Check(bex_pool_set_order(pool_, c, order.data()), "SetOrder");
  }
  void OrderByCost(int c) { Check(bex_pool_order_by_cost(pool_, c), "OrderByCost"); }
  std::vector<int> GetOrder(int c) {
    std::vector<int> order((size_t)ns_.at(c));
    Check(bex_pool_get_order(pool_, c, order.data()), "GetOrder");
    return order;
  }
  void Autotune(double probe_fraction) {
    Check(bex_pool_autotune(pool_, probe_fraction, nullptr), "Autotune");
  }
  void Solve() { Check(bex_pool_solve(pool_), "Solve"); }
  void SolvePeriods(int periods) { Check(bex_pool_solve_periods(pool_, periods), "SolvePeriods"); }
  void Synchronise() { Check(bex_pool_synchronise(pool_), "Synchronise"); }
  bex_pool* Handle() { return pool_; }
};

#endif
