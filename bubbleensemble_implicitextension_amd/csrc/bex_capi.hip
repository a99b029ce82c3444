// Host side of libbex.so: the C ABI declared in include/bex_capi.h.
//
// A bex_solver is the MI355X counterpart of one reference ProblemSolver object
// (SourceCodes/CoupledSystems_PerBlock_Interface.cuh): it owns pinned host buffers and device
// buffers in the reference's structure-of-arrays layout, an in-order HIP stream and an event,
// and launches the solver kernel of bex_solver_device.h once per Solve().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bex_capi.h"
#include "bex_solver_device.h"

namespace {

std::string g_create_error;

struct Buf {
  void* h = nullptr;
  void* d = nullptr;
  long long n = 0;
  int esz = 8;
  size_t bytes() const { return (size_t)(n > 0 ? n : 1) * esz; }
};

}  // namespace

struct bex_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  hipEvent_t timer0 = nullptr, timer1 = nullptr;
  bex_problem_desc d{};
  // Thread configuration of the reference layout (Interface.cuh:386-391).
  int LogicalThreadsPerBlock = 0, ThreadPaddingPerBlock = 0, GridSize = 0;
  long long TLT = 0;
  long long CompactCouplingSize = 0;  // per coupling, reference storage (Interface.cuh:428-438)

  Buf TimeDomain, ActualState, ActualTime, UnitParameters, SystemParameters, GlobalParameters,
      IntegerGlobalParameters, UnitAccessories, IntegerUnitAccessories, SystemAccessories,
      IntegerSystemAccessories, CouplingMatrix, CouplingStrength, CouplingIndex,
      DenseOutputIndex, DenseOutputTimeInstances, DenseOutputStates;
  double* h_CouplingOriginal = nullptr;  // full [c][N][N], host only (Interface.cuh:1076)
  double* d_CouplingFull = nullptr;      // full [c][N][N] expanded from the compact storage
  double* d_CouplingFullT = nullptr;     // transpose of coupling 0
  bool PerSystemCoupling = false;        // pool classes: one full matrix per system
  struct bex_pool* pool = nullptr;       // pool classes: the owning pool and class index
  int pool_class = -1;
  double* d_CouplingSys = nullptr;       // [NS][N][N] column-major (pool classes)
  double* d_CouplingSysT = nullptr;      // their transposes (pool classes that stream M)
  double* h_tol = nullptr;               // [2*UD + NE] rtol, atol, event tolerance
  int* h_dir = nullptr;                  // [NE]
  double* d_tol = nullptr;
  int* d_dir = nullptr;
  long long* d_stats = nullptr;
  // [NS][2] start/end wall-clock ticks, when enabled: pinned host memory mapped into the
  // device's address space (the kernel stores to it over the bus, the host reads it without
  // any GPU command, also while a launch runs); d_timeline is its device address.
  long long* h_timeline = nullptr;
  long long* d_timeline = nullptr;
  int* h_stop = nullptr;  // stop flag in host-mapped memory (bex_request_stop)
  int* d_stop = nullptr;  // its device address

  double InitialTimeStep = 1e-2, MaximumTimeStep = 1.0e6, MinimumTimeStep = 1.0e-12;
  double TimeStepGrowLimit = 5.0, TimeStepShrinkLimit = 0.1, DenseOutputMinimumTimeStep = 0.0;
  int ActiveSystems = 0, DenseOutputSaveFrequency = 1;
  double BiCGRhoTolerance = 1e-6, BiCGResidualTolerance = 1e-6;
  int BiCGMaxIterations = 19;
  int KernelLayout = BEX_LAYOUT_AUTO;
  int PersistentLaunch = 1;  // 1: persistent waves over a device work queue (bex_solve)
  int PersistentWorkgroups = 0;  // > 0: cap on the persistent launch's workgroups
  unsigned long long* d_queue = nullptr;  // its counter
  int* d_order = nullptr;                 // hand-out order of the systems (nullptr: index order)
  unsigned char* d_warm_systems = nullptr;  // [NS] which systems take the learned start
  int cu_count = 0;
  int PriorityStepThreshold = 3000;
  long long MaxStepAttempts = 0;  // 0: no cap
  int WarmStart = 0;                          // 1: learned BiCG initial guess
  BexMlpDevice* d_net = nullptr;              // network in the kernel's operand layout
  float* d_samples = nullptr;                 // [SampleCapacity][BEX_SAMPLE_FLOATS] samples
  unsigned long long* d_sample_count = nullptr;
  long long SampleCapacity = 0;
  int SampleEvery = 0;
  int DenseInterpolation = 0;         // 1: Hermite dense output at a uniform grid
  double* d_DenseAux = nullptr;       // [2 UD][TLT], allocated with the option
  bool SharedGlobalVariables = true, SharedCouplingMatrices = false;
  unsigned char* d_ckpt = nullptr;    // device checkpoint (bex_checkpoint_save/restore)
  int config_epoch = 0;  // bumped by every option change (pools re-capture their graph)
  std::string err;
};

namespace {

int fail(bex_solver* s, int code, const std::string& msg) {
  if (s)
    s->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define BEX_HIP(s, call)                                                              \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((s), BEX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int bound(bex_solver* s, const char* fn, const char* what, long long v, long long lo, long long hi) {
  if (v < lo || v > hi) {
    char m[256];
    if (lo > hi)
      snprintf(m, sizeof m, "%s: %s index %lld out of range (acceptable index: none)", fn, what, v);
    else
      snprintf(m, sizeof m, "%s: %s index %lld out of range (acceptable index: %lld-%lld)", fn,
               what, v, lo, hi);
    return fail(s, BEX_E_BOUND, m);
  }
  return BEX_OK;
}

// Clears a work-queue counter (a kernel rather than a memset, so that launch sequences can be
// captured into graphs).
__global__ void zero_queue_kernel(unsigned long long* c) { *c = 0ull; }

using KernelFn = void (*)(BexKernelArgs, long long);
// Persistent kernel: (args, order, systems, work items, queue counter).
using QueueFn = void (*)(BexKernelArgs, const int*, long long, long long, unsigned long long*);

struct KernelChoice {
  KernelFn fn;      // one workgroup per kSystemsPerBlock systems
  QueueFn queue;    // persistent waves over a device work queue (same results)
  size_t lds;
  int systems_per_block;
  int block;
  int items_per_block;
  int systems_per_item;
};

// Dense output compiles into separate kernels (none, step points, Hermite), so launches without
// it (NDO = 0) carry none of its registers and each mode only its own.
template <class L, int ALG>
KernelFn kernel_of(int dense) {
  switch (dense) {
    case bexk::kDenseSteps: return (KernelFn)bexk::solve_kernel<L, ALG, bexk::kDenseSteps>;
    case bexk::kDenseHermite: return (KernelFn)bexk::solve_kernel<L, ALG, bexk::kDenseHermite>;
    default: return (KernelFn)bexk::solve_kernel<L, ALG, bexk::kDenseNone>;
  }
}
template <class L, int ALG>
QueueFn queue_kernel_of(int dense) {
  switch (dense) {
    case bexk::kDenseSteps: return (QueueFn)bexk::queue_kernel<L, ALG, bexk::kDenseSteps>;
    case bexk::kDenseHermite: return (QueueFn)bexk::queue_kernel<L, ALG, bexk::kDenseHermite>;
    default: return (QueueFn)bexk::queue_kernel<L, ALG, bexk::kDenseNone>;
  }
}
template <class L>
KernelChoice choice(int alg, int dense) {
  const bool rk4 = alg == BEX_RK4;
  const KernelFn fn = rk4 ? kernel_of<L, bexk::kRK4>(dense) : kernel_of<L, bexk::kRKCK45>(dense);
  const QueueFn q = rk4 ? queue_kernel_of<L, bexk::kRK4>(dense) : queue_kernel_of<L, bexk::kRKCK45>(dense);
  return {fn, q, L::lds_bytes(), L::kSystemsPerBlock, L::kBlock, L::kItemsPerBlock,
          L::kSystemsPerItem};
}

// Dense-output kernel variant of a solver's current settings.
int dense_mode(const bex_solver* s);

// layout: BEX_LAYOUT_AUTO picks the role-pair layouts (32 units: one system per wave; 64 and
// 128 units: two and eight waves per system, the matrix in registers) and the lane-per-unit
// layouts otherwise; BEX_LAYOUT_GROUP forces lane-per-unit (the matrix in LDS);
// BEX_LAYOUT_SYSTEM_PAIR takes two 32-unit systems per wave with both matrix rows in registers
// (plain start only). mode: bexk::WarmStartMode; the learned guess and sample collection exist
// for 32 units only.
bool kernel_for(int ups, int alg, int layout, int mode, int dense, KernelChoice* k) {
  if (layout == BEX_LAYOUT_SYSTEM_PAIR) {
    if (ups != 32 || mode != bexk::kPlain) return false;
    *k = choice<bexk::PairLayout32>(alg, dense);
    return true;
  }
  if ((ups == 64 || ups == 128) && layout != BEX_LAYOUT_GROUP) {
    if (mode != bexk::kPlain) return false;
    if (ups == 64)
      *k = choice<bexk::RolePairLayout64<false>>(alg, dense);
    else
      *k = choice<bexk::ChainLayout128<false>>(alg, dense);
    return true;
  }
  if (ups == 32 && layout != BEX_LAYOUT_GROUP) {
    switch (mode) {
      case bexk::kPlain: *k = choice<bexk::RolePairLayout32<bexk::kPlain>>(alg, dense); return true;
      case bexk::kSampling: *k = choice<bexk::RolePairLayout32<bexk::kSampling>>(alg, dense); return true;
      case bexk::kLearned: *k = choice<bexk::RolePairLayout32<bexk::kLearned>>(alg, dense); return true;
      default: return false;
    }
  }
  if (layout == BEX_LAYOUT_ROLE_PAIR || mode != bexk::kPlain) return false;
  switch (ups) {
    case 8: *k = choice<bexk::GroupLayout<8>>(alg, dense); return true;
    case 16: *k = choice<bexk::GroupLayout<16>>(alg, dense); return true;
    case 32: *k = choice<bexk::GroupLayout<32>>(alg, dense); return true;
    case 64: *k = choice<bexk::GroupLayout<64>>(alg, dense); return true;
    case 96: *k = choice<bexk::WideLayout<96, false, 2>>(alg, dense); return true;
    case 128: *k = choice<bexk::WideLayout<128, false, 2>>(alg, dense); return true;
    case 160: *k = choice<bexk::WideLayout<160, false, 2>>(alg, dense); return true;
    case 192: *k = choice<bexk::WideLayout<192, false, 2>>(alg, dense); return true;
    case 224: *k = choice<bexk::WideLayout<224, false, 2>>(alg, dense); return true;
    case 256: *k = choice<bexk::WideLayout<256, false, 2>>(alg, dense); return true;
    default: return false;
  }
}

// Batch inference through the solver's own MFMA path: each wave evaluates 32 rows, row r on
// lanes r % 32 and 32 + r % 32 exactly as a role-pair wave evaluates its 32 units.
__global__ __launch_bounds__(64) void mlp_mfma_kernel(const float* __restrict__ rows, int stride,
                                                      long long n, const BexMlpDevice* net,
                                                      float* __restrict__ out) {
  __shared__ BexMlpDevice snet;
  const uint4* from = reinterpret_cast<const uint4*>(net);
  uint4* to = reinterpret_cast<uint4*>(&snet);
  for (int k = threadIdx.x; k < (int)(sizeof(BexMlpDevice) / sizeof(uint4)); k += 64) to[k] = from[k];
  __syncthreads();
  const int lane = threadIdx.x;
  const long long r = (long long)blockIdx.x * 32 + (lane & 31);
  double wf[BEX_MLP_IN];
#pragma unroll
  for (int k = 0; k < BEX_MLP_IN; k++) wf[k] = (r < n) ? (double)rows[r * stride + k] : 0.0;
  const float y = bexk::mlp_factor(&snet, wf, lane);
  if (lane < 32 && r < n) out[r] = y;
}

int dense_mode(const bex_solver* s) {
  if (s->d.NDO <= 0) return bexk::kDenseNone;
  return (s->DenseInterpolation && s->d_DenseAux) ? bexk::kDenseHermite : bexk::kDenseSteps;
}

// Kernel arguments of one launch over the solver's buffers (also used by pools).
void fill_args(const bex_solver* s, BexKernelArgs* ap) {
  BexKernelArgs& a = *ap;
  const int ND = s->d.UD;
  a.NS = s->d.NS;
  a.ActiveSystems = s->ActiveSystems;
  a.NDO = s->d.NDO;
  a.first_system = 0;
  a.TLT = s->TLT;
  a.TimeDomain = (double*)s->TimeDomain.d;
  a.ActualState = (double*)s->ActualState.d;
  a.ActualTime = (double*)s->ActualTime.d;
  a.UnitParameters = (double*)s->UnitParameters.d;
  a.SystemParameters = (double*)s->SystemParameters.d;
  a.GlobalParameters = (double*)s->GlobalParameters.d;
  a.IntegerGlobalParameters = (int*)s->IntegerGlobalParameters.d;
  a.UnitAccessories = (double*)s->UnitAccessories.d;
  a.IntegerUnitAccessories = (int*)s->IntegerUnitAccessories.d;
  a.SystemAccessories = (double*)s->SystemAccessories.d;
  a.IntegerSystemAccessories = (int*)s->IntegerSystemAccessories.d;
  a.CouplingMatrix = s->PerSystemCoupling ? s->d_CouplingSys : s->d_CouplingFull;
  a.CouplingMatrixT = s->PerSystemCoupling ? s->d_CouplingSysT : s->d_CouplingFullT;
  a.DenseOutputIndex = (int*)s->DenseOutputIndex.d;
  a.DenseOutputTimeInstances = (double*)s->DenseOutputTimeInstances.d;
  a.DenseOutputStates = (double*)s->DenseOutputStates.d;
  a.RelativeTolerance = s->d_tol;
  a.AbsoluteTolerance = s->d_tol + ND;
  a.EventTolerance = s->d_tol + 2 * ND;
  a.EventDirection = s->d_dir;
  a.Stats = s->d_stats;
  a.Timeline = s->d_timeline;
  a.StopFlag = s->d_stop;
  a.MaxStepAttempts = s->MaxStepAttempts;
  a.InitialTimeStep = s->InitialTimeStep;
  a.MaximumTimeStep = s->MaximumTimeStep;
  a.MinimumTimeStep = s->MinimumTimeStep;
  a.TimeStepGrowLimit = s->TimeStepGrowLimit;
  a.TimeStepShrinkLimit = s->TimeStepShrinkLimit;
  a.DenseOutputMinimumTimeStep = s->DenseOutputMinimumTimeStep;
  a.DenseOutputSaveFrequency = s->DenseOutputSaveFrequency;
  a.BiCGRhoTolerance = s->BiCGRhoTolerance;
  a.BiCGResidualTolerance = s->BiCGResidualTolerance;
  a.BiCGMaxIterations = s->BiCGMaxIterations;
  a.PriorityStepThreshold = s->PriorityStepThreshold > 0 ? s->PriorityStepThreshold : 0x3fffffff;
  a.WarmStart = s->d_net;
  a.WarmStartSystems = s->d_warm_systems;
  a.Samples = s->d_samples;
  a.SampleCount = s->d_sample_count;
  a.SampleCapacity = s->SampleCapacity;
  a.SampleEvery = s->SampleEvery > 0 ? s->SampleEvery : 1;
  a.DenseInterpolation = (s->DenseInterpolation && s->d.NDO > 0 && s->d_DenseAux) ? 1 : 0;
  a.DenseAux = s->d_DenseAux;
  a.ref_spb = s->d.SPB;
  a.ref_block_stride = s->LogicalThreadsPerBlock + s->ThreadPaddingPerBlock;
}

// Logical thread index of (sid, uid) in the reference layout (Interface.cuh:875-879).
long long unit_index(const bex_solver* s, int sid, int uid) {
  int block = sid / s->d.SPB, lsid = sid % s->d.SPB;
  return (long long)block * (s->LogicalThreadsPerBlock + s->ThreadPaddingPerBlock) +
         (long long)lsid * s->d.UPS + uid;
}

Buf* buffer_of(bex_solver* s, int var) {
  switch (var) {
    case BEX_TimeDomain: return &s->TimeDomain;
    case BEX_ActualTime: return &s->ActualTime;
    case BEX_ActualState: return &s->ActualState;
    case BEX_UnitParameters: return &s->UnitParameters;
    case BEX_SystemParameters: return &s->SystemParameters;
    case BEX_GlobalParameters: return &s->GlobalParameters;
    case BEX_IntegerGlobalParameters: return &s->IntegerGlobalParameters;
    case BEX_UnitAccessories: return &s->UnitAccessories;
    case BEX_IntegerUnitAccessories: return &s->IntegerUnitAccessories;
    case BEX_SystemAccessories: return &s->SystemAccessories;
    case BEX_IntegerSystemAccessories: return &s->IntegerSystemAccessories;
    case BEX_CouplingMatrix: return &s->CouplingMatrix;
    case BEX_CouplingStrength: return &s->CouplingStrength;
    case BEX_CouplingIndex: return &s->CouplingIndex;
    case BEX_DenseIndex: return &s->DenseOutputIndex;
    case BEX_DenseTime: return &s->DenseOutputTimeInstances;
    case BEX_DenseState: return &s->DenseOutputStates;
    default: return nullptr;
  }
}

void put(Buf& b, long long i, double v) {
  if (b.esz == 8)
    static_cast<double*>(b.h)[i] = v;
  else
    static_cast<int*>(b.h)[i] = (int)v;
}
double get(const Buf& b, long long i) {
  return b.esz == 8 ? static_cast<const double*>(b.h)[i] : (double)static_cast<const int*>(b.h)[i];
}

// Compact index of coupling element (row, col) for the reference's four storage schemes
// (Interface.cuh:1001-1070); returns -1 outside the stored band.
long long compact_index(const bex_solver* s, int c, int row, int col) {
  const int N = s->d.UPS, CBW = s->d.CBW, CCI = s->d.CCI;
  if (CCI == 0 && CBW == 0) return row + (long long)col * N + (long long)c * N * N;
  int mod = col - row + CBW;
  if (mod >= N) mod -= N;
  if (mod < 0) mod += N;
  int lim = (CBW > 0) ? 2 * CBW + 1 : N;
  if (mod < 0 || mod >= lim) return -1;
  if (CCI == 0) return row + (long long)mod * N + (long long)c * N * lim;
  return mod + (long long)c * lim;
}

int copy_h2d(bex_solver* s, Buf& b) {
  if (b.n > 0) BEX_HIP(s, hipMemcpyAsync(b.d, b.h, (size_t)b.n * b.esz, hipMemcpyHostToDevice, s->stream));
  return BEX_OK;
}
int copy_d2h(bex_solver* s, Buf& b) {
  if (b.n > 0) BEX_HIP(s, hipMemcpyAsync(b.h, b.d, (size_t)b.n * b.esz, hipMemcpyDeviceToHost, s->stream));
  return BEX_OK;
}

// The kernel reads a full N x N matrix per coupling; expand the compact storage into it.
int upload_coupling(bex_solver* s) {
  const int N = s->d.UPS, NC = s->d.NC;
  double* full = nullptr;
  BEX_HIP(s, hipHostMalloc((void**)&full, sizeof(double) * (size_t)NC * N * N, hipHostMallocDefault));
  for (int c = 0; c < NC; c++)
    for (int col = 0; col < N; col++)
      for (int row = 0; row < N; row++) {
        long long k = compact_index(s, c, row, col);
        full[row + (long long)col * N + (long long)c * N * N] =
            (k >= 0) ? static_cast<double*>(s->CouplingMatrix.h)[k] : 0.0;
      }
  hipError_t e = hipMemcpyAsync(s->d_CouplingFull, full, sizeof(double) * (size_t)NC * N * N,
                                hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  // Transpose of coupling 0 (the one the implicit solve uses), for coalesced row loads of M^T.
  for (int col = 0; col < N && e == hipSuccess; col++)
    for (int row = 0; row < N; row++) {
      long long k = compact_index(s, 0, col, row);
      full[row + (long long)col * N] = (k >= 0) ? static_cast<double*>(s->CouplingMatrix.h)[k] : 0.0;
    }
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->d_CouplingFullT, full, sizeof(double) * (size_t)N * N,
                       hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  hipHostFree(full);
  if (e != hipSuccess) return fail(s, BEX_E_HIP, std::string("coupling upload: ") + hipGetErrorString(e));
  return BEX_OK;
}

int upload_tolerances(bex_solver* s) {
  const int n = 2 * s->d.UD + (s->d.NE > 0 ? s->d.NE : 1);
  BEX_HIP(s, hipMemcpyAsync(s->d_tol, s->h_tol, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
  BEX_HIP(s, hipMemcpyAsync(s->d_dir, s->h_dir, sizeof(int) * (s->d.NE > 0 ? s->d.NE : 1),
                            hipMemcpyHostToDevice, s->stream));
  // The staging arrays are reused by later calls: finish the copy before returning.
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  return BEX_OK;
}

}  // namespace

extern "C" {

int bex_compiled_dims(bex_problem_desc* out) {
  if (!out) return BEX_E_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  using D = BexSystemDims;
  out->UD = D::UD;
  out->NUP = D::NUP;
  out->NSP = D::NSP;
  out->NGP = D::NGP;
  out->NiGP = D::NiGP;
  out->NUA = D::NUA;
  out->NiUA = D::NiUA;
  out->NSA = D::NSA;
  out->NiSA = D::NiSA;
  out->NE = D::NE;
  out->NC = D::NC;
  return BEX_OK;
}

const char* bex_last_error(const bex_solver* s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

int bex_device_count(int* count) {
  if (!count) return fail(nullptr, BEX_E_ARGUMENT, "bex_device_count: null argument");
  *count = 0;
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice) return BEX_OK;
  if (e != hipSuccess) return fail(nullptr, BEX_E_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  *count = n;
  return BEX_OK;
}

int bex_get_device_properties(int device, bex_device_properties* out) {
  if (!out) return fail(nullptr, BEX_E_ARGUMENT, "bex_get_device_properties: null argument");
  int n = 0;
  int rc = bex_device_count(&n);
  if (rc != BEX_OK) return rc;
  if (device < 0 || device >= n)
    return fail(nullptr, BEX_E_BOUND, "device index " + std::to_string(device) + " out of range (" +
                                          std::to_string(n) + " devices)");
  hipDeviceProp_t p;
  const hipError_t e = hipGetDeviceProperties(&p, device);
  if (e != hipSuccess) return fail(nullptr, BEX_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof out->name, "%s", p.name);
  std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
  out->total_global_mem = (long long)p.totalGlobalMem;
  out->shared_mem_per_block = (long long)p.sharedMemPerBlock;
  out->total_const_mem = (long long)p.totalConstMem;
  out->regs_per_block = p.regsPerBlock;
  out->multiprocessor_count = p.multiProcessorCount;
  out->major = p.major;
  out->minor = p.minor;
  out->clock_rate_khz = p.clockRate;
  out->memory_clock_rate_khz = p.memoryClockRate;
  out->memory_bus_width = p.memoryBusWidth;
  out->warp_size = p.warpSize;
  out->max_threads_per_multiprocessor = p.maxThreadsPerMultiProcessor;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  for (int i = 0; i < 3; i++) {
    out->max_threads_dim[i] = p.maxThreadsDim[i];
    out->max_grid_size[i] = p.maxGridSize[i];
  }
  out->concurrent_copy = p.asyncEngineCount > 0 ? 1 : 0;
  out->concurrent_kernels = p.concurrentKernels;
  out->ecc_enabled = p.ECCEnabled;
  return BEX_OK;
}

int bex_select_device_by_closest_revision(int major, int minor, int* device) {
  if (!device) return fail(nullptr, BEX_E_ARGUMENT, "bex_select_device_by_closest_revision: null argument");
  int n = 0;
  int rc = bex_device_count(&n);
  if (rc != BEX_OK) return rc;
  if (n == 0) return fail(nullptr, BEX_E_HIP, "no HIP device");
  hipDeviceProp_t want;
  std::memset(&want, 0, sizeof want);
  want.major = major;
  want.minor = minor;
  const hipError_t e = hipChooseDevice(device, &want);
  if (e != hipSuccess) return fail(nullptr, BEX_E_HIP, std::string("hipChooseDevice: ") + hipGetErrorString(e));
  return BEX_OK;
}

int bex_create(const bex_problem_desc* desc, int device, bex_solver** out) {
  if (!desc || !out) return fail(nullptr, BEX_E_ARGUMENT, "bex_create: null argument");
  *out = nullptr;
  const bex_problem_desc& d = *desc;
  using D = BexSystemDims;
  if (d.UD != D::UD || d.NUP != D::NUP || d.NSP != D::NSP || d.NGP != D::NGP ||
      d.NiGP != D::NiGP || d.NUA != D::NUA || d.NiUA != D::NiUA || d.NSA != D::NSA ||
      d.NiSA != D::NiSA || d.NE != D::NE) {
    char m[512];
    snprintf(m, sizeof m,
             "bex_create: dimensions (UD %d NUP %d NSP %d NGP %d NiGP %d NUA %d NiUA %d NSA %d "
             "NiSA %d NE %d) differ from the compiled system definition (UD %d NUP %d NSP %d "
             "NGP %d NiGP %d NUA %d NiUA %d NSA %d NiSA %d NE %d)",
             d.UD, d.NUP, d.NSP, d.NGP, d.NiGP, d.NUA, d.NiUA, d.NSA, d.NiSA, d.NE, D::UD,
             D::NUP, D::NSP, D::NGP, D::NiGP, D::NUA, D::NiUA, D::NSA, D::NiSA, D::NE);
    return fail(nullptr, BEX_E_MISMATCH, m);
  }
  if (d.NS <= 0 || d.TPB <= 0 || d.SPB <= 0 || d.NC <= 0 || d.NDO < 0 || d.CBW < 0 ||
      (d.CCI != 0 && d.CCI != 1) || (d.Algorithm != BEX_RK4 && d.Algorithm != BEX_RKCK45))
    return fail(nullptr, BEX_E_ARGUMENT, "bex_create: invalid problem description");
  KernelChoice kc;
  if (!kernel_for(d.UPS, d.Algorithm, BEX_LAYOUT_AUTO, bexk::kPlain, bexk::kDenseNone, &kc))
    return fail(nullptr, BEX_E_UNSUPPORTED,
                "bex_create: units per system must be 8, 16 or a multiple of 32 up to 256 in "
                "this build");

  bex_solver* s = new bex_solver();
  s->d = d;
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&s->event);
  if (e == hipSuccess) e = hipEventCreate(&s->timer0);
  if (e == hipSuccess) e = hipEventCreate(&s->timer1);
  if (e != hipSuccess) {
    g_create_error = std::string("bex_create: ") + hipGetErrorString(e);
    delete s;
    return BEX_E_HIP;
  }

  // Thread configuration (Interface.cuh:386-391).
  s->LogicalThreadsPerBlock = d.SPB * d.UPS;
  int launches = s->LogicalThreadsPerBlock / d.TPB + (s->LogicalThreadsPerBlock % d.TPB ? 1 : 0);
  s->ThreadPaddingPerBlock = launches * d.TPB - s->LogicalThreadsPerBlock;
  s->GridSize = d.NS / d.SPB + (d.NS % d.SPB ? 1 : 0);
  s->TLT = (long long)(s->LogicalThreadsPerBlock + s->ThreadPaddingPerBlock) * s->GridSize;
  s->ActiveSystems = d.NS;

  // Sizes (Interface.cuh:408-438).
  const int N = d.UPS;
  if (d.CCI == 0 && d.CBW == 0) s->CompactCouplingSize = (long long)N * N;
  else if (d.CCI == 0) s->CompactCouplingSize = (long long)N * (2 * d.CBW + 1);
  else if (d.CBW > 0) s->CompactCouplingSize = 2 * d.CBW + 1;
  else s->CompactCouplingSize = N;
  struct Spec { Buf* b; long long n; int esz; };
  Spec specs[] = {
      {&s->TimeDomain, 2LL * d.NS, 8},
      {&s->ActualState, s->TLT * d.UD, 8},
      {&s->ActualTime, d.NS, 8},
      {&s->UnitParameters, s->TLT * d.NUP, 8},
      {&s->SystemParameters, (long long)d.NS * d.NSP, 8},
      {&s->GlobalParameters, d.NGP, 8},
      {&s->IntegerGlobalParameters, d.NiGP, 4},
      {&s->UnitAccessories, s->TLT * d.NUA, 8},
      {&s->IntegerUnitAccessories, s->TLT * d.NiUA, 4},
      {&s->SystemAccessories, (long long)d.NS * d.NSA, 8},
      {&s->IntegerSystemAccessories, (long long)d.NS * d.NiSA, 4},
      {&s->CouplingMatrix, (long long)d.NC * s->CompactCouplingSize, 8},
      {&s->CouplingStrength, (long long)d.NC * d.NS, 8},
      {&s->CouplingIndex, d.NC, 4},
      {&s->DenseOutputIndex, d.NS, 4},
      {&s->DenseOutputTimeInstances, (long long)d.NS * d.NDO, 8},
      {&s->DenseOutputStates, s->TLT * d.UD * d.NDO, 8},
  };
  size_t required = 0;
  for (auto& sp : specs) {
    sp.b->n = sp.n;
    sp.b->esz = sp.esz;
    required += sp.b->bytes();
  }
  required += sizeof(double) * (size_t)d.NC * N * N + sizeof(long long) * (size_t)d.NS * BEX_NSTAT;
  size_t freeMem = 0, totalMem = 0;
  if (hipMemGetInfo(&freeMem, &totalMem) == hipSuccess && required >= freeMem) {
    g_create_error = "bex_create: the required amount of global memory is larger than the free";
    bex_destroy(s);
    return BEX_E_MEMORY;
  }
  bool ok = true;
  for (auto& sp : specs) {
    if (hipHostMalloc(&sp.b->h, sp.b->bytes(), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&sp.b->d, sp.b->bytes()) != hipSuccess) {
      ok = false;
      break;
    }
    std::memset(sp.b->h, 0, sp.b->bytes());
  }
  const int ntol = 2 * d.UD + (d.NE > 0 ? d.NE : 1);
  ok = ok && hipHostMalloc((void**)&s->h_CouplingOriginal, sizeof(double) * (size_t)d.NC * N * N,
                           hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_CouplingFull, sizeof(double) * (size_t)d.NC * N * N) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_CouplingFullT, sizeof(double) * (size_t)N * N) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&s->h_tol, sizeof(double) * ntol, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&s->h_dir, sizeof(int) * (d.NE > 0 ? d.NE : 1), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_tol, sizeof(double) * ntol) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_dir, sizeof(int) * (d.NE > 0 ? d.NE : 1)) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_stats, sizeof(long long) * (size_t)d.NS * BEX_NSTAT) == hipSuccess;
  ok = ok && hipMalloc((void**)&s->d_queue, sizeof(unsigned long long)) == hipSuccess;
  ok = ok && hipDeviceGetAttribute(&s->cu_count, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&s->h_stop, sizeof(int),
                           hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess;
  ok = ok && hipHostGetDevicePointer((void**)&s->d_stop, s->h_stop, 0) == hipSuccess;
  if (ok) *s->h_stop = 0;
  if (!ok) {
    g_create_error = "bex_create: memory allocation failed";
    bex_destroy(s);
    return BEX_E_MEMORY;
  }
  std::memset(s->h_CouplingOriginal, 0, sizeof(double) * (size_t)d.NC * N * N);
  // Default options (Interface.cuh:613-635).
  for (int i = 0; i < d.UD; i++) {
    s->h_tol[i] = 1e-8;
    s->h_tol[d.UD + i] = 1e-8;
  }
  for (int i = 0; i < (d.NE > 0 ? d.NE : 1); i++) {
    s->h_tol[2 * d.UD + i] = 1e-6;
    s->h_dir[i] = 0;
  }
  int rc = upload_tolerances(s);
  if (rc == BEX_OK) rc = upload_coupling(s);
  if (rc == BEX_OK && hipMemsetAsync(s->d_stats, 0, sizeof(long long) * (size_t)d.NS * BEX_NSTAT,
                                     s->stream) != hipSuccess)
    rc = BEX_E_HIP;
  if (rc == BEX_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = BEX_E_HIP;
  if (rc != BEX_OK) {
    g_create_error = s->err;
    bex_destroy(s);
    return rc;
  }
  *out = s;
  return BEX_OK;
}

int bex_destroy(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  hipSetDevice(s->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  Buf* bufs[] = {&s->TimeDomain, &s->ActualState, &s->ActualTime, &s->UnitParameters,
                 &s->SystemParameters, &s->GlobalParameters, &s->IntegerGlobalParameters,
                 &s->UnitAccessories, &s->IntegerUnitAccessories, &s->SystemAccessories,
                 &s->IntegerSystemAccessories, &s->CouplingMatrix, &s->CouplingStrength,
                 &s->CouplingIndex, &s->DenseOutputIndex, &s->DenseOutputTimeInstances,
                 &s->DenseOutputStates};
  for (Buf* b : bufs) {
    if (b->h) hipHostFree(b->h);
    if (b->d) hipFree(b->d);
  }
  if (s->h_CouplingOriginal) hipHostFree(s->h_CouplingOriginal);
  if (s->d_CouplingFull) hipFree(s->d_CouplingFull);
  if (s->d_CouplingFullT) hipFree(s->d_CouplingFullT);
  if (s->d_CouplingSys) hipFree(s->d_CouplingSys);
  if (s->d_CouplingSysT) hipFree(s->d_CouplingSysT);
  if (s->h_tol) hipHostFree(s->h_tol);
  if (s->h_dir) hipHostFree(s->h_dir);
  if (s->d_tol) hipFree(s->d_tol);
  if (s->d_dir) hipFree(s->d_dir);
  if (s->d_stats) hipFree(s->d_stats);
  if (s->d_queue) hipFree(s->d_queue);
  if (s->d_order) hipFree(s->d_order);
  if (s->d_warm_systems) hipFree(s->d_warm_systems);
  if (s->h_timeline) hipHostFree(s->h_timeline);
  if (s->h_stop) hipHostFree(s->h_stop);
  if (s->d_DenseAux) hipFree(s->d_DenseAux);
  if (s->d_net) hipFree(s->d_net);
  if (s->d_samples) hipFree(s->d_samples);
  if (s->d_sample_count) hipFree(s->d_sample_count);
  if (s->d_ckpt) hipFree(s->d_ckpt);
  if (s->event) hipEventDestroy(s->event);
  if (s->timer0) hipEventDestroy(s->timer0);
  if (s->timer1) hipEventDestroy(s->timer1);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return BEX_OK;
}

int bex_solver_option(bex_solver* s, int option, double v) {
  if (!s) return BEX_E_ARGUMENT;
  s->config_epoch++;
  switch (option) {
    case BEX_InitialTimeStep: s->InitialTimeStep = v; break;
    case BEX_ActiveSystems: s->ActiveSystems = (int)v; break;
    case BEX_MaximumTimeStep: s->MaximumTimeStep = v; break;
    case BEX_MinimumTimeStep: s->MinimumTimeStep = v; break;
    case BEX_TimeStepGrowLimit: s->TimeStepGrowLimit = v; break;
    case BEX_TimeStepShrinkLimit: s->TimeStepShrinkLimit = v; break;
    case BEX_DenseOutputMinimumTimeStep: s->DenseOutputMinimumTimeStep = v; break;
    case BEX_DenseOutputSaveFrequency: s->DenseOutputSaveFrequency = (int)v; break;
    case BEX_DenseOutputInterpolation: {
      if (v != 0.0 && !s->d_DenseAux && s->d.NDO > 0) {
        BEX_HIP(s, hipSetDevice(s->device));
        BEX_HIP(s, hipMalloc((void**)&s->d_DenseAux, sizeof(double) * 2 * s->d.UD * (size_t)s->TLT));
      }
      s->DenseInterpolation = (v != 0.0) ? 1 : 0;
      break;
    }
    case BEX_SharedGlobalVariables: s->SharedGlobalVariables = (v != 0.0); break;
    case BEX_SharedCouplingMatrices: s->SharedCouplingMatrices = (v != 0.0); break;
    case BEX_BiCGRhoTolerance: s->BiCGRhoTolerance = v; break;
    case BEX_BiCGResidualTolerance: s->BiCGResidualTolerance = v; break;
    case BEX_BiCGMaxIterations: s->BiCGMaxIterations = (int)v; break;
    case BEX_PriorityStepThreshold: s->PriorityStepThreshold = (int)v; break;
    case BEX_MaximumStepAttempts:
      if (!(v >= 0.0) || v > 9.0e18)
        return fail(s, BEX_E_ARGUMENT, "SolverOption: MaximumStepAttempts must be >= 0");
      s->MaxStepAttempts = (long long)v;
      break;
    case BEX_PersistentLaunch: s->PersistentLaunch = (v != 0.0) ? 1 : 0; break;
    case BEX_PersistentWorkgroups:
      if (!(v >= 0.0) || v > 2.0e9)
        return fail(s, BEX_E_ARGUMENT, "SolverOption: PersistentWorkgroups must be >= 0");
      s->PersistentWorkgroups = (int)v;
      break;
    case BEX_KernelLayout: {
      KernelChoice kc;
      if (!kernel_for(s->d.UPS, s->d.Algorithm, (int)v, bexk::kPlain, bexk::kDenseNone, &kc))
        return fail(s, BEX_E_UNSUPPORTED, "SolverOption: no kernel layout for this unit count");
      s->KernelLayout = (int)v;
      break;
    }
    case BEX_WarmStart: {
      KernelChoice kc;
      if (v != 0.0 && !kernel_for(s->d.UPS, s->d.Algorithm, s->KernelLayout, bexk::kLearned, bexk::kDenseNone, &kc))
        return fail(s, BEX_E_UNSUPPORTED,
                    "SolverOption: the learned warm start needs 32 units and the role-pair layout");
      s->WarmStart = (v != 0.0) ? 1 : 0;
      break;
    }
    default:
      return fail(s, BEX_E_OPTION,
                  "SolverOption: this option needs 2 input arguments or is not applicable");
  }
  return BEX_OK;
}

int bex_solver_option_indexed(bex_solver* s, int option, int index, double v) {
  if (!s) return BEX_E_ARGUMENT;
  s->config_epoch++;
  const int UD = s->d.UD, NE = s->d.NE;
  int rc;
  switch (option) {
    case BEX_RelativeTolerance:
      if ((rc = bound(s, "SolverOption", "RelativeTolerance", index, 0, UD - 1))) return rc;
      s->h_tol[index] = v;
      break;
    case BEX_AbsoluteTolerance:
      if ((rc = bound(s, "SolverOption", "AbsoluteTolerance", index, 0, UD - 1))) return rc;
      s->h_tol[UD + index] = v;
      break;
    case BEX_EventTolerance:
      if ((rc = bound(s, "SolverOption", "EventTolerance", index, 0, NE - 1))) return rc;
      s->h_tol[2 * UD + index] = v;
      break;
    case BEX_EventDirection:
      if ((rc = bound(s, "SolverOption", "EventDirection", index, 0, NE - 1))) return rc;
      s->h_dir[index] = (int)v;
      break;
    default:
      return fail(s, BEX_E_OPTION,
                  "SolverOption: this option needs 1 input argument or is not applicable");
  }
  return upload_tolerances(s);
}

int bex_set_host_unit(bex_solver* s, int sid, int uid, int var, int k, double v) {
  if (!s) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "SetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  if ((rc = bound(s, "SetHost", "UnitNumber", uid, 0, s->d.UPS - 1))) return rc;
  long long g = unit_index(s, sid, uid);
  int lim;
  Buf* b;
  switch (var) {
    case BEX_ActualState: lim = s->d.UD; b = &s->ActualState; break;
    case BEX_UnitParameters: lim = s->d.NUP; b = &s->UnitParameters; break;
    case BEX_UnitAccessories: lim = s->d.NUA; b = &s->UnitAccessories; break;
    case BEX_IntegerUnitAccessories: lim = s->d.NiUA; b = &s->IntegerUnitAccessories; break;
    default: return fail(s, BEX_E_VARIABLE, "SetHost: this option needs different argument configuration or not applicable");
  }
  if ((rc = bound(s, "SetHost", "SerialNumber", k, 0, lim - 1))) return rc;
  put(*b, g + (long long)k * s->TLT, v);
  return BEX_OK;
}

int bex_get_host_unit(bex_solver* s, int sid, int uid, int var, int k, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "GetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  if ((rc = bound(s, "GetHost", "UnitNumber", uid, 0, s->d.UPS - 1))) return rc;
  long long g = unit_index(s, sid, uid);
  int lim;
  Buf* b;
  switch (var) {
    case BEX_ActualState: lim = s->d.UD; b = &s->ActualState; break;
    case BEX_UnitParameters: lim = s->d.NUP; b = &s->UnitParameters; break;
    case BEX_UnitAccessories: lim = s->d.NUA; b = &s->UnitAccessories; break;
    case BEX_IntegerUnitAccessories: lim = s->d.NiUA; b = &s->IntegerUnitAccessories; break;
    default: return fail(s, BEX_E_VARIABLE, "GetHost: this option needs different argument configuration or not applicable");
  }
  if ((rc = bound(s, "GetHost", "SerialNumber", k, 0, lim - 1))) return rc;
  *v = get(*b, g + (long long)k * s->TLT);
  return BEX_OK;
}

static int system_slot(bex_solver* s, const char* fn, int var, int k, Buf** b) {
  int lim;
  switch (var) {
    case BEX_TimeDomain: lim = 2; *b = &s->TimeDomain; break;
    case BEX_ActualTime: lim = 1; *b = &s->ActualTime; break;
    case BEX_SystemParameters: lim = s->d.NSP; *b = &s->SystemParameters; break;
    case BEX_SystemAccessories: lim = s->d.NSA; *b = &s->SystemAccessories; break;
    case BEX_IntegerSystemAccessories: lim = s->d.NiSA; *b = &s->IntegerSystemAccessories; break;
    case BEX_DenseTime: lim = s->d.NDO; *b = &s->DenseOutputTimeInstances; break;
    case BEX_CouplingStrength: lim = s->d.NC; *b = &s->CouplingStrength; break;
    default: return fail(s, BEX_E_VARIABLE, std::string(fn) + ": this option needs different argument configuration or not applicable");
  }
  return bound(s, fn, "SerialNumber", k, 0, lim - 1);
}

int bex_set_host_system(bex_solver* s, int sid, int var, int k, double v) {
  if (!s) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "SetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  Buf* b;
  if ((rc = system_slot(s, "SetHost", var, k, &b))) return rc;
  put(*b, sid + (long long)k * s->d.NS, v);
  return BEX_OK;
}

int bex_get_host_system(bex_solver* s, int sid, int var, int k, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "GetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  Buf* b;
  if ((rc = system_slot(s, "GetHost", var, k, &b))) return rc;
  *v = get(*b, sid + (long long)k * s->d.NS);
  return BEX_OK;
}

static int global_slot(bex_solver* s, const char* fn, int var, int k, Buf** b) {
  int lim;
  switch (var) {
    case BEX_GlobalParameters: lim = s->d.NGP; *b = &s->GlobalParameters; break;
    case BEX_IntegerGlobalParameters: lim = s->d.NiGP; *b = &s->IntegerGlobalParameters; break;
    case BEX_CouplingIndex: lim = s->d.NC; *b = &s->CouplingIndex; break;
    default: return fail(s, BEX_E_VARIABLE, std::string(fn) + ": this option needs different argument configuration or not applicable");
  }
  return bound(s, fn, "SerialNumber", k, 0, lim - 1);
}

int bex_set_host_global(bex_solver* s, int var, int k, double v) {
  if (!s) return BEX_E_ARGUMENT;
  Buf* b;
  int rc;
  if ((rc = global_slot(s, "SetHost", var, k, &b))) return rc;
  put(*b, k, v);
  return BEX_OK;
}

int bex_get_host_global(bex_solver* s, int var, int k, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  Buf* b;
  int rc;
  if ((rc = global_slot(s, "GetHost", var, k, &b))) return rc;
  *v = get(*b, k);
  return BEX_OK;
}

int bex_set_host_coupling(bex_solver* s, int c, int var, int row, int col, double v) {
  if (!s) return BEX_E_ARGUMENT;
  int rc;
  const int N = s->d.UPS;
  if ((rc = bound(s, "SetHost", "CouplingNumber", c, 0, s->d.NC - 1))) return rc;
  if ((rc = bound(s, "SetHost", "Coupling Matrix Row", row, 0, N - 1))) return rc;
  if ((rc = bound(s, "SetHost", "Coupling Matrix Col", col, 0, N - 1))) return rc;
  if (var != BEX_CouplingMatrix)
    return fail(s, BEX_E_VARIABLE, "SetHost: this option needs different argument configuration or not applicable");
  long long k = compact_index(s, c, row, col);
  if (k < 0) {
    int lim = (s->d.CBW > 0) ? 2 * s->d.CBW + 1 : N;
    int mod = ((col - row + s->d.CBW) % N + N) % N;
    return bound(s, "SetHost", "Coupling Matrix Shifted Diagonal", mod, 0, lim - 1);
  }
  s->h_CouplingOriginal[row + (long long)col * N + (long long)c * N * N] = v;
  static_cast<double*>(s->CouplingMatrix.h)[k] = v;
  return BEX_OK;
}

int bex_get_host_coupling(bex_solver* s, int c, int var, int row, int col, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  int rc;
  const int N = s->d.UPS;
  if ((rc = bound(s, "GetHost", "CouplingNumber", c, 0, s->d.NC - 1))) return rc;
  if ((rc = bound(s, "GetHost", "Coupling Matrix Row", row, 0, N - 1))) return rc;
  if ((rc = bound(s, "GetHost", "Coupling Matrix Col", col, 0, N - 1))) return rc;
  if (var != BEX_CouplingMatrix)
    return fail(s, BEX_E_VARIABLE, "GetHost: this option needs different argument configuration or not applicable");
  *v = s->h_CouplingOriginal[row + (long long)col * N + (long long)c * N * N];  // Interface.cuh:1469
  return BEX_OK;
}

int bex_set_host_dense_index(bex_solver* s, int sid, int var, double v) {
  if (!s) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "SetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  if (var != BEX_DenseIndex)
    return fail(s, BEX_E_VARIABLE, "SetHost: this option needs different argument configuration or not applicable");
  // The index counts stored points: 0 (empty) to NDO (full). Anything else would address
  // outside the dense-output buffers.
  if (!(v >= 0.0 && v <= (double)s->d.NDO)) {
    char m[192];
    snprintf(m, sizeof m, "SetHost: DenseIndex value %g out of range (acceptable values: 0-%d)", v,
             s->d.NDO);
    return fail(s, BEX_E_BOUND, m);
  }
  static_cast<int*>(s->DenseOutputIndex.h)[sid] = (int)v;
  return BEX_OK;
}

int bex_get_host_dense_index(bex_solver* s, int sid, int var, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  int rc;
  if ((rc = bound(s, "GetHost", "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  if (var != BEX_DenseIndex)
    return fail(s, BEX_E_VARIABLE, "GetHost: this option needs different argument configuration or not applicable");
  *v = static_cast<int*>(s->DenseOutputIndex.h)[sid];
  return BEX_OK;
}

static int dense_state_slot(bex_solver* s, const char* fn, int sid, int uid, int var, int comp,
                            int k, long long* idx) {
  int rc;
  if ((rc = bound(s, fn, "SystemNumber", sid, 0, s->d.NS - 1))) return rc;
  if ((rc = bound(s, fn, "UnitNumber", uid, 0, s->d.UPS - 1))) return rc;
  if ((rc = bound(s, fn, "ComponentNumber in dense state", comp, 0, s->d.UD - 1))) return rc;
  if (var != BEX_DenseState)
    return fail(s, BEX_E_VARIABLE, std::string(fn) + ": this option needs different argument configuration or not applicable");
  if ((rc = bound(s, fn, "DenseState", k, 0, s->d.NDO - 1))) return rc;
  *idx = unit_index(s, sid, uid) + (long long)comp * s->TLT + (long long)k * s->TLT * s->d.UD;
  return BEX_OK;
}

int bex_set_host_dense_state(bex_solver* s, int sid, int uid, int var, int comp, int k, double v) {
  if (!s) return BEX_E_ARGUMENT;
  long long i;
  int rc;
  if ((rc = dense_state_slot(s, "SetHost", sid, uid, var, comp, k, &i))) return rc;
  static_cast<double*>(s->DenseOutputStates.h)[i] = v;
  return BEX_OK;
}

int bex_get_host_dense_state(bex_solver* s, int sid, int uid, int var, int comp, int k, double* v) {
  if (!s || !v) return BEX_E_ARGUMENT;
  long long i;
  int rc;
  if ((rc = dense_state_slot(s, "GetHost", sid, uid, var, comp, k, &i))) return rc;
  *v = static_cast<double*>(s->DenseOutputStates.h)[i];
  return BEX_OK;
}

int bex_host_view(bex_solver* s, int var, void** ptr, long long* count, int* elem_bytes) {
  if (!s || !ptr || !count || !elem_bytes) return BEX_E_ARGUMENT;
  Buf* b = buffer_of(s, var);
  if (!b) return fail(s, BEX_E_VARIABLE, "bex_host_view: no single buffer for this variable");
  *ptr = b->h;
  *count = b->n;
  *elem_bytes = b->esz;
  return BEX_OK;
}

// Host-side dense-output indices (also writable through host views) must lie in [0, NDO]
// before they reach the kernel, which stores point idx of a system at offset idx.
static int check_dense_index(bex_solver* s) {
  const int* idx = static_cast<const int*>(s->DenseOutputIndex.h);
  for (long long k = 0; k < s->DenseOutputIndex.n; k++)
    if (idx[k] < 0 || idx[k] > s->d.NDO) {
      char m[192];
      snprintf(m, sizeof m,
               "SynchroniseFromHostToDevice: DenseIndex of system %lld is %d (acceptable values: 0-%d)",
               k, idx[k], s->d.NDO);
      return fail(s, BEX_E_BOUND, m);
    }
  return BEX_OK;
}

int bex_sync_h2d(bex_solver* s, int var) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  int rc = BEX_OK;
  switch (var) {
    case BEX_CouplingMatrix: return upload_coupling(s);
    case BEX_DenseOutput:
      if ((rc = check_dense_index(s))) return rc;
      if ((rc = copy_h2d(s, s->DenseOutputIndex))) return rc;
      if ((rc = copy_h2d(s, s->DenseOutputTimeInstances))) return rc;
      return copy_h2d(s, s->DenseOutputStates);
    case BEX_All: {
      Buf* all[] = {&s->TimeDomain, &s->ActualState, &s->ActualTime, &s->UnitParameters,
                    &s->SystemParameters, &s->GlobalParameters, &s->IntegerGlobalParameters,
                    &s->UnitAccessories, &s->IntegerUnitAccessories, &s->SystemAccessories,
                    &s->IntegerSystemAccessories, &s->CouplingStrength, &s->CouplingIndex,
                    &s->DenseOutputIndex, &s->DenseOutputTimeInstances, &s->DenseOutputStates};
      if ((rc = check_dense_index(s))) return rc;
      for (Buf* b : all)
        if ((rc = copy_h2d(s, *b))) return rc;
      return upload_coupling(s);
    }
    default: {
      Buf* b = buffer_of(s, var);
      if (!b || var == BEX_DenseIndex || var == BEX_DenseTime || var == BEX_DenseState)
        return fail(s, BEX_E_VARIABLE, "SynchroniseFromHostToDevice: this option is not applicable");
      return copy_h2d(s, *b);
    }
  }
}

int bex_sync_d2h(bex_solver* s, int var) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  int rc = BEX_OK;
  switch (var) {
    case BEX_CouplingMatrix: return BEX_OK;  // the kernel never writes the coupling matrix
    case BEX_DenseOutput:
      if ((rc = copy_d2h(s, s->DenseOutputIndex))) return rc;
      if ((rc = copy_d2h(s, s->DenseOutputTimeInstances))) return rc;
      return copy_d2h(s, s->DenseOutputStates);
    case BEX_All: {
      Buf* all[] = {&s->TimeDomain, &s->ActualState, &s->ActualTime, &s->UnitParameters,
                    &s->SystemParameters, &s->GlobalParameters, &s->IntegerGlobalParameters,
                    &s->UnitAccessories, &s->IntegerUnitAccessories, &s->SystemAccessories,
                    &s->IntegerSystemAccessories, &s->CouplingStrength, &s->CouplingIndex,
                    &s->DenseOutputIndex, &s->DenseOutputTimeInstances, &s->DenseOutputStates};
      for (Buf* b : all)
        if ((rc = copy_d2h(s, *b))) return rc;
      return BEX_OK;
    }
    default: {
      Buf* b = buffer_of(s, var);
      if (!b || var == BEX_DenseIndex || var == BEX_DenseTime || var == BEX_DenseState)
        return fail(s, BEX_E_VARIABLE, "SynchroniseFromDeviceToHost: this option is not applicable");
      return copy_d2h(s, *b);
    }
  }
}

static int pool_class_solve(bex_solver* s);  // pool section below

// Systems a persistent launch hands out (with a hand-out order every system is listed; without
// one, systems 0 .. nsys-1) and the work items they make (systems_per_item systems each).
static void persistent_items(const bex_solver* s, const KernelChoice& kc, int nsys,
                             long long* count, long long* nitems) {
  *count = s->d_order ? (long long)s->d.NS : (long long)nsys;
  *nitems = (*count + kc.systems_per_item - 1) / kc.systems_per_item;
}

// Workgroups of a persistent launch of kernel kc over nitems work items: as many as fit on the
// device, at most PersistentWorkgroups when that is set, and no more than the items need.
static int persistent_blocks(bex_solver* s, const KernelChoice& kc, long long nitems,
                             long long* blocks) {
  int per_cu = 0;
  BEX_HIP(s, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &per_cu, reinterpret_cast<const void*>(kc.queue), kc.block, kc.lds));
  if (per_cu < 1) return fail(s, BEX_E_UNSUPPORTED, "Solve: the kernel does not fit on a CU");
  long long b = (long long)s->cu_count * per_cu;
  if (s->PersistentWorkgroups > 0 && b > s->PersistentWorkgroups) b = s->PersistentWorkgroups;
  const long long need = (nitems + kc.items_per_block - 1) / kc.items_per_block;
  *blocks = b < need ? b : need;
  return BEX_OK;
}

int bex_persistent_grid(bex_solver* s, int* workgroups, int* waves_per_workgroup) {
  if (!s || !workgroups || !waves_per_workgroup) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  KernelChoice kc;
  const int mode = s->WarmStart ? bexk::kLearned : (s->d_samples ? bexk::kSampling : bexk::kPlain);
  if (s->pool || !s->PersistentLaunch ||
      !kernel_for(s->d.UPS, s->d.Algorithm, s->KernelLayout, mode, dense_mode(s), &kc))
    return fail(s, BEX_E_UNSUPPORTED, "bex_persistent_grid: no persistent kernel for this solver");
  const int nsys = s->ActiveSystems < s->d.NS ? s->ActiveSystems : s->d.NS;
  long long count = 0, nitems = 0, blocks = 0;
  persistent_items(s, kc, nsys, &count, &nitems);
  if (nsys > 0) {
    const int rc = persistent_blocks(s, kc, nitems, &blocks);
    if (rc != BEX_OK) return rc;
  }
  *workgroups = (int)blocks;
  *waves_per_workgroup = kc.block / 64;
  return BEX_OK;
}

int bex_solve(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  if (s->pool) return pool_class_solve(s);  // per-system matrices: the pool's class kernel
  KernelChoice kc;
  const int mode = s->WarmStart ? bexk::kLearned : (s->d_samples ? bexk::kSampling : bexk::kPlain);
  if (mode == bexk::kLearned && !s->d_net)
    return fail(s, BEX_E_ARGUMENT, "Solve: warm start enabled but no network was set");
  if (!kernel_for(s->d.UPS, s->d.Algorithm, s->KernelLayout, mode, dense_mode(s), &kc))
    return fail(s, BEX_E_UNSUPPORTED, "Solve: no kernel for this unit count, layout and mode");
  BexKernelArgs a{};
  fill_args(s, &a);
  if (a.DenseInterpolation && !(s->DenseOutputMinimumTimeStep > 0.0))
    return fail(s, BEX_E_OPTION,
                "Solve: interpolated dense output needs DenseOutputMinimumTimeStep > 0");
  int nsys = s->ActiveSystems < s->d.NS ? s->ActiveSystems : s->d.NS;
  if (nsys <= 0) return BEX_OK;
  if (s->PersistentLaunch) {
    // As many workgroups as fit on the device at once, each taking work items from the queue
    // (the counter is cleared by a kernel, so the sequence can be captured into a graph).
    if (kc.lds > 65536)
      BEX_HIP(s, hipFuncSetAttribute(reinterpret_cast<const void*>(kc.queue),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kc.lds));
    long long count = 0, nitems = 0, blocks = 0;
    persistent_items(s, kc, nsys, &count, &nitems);
    const int rc = persistent_blocks(s, kc, nitems, &blocks);
    if (rc != BEX_OK) return rc;
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(1), 0, s->stream, s->d_queue);
    hipLaunchKernelGGL(kc.queue, dim3((unsigned)blocks), dim3(kc.block), kc.lds, s->stream, a,
                       (const int*)s->d_order, count, nitems, s->d_queue);
    BEX_HIP(s, hipGetLastError());
    return BEX_OK;
  }
  if (kc.lds > 65536)
    BEX_HIP(s, hipFuncSetAttribute(reinterpret_cast<const void*>(kc.fn),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kc.lds));
  dim3 grid((nsys + kc.systems_per_block - 1) / kc.systems_per_block), block(kc.block);
  hipLaunchKernelGGL(kc.fn, grid, block, kc.lds, s->stream, a, (long long)nsys);
  BEX_HIP(s, hipGetLastError());
  return BEX_OK;
}

int bex_set_warm_start_systems(bex_solver* s, const unsigned char* flags, int n) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  if (!flags || n == 0) {
    if (s->d_warm_systems) {
      BEX_HIP(s, hipStreamSynchronize(s->stream));
      BEX_HIP(s, hipFree(s->d_warm_systems));
      s->d_warm_systems = nullptr;
    }
    return BEX_OK;
  }
  if (n != s->d.NS)
    return fail(s, BEX_E_ARGUMENT, "bex_set_warm_start_systems: one flag per system is needed");
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  if (!s->d_warm_systems) BEX_HIP(s, hipMalloc((void**)&s->d_warm_systems, (size_t)n));
  BEX_HIP(s, hipMemcpyAsync(s->d_warm_systems, flags, (size_t)n, hipMemcpyHostToDevice, s->stream));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  return BEX_OK;
}

int bex_set_launch_order(bex_solver* s, const int* order, int n) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  if (!order || n == 0) {
    if (s->d_order) {
      BEX_HIP(s, hipStreamSynchronize(s->stream));
      BEX_HIP(s, hipFree(s->d_order));
      s->d_order = nullptr;
    }
    return BEX_OK;
  }
  if (n != s->d.NS)
    return fail(s, BEX_E_ARGUMENT, "bex_set_launch_order: the order must list every system once");
  std::vector<char> seen((size_t)n, 0);
  for (int k = 0; k < n; k++) {
    if (order[k] < 0 || order[k] >= n || seen[(size_t)order[k]])
      return fail(s, BEX_E_ARGUMENT, "bex_set_launch_order: the order must list every system once");
    seen[(size_t)order[k]] = 1;
  }
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  if (!s->d_order) BEX_HIP(s, hipMalloc((void**)&s->d_order, sizeof(int) * (size_t)n));
  // On the solver's stream, which the next Solve() follows; complete before `order` is
  // returned to the caller.
  BEX_HIP(s, hipMemcpyAsync(s->d_order, order, sizeof(int) * (size_t)n, hipMemcpyHostToDevice,
                            s->stream));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  return BEX_OK;
}

int bex_synchronise_device(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipDeviceSynchronize());
  return BEX_OK;
}

int bex_insert_synchronisation_point(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipEventRecord(s->event, s->stream));
  return BEX_OK;
}

int bex_synchronise_solver(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipEventSynchronize(s->event));
  return BEX_OK;
}

// Buffers one Solve() writes back (Solver.cuh:601-648); each starts at a 256-byte offset of the
// snapshot.
static int checkpoint_copy(bex_solver* s, bool save) {
  Buf* bufs[] = {&s->TimeDomain, &s->ActualState, &s->ActualTime, &s->UnitAccessories,
                 &s->IntegerUnitAccessories, &s->SystemAccessories,
                 &s->IntegerSystemAccessories, &s->DenseOutputIndex};
  auto span = [](const Buf* b) { return b->n > 0 ? ((b->bytes() + 255) / 256) * 256 : (size_t)0; };
  BEX_HIP(s, hipSetDevice(s->device));
  if (!s->d_ckpt) {
    if (!save) return fail(s, BEX_E_ARGUMENT, "bex_checkpoint_restore: no checkpoint was saved");
    size_t total = 0;
    for (const Buf* b : bufs) total += span(b);
    BEX_HIP(s, hipMalloc((void**)&s->d_ckpt, total > 0 ? total : 256));
  }
  size_t off = 0;
  for (Buf* b : bufs) {
    if (b->n > 0) {
      unsigned char* snap = s->d_ckpt + off;
      const size_t bytes = (size_t)b->n * b->esz;
      BEX_HIP(s, hipMemcpyAsync(save ? (void*)snap : b->d, save ? b->d : (const void*)snap, bytes,
                                hipMemcpyDeviceToDevice, s->stream));
    }
    off += span(b);
  }
  return BEX_OK;
}

int bex_checkpoint_save(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  return checkpoint_copy(s, true);
}

int bex_checkpoint_restore(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  return checkpoint_copy(s, false);
}

int bex_get_stats(bex_solver* s, long long* out) {
  if (!s || !out) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipMemcpyAsync(out, s->d_stats, sizeof(long long) * (size_t)s->d.NS * BEX_NSTAT,
                            hipMemcpyDeviceToHost, s->stream));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  return BEX_OK;
}

int bex_timer_start(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipEventRecord(s->timer0, s->stream));
  return BEX_OK;
}

int bex_timer_stop(bex_solver* s, double* elapsed_ms) {
  if (!s || !elapsed_ms) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipEventRecord(s->timer1, s->stream));
  BEX_HIP(s, hipEventSynchronize(s->timer1));
  float ms = 0.0f;
  BEX_HIP(s, hipEventElapsedTime(&ms, s->timer0, s->timer1));
  *elapsed_ms = ms;
  return BEX_OK;
}

int bex_enable_timeline(bex_solver* s, int on) {
  if (!s) return BEX_E_ARGUMENT;
  s->config_epoch++;
  BEX_HIP(s, hipSetDevice(s->device));
  // No launch may still write the timeline while it is cleared or freed (a pool's graph runs
  // on the pool's streams, hence the device-wide wait).
  BEX_HIP(s, hipDeviceSynchronize());
  const size_t bytes = sizeof(long long) * 2 * (size_t)s->d.NS;
  if (on) {
    if (!s->h_timeline) {
      BEX_HIP(s, hipHostMalloc((void**)&s->h_timeline, bytes,
                               hipHostMallocMapped | hipHostMallocCoherent));
      BEX_HIP(s, hipHostGetDevicePointer((void**)&s->d_timeline, s->h_timeline, 0));
    }
    std::memset(s->h_timeline, 0, bytes);  // (re)enabling clears it
  } else if (s->h_timeline) {
    BEX_HIP(s, hipHostFree(s->h_timeline));
    s->h_timeline = nullptr;
    s->d_timeline = nullptr;
  }
  return BEX_OK;
}

int bex_get_timeline(bex_solver* s, long long* out, double* ticks_per_second) {
  if (!s || !out || !ticks_per_second) return BEX_E_ARGUMENT;
  if (!s->h_timeline) return fail(s, BEX_E_ARGUMENT, "bex_get_timeline: timeline not enabled");
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipStreamSynchronize(s->stream));  // pool launches: the class stream waits for them
  std::memcpy(out, s->h_timeline, sizeof(long long) * 2 * (size_t)s->d.NS);
  int khz = 0;
  BEX_HIP(s, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device));
  *ticks_per_second = 1e3 * khz;
  return BEX_OK;
}

int bex_peek_timeline(bex_solver* s, long long* out) {
  if (!s || !out) return BEX_E_ARGUMENT;
  if (!s->h_timeline) return fail(s, BEX_E_ARGUMENT, "bex_peek_timeline: timeline not enabled");
  // Plain reads of the mapped host buffer: no stream, no wait for the running launch.
  const volatile long long* src = s->h_timeline;
  for (long long k = 0; k < 2LL * s->d.NS; k++) out[k] = src[k];
  return BEX_OK;
}

int bex_request_stop(bex_solver* s) {
  if (!s || !s->h_stop) return BEX_E_ARGUMENT;
  *reinterpret_cast<volatile int*>(s->h_stop) = 1;
  return BEX_OK;
}

int bex_clear_stop(bex_solver* s) {
  if (!s || !s->h_stop) return BEX_E_ARGUMENT;
  *reinterpret_cast<volatile int*>(s->h_stop) = 0;
  return BEX_OK;
}

int bex_reset_stats(bex_solver* s) {
  if (!s) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipMemsetAsync(s->d_stats, 0, sizeof(long long) * (size_t)s->d.NS * BEX_NSTAT, s->stream));
  return BEX_OK;
}

}  // extern "C"

// Scratch of bex_mlp_predict_mfma_noalloc: one network packed in the MFMA operand layout
// (sizeof(BexMlpDevice) = 9744 bytes), the staging buffer that bex_mlp_predict_mfma allocates
// and frees on every call. Each device holds its own copy until the process ends; every call on
// a device overwrites it, so calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[sizeof(BexMlpDevice)];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};  // device address of `scratch`, per device

// Device address of the current device's `scratch`, looked up once per device (with that
// device current) by hipGetSymbolAddress; host code never uses the symbol itself.
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return BEX_E_HIP;
  if (dev < 0 || dev >= kScratchDevices) return BEX_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BEX_E_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BEX_OK;
}

}  // namespace

extern "C" {

int bex_mlp_predict_mfma_noalloc(const float* rows, int stride, long long n, const float* params,
                                 float* out, void* stream) {
  if (!rows || !params || !out || n <= 0 || stride < BEX_MLP_IN) return BEX_E_ARGUMENT;
  const long long nb = (n + 31) / 32;
  if (nb > 0x7fffffff) return BEX_E_UNSUPPORTED;  // more rows than one launch covers
  void* d = nullptr;
  int rc = scratch_address(&d);
  if (rc != BEX_OK) return rc;
  BexMlpDevice host;
  bex_mlp_pack_device(params, &host);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(d, &host, sizeof(BexMlpDevice), hipMemcpyHostToDevice, st) != hipSuccess)
    return BEX_E_HIP;
  hipLaunchKernelGGL(mlp_mfma_kernel, dim3((unsigned)nb), dim3(64), 0, st, rows, stride, n,
                     static_cast<const BexMlpDevice*>(d), out);
  // The packed copy lives in this frame: finish before returning, like bex_mlp_predict_mfma.
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return BEX_E_HIP;
  return BEX_OK;
}

int bex_mlp_predict_mfma(const float* rows, int stride, long long n, const float* params,
                         float* out, void* stream) {
  if (!rows || !params || !out || n <= 0 || stride < BEX_MLP_IN) return BEX_E_ARGUMENT;
  const long long nb = (n + 31) / 32;
  if (nb > 0x7fffffff) return BEX_E_UNSUPPORTED;
  BexMlpDevice host;
  bex_mlp_pack_device(params, &host);
  BexMlpDevice* d = nullptr;
  if (hipMalloc((void**)&d, sizeof(BexMlpDevice)) != hipSuccess) return BEX_E_MEMORY;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int rc = BEX_OK;
  if (hipMemcpyAsync(d, &host, sizeof(BexMlpDevice), hipMemcpyHostToDevice, st) != hipSuccess)
    rc = BEX_E_HIP;
  if (rc == BEX_OK) {
    hipLaunchKernelGGL(mlp_mfma_kernel, dim3((unsigned)nb), dim3(64), 0, st, rows, stride, n, d, out);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = BEX_E_HIP;
  }
  hipFree(d);
  return rc;
}

int bex_set_warm_start_network(bex_solver* s, const float* params) {
  if (!s || !params) return BEX_E_ARGUMENT;
  for (int k = 0; k < BEX_MLP_NPARAM; k++)
    if (!std::isfinite(params[k]))
      return fail(s, BEX_E_ARGUMENT, "bex_set_warm_start_network: non-finite parameter");
  BEX_HIP(s, hipSetDevice(s->device));
  BexMlpDevice host;
  bex_mlp_pack_device(params, &host);
  if (!s->d_net) BEX_HIP(s, hipMalloc((void**)&s->d_net, sizeof(BexMlpDevice)));
  // On the solver's stream, complete before return: the host staging copy goes out of scope.
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  BEX_HIP(s, hipMemcpyAsync(s->d_net, &host, sizeof(BexMlpDevice), hipMemcpyHostToDevice,
                            s->stream));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  return BEX_OK;
}

int bex_enable_sampling(bex_solver* s, long long capacity, int every) {
  if (!s || capacity < 0 || (capacity > 0 && every < 1)) return BEX_E_ARGUMENT;
  BEX_HIP(s, hipSetDevice(s->device));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  if (s->d_samples) BEX_HIP(s, hipFree(s->d_samples));
  if (s->d_sample_count) BEX_HIP(s, hipFree(s->d_sample_count));
  s->d_samples = nullptr;
  s->d_sample_count = nullptr;
  s->SampleCapacity = 0;
  s->SampleEvery = 0;
  if (capacity == 0) return BEX_OK;
  KernelChoice kc;
  if (!kernel_for(s->d.UPS, s->d.Algorithm, s->KernelLayout, bexk::kSampling, bexk::kDenseNone, &kc))
    return fail(s, BEX_E_UNSUPPORTED,
                "bex_enable_sampling: sample collection needs 32 units and the role-pair layout");
  BEX_HIP(s, hipMalloc((void**)&s->d_samples, sizeof(float) * BEX_SAMPLE_FLOATS * (size_t)capacity));
  BEX_HIP(s, hipMalloc((void**)&s->d_sample_count, sizeof(unsigned long long)));
  BEX_HIP(s, hipMemsetAsync(s->d_sample_count, 0, sizeof(unsigned long long), s->stream));
  s->SampleCapacity = capacity;
  s->SampleEvery = every;
  return BEX_OK;
}

int bex_get_samples(bex_solver* s, float* out, long long max_rows, long long* rows) {
  if (!s || !rows || (max_rows > 0 && !out)) return BEX_E_ARGUMENT;
  if (!s->d_samples) return fail(s, BEX_E_ARGUMENT, "bex_get_samples: sampling not enabled");
  BEX_HIP(s, hipSetDevice(s->device));
  unsigned long long reserved = 0;
  BEX_HIP(s, hipMemcpyAsync(&reserved, s->d_sample_count, sizeof reserved, hipMemcpyDeviceToHost,
                            s->stream));
  BEX_HIP(s, hipStreamSynchronize(s->stream));
  long long n = (long long)reserved;
  if (n > s->SampleCapacity) n = s->SampleCapacity - s->SampleCapacity % 32;
  *rows = n;
  if (max_rows > 0 && n > 0) {
    const long long m = n < max_rows ? n : max_rows;
    BEX_HIP(s, hipMemcpy(out, s->d_samples, sizeof(float) * BEX_SAMPLE_FLOATS * (size_t)m,
                         hipMemcpyDeviceToHost));
  }
  return BEX_OK;
}

}  // extern "C"

// =========================================================================================
// Pools of uneven ensembles (no reference counterpart: the reference's per-block solver holds
// one N and one shared coupling matrix per object). A pool holds size classes; class c is a
// solver of NS[c] systems of N[c] units with one coupling matrix per system. Each class runs
// as a persistent kernel whose waves take systems from a device work queue in a given order
// (longest expected run first); one period of all classes is one hipGraph whose branches run
// the classes concurrently, replayed once per period.

namespace {

using PoolFn = void (*)(BexKernelArgs, const int*, long long, long long, unsigned long long*);

struct PoolVariant {
  const char* name;
  PoolFn fn;
  size_t lds;
  int block;
  int items_per_block;
  int systems_per_item;
};

template <class L>
PoolVariant pool_variant(const char* name) {
  return {name, (PoolFn)bexk::queue_kernel<L, bexk::kRKCK45>, L::lds_bytes(), L::kBlock,
          L::kItemsPerBlock, L::kSystemsPerItem};
}

const std::vector<PoolVariant>& pool_variants(int n) {
  static const std::vector<PoolVariant> v8 = {
      pool_variant<bexk::GroupLayout<8, 1, true>>("group8 x1 wave"),
      pool_variant<bexk::GroupLayout<8, 2, true>>("group8 x2 waves"),
      pool_variant<bexk::GroupLayout<8, 4, true>>("group8 x4 waves")};
  static const std::vector<PoolVariant> v16 = {
      pool_variant<bexk::GroupLayout<16, 1, true>>("group16 x1 wave"),
      pool_variant<bexk::GroupLayout<16, 2, true>>("group16 x2 waves"),
      pool_variant<bexk::GroupLayout<16, 4, true>>("group16 x4 waves")};
  static const std::vector<PoolVariant> v32 = {
      pool_variant<bexk::RolePairLayout32<bexk::kPlain, true>>("rolepair32"),
      pool_variant<bexk::GroupLayout<32, 1, true>>("group32 x1 wave"),
      pool_variant<bexk::GroupLayout<32, 2, true>>("group32 x2 waves")};
  static const std::vector<PoolVariant> v64 = {
      pool_variant<bexk::RolePairLayout64<true>>("rolepair64 x2 waves, M in registers"),
      pool_variant<bexk::GroupLayout<64, 1, true>>("group64 x1 wave"),
      pool_variant<bexk::GroupLayout<64, 2, true>>("group64 x2 waves")};
  static const std::vector<PoolVariant> v128 = {
      pool_variant<bexk::ChainLayout128<true>>("chain128 x8 waves, M in registers"),
      pool_variant<bexk::WideLayout<128, true, 2>>("wide128 lane pairs x4 waves"),
      pool_variant<bexk::WideLayout<128, true, 1>>("wide128 x2 waves")};
  // One lane per unit first: two systems per CU keep twice the matrix loads in flight.
  static const std::vector<PoolVariant> v256 = {
      pool_variant<bexk::WideLayout<256, true, 1>>("wide256 x4 waves, M streamed"),
      pool_variant<bexk::WideLayout<256, true, 2>>("wide256 lane pairs x8 waves, M streamed")};
  static const std::vector<PoolVariant> none;
  switch (n) {
    case 8: return v8;
    case 16: return v16;
    case 32: return v32;
    case 64: return v64;
    case 128: return v128;
    case 256: return v256;
    default: return none;
  }
}

}  // namespace

struct bex_pool_class {
  bex_solver* s = nullptr;
  int N = 0, NS = 0;
  int variant = 0;
  int* d_order = nullptr;
  unsigned long long* d_counter = nullptr;
  int config_epoch = -1;  // class solver's configuration at graph capture
};

struct bex_pool {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> side;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join, ready;
  std::vector<bex_pool_class> cls;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  int cu_count = 0;
  std::string err;
};

namespace {

int pfail(bex_pool* p, int code, const std::string& msg) {
  if (p)
    p->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define BEX_PHIP(p, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return pfail((p), BEX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Persistent launch of class c on stream st (queue counter cleared by a kernel first, so the
// sequence can be captured into a graph).
int launch_class(bex_pool* p, int c, hipStream_t st) {
  bex_pool_class& k = p->cls[c];
  const PoolVariant& v = pool_variants(k.N)[k.variant];
  BexKernelArgs a{};
  fill_args(k.s, &a);
  const int nsys = k.s->ActiveSystems < k.NS ? k.s->ActiveSystems : k.NS;
  if (nsys <= 0) return BEX_OK;
  const long long nitems = (k.NS + v.systems_per_item - 1) / v.systems_per_item;
  int per_cu = 0;
  BEX_PHIP(p, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(v.fn),
                                                           v.block, v.lds));
  if (per_cu < 1) return pfail(p, BEX_E_UNSUPPORTED, "pool: a class kernel does not fit on a CU");
  long long blocks = (long long)p->cu_count * per_cu;
  const long long need = (nitems + v.items_per_block - 1) / v.items_per_block;
  if (blocks > need) blocks = need;
  hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(1), 0, st, k.d_counter);
  hipLaunchKernelGGL(v.fn, dim3((unsigned)blocks), dim3(v.block), v.lds, st, a,
                     (const int*)k.d_order, (long long)k.NS, nitems, k.d_counter);
  BEX_PHIP(p, hipGetLastError());
  return BEX_OK;
}

bool graph_stale(const bex_pool* p) {
  if (!p->graph_valid) return true;
  for (const auto& k : p->cls)
    if (k.config_epoch != k.s->config_epoch) return true;
  return false;
}

int capture(bex_pool* p) {
  if (p->exec) {
    hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  // Branches in descending estimated cost (N^2 per unit of BiCG work, times the class size):
  // when the runtime maps more branches than it has hardware queues onto one queue, the
  // costliest class is then never the one left waiting behind another.
  std::vector<size_t> order(p->cls.size());
  for (size_t c = 0; c < order.size(); c++) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    const double cx = (double)p->cls[x].N * p->cls[x].N * p->cls[x].NS;
    const double cy = (double)p->cls[y].N * p->cls[y].N * p->cls[y].NS;
    return cx > cy;
  });
  BEX_PHIP(p, hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  int rc = BEX_OK;
  if (hipEventRecord(p->fork, p->stream) != hipSuccess) rc = BEX_E_HIP;
  for (size_t c : order) {
    if (rc != BEX_OK) break;
    if (hipStreamWaitEvent(p->side[c], p->fork, 0) != hipSuccess) {
      rc = BEX_E_HIP;
      break;
    }
    rc = launch_class(p, (int)c, p->side[c]);
    if (rc == BEX_OK && (hipEventRecord(p->join[c], p->side[c]) != hipSuccess ||
                         hipStreamWaitEvent(p->stream, p->join[c], 0) != hipSuccess))
      rc = BEX_E_HIP;
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (rc != BEX_OK) {
    if (g) hipGraphDestroy(g);
    return rc == BEX_E_HIP ? pfail(p, rc, "pool: graph capture failed") : rc;
  }
  if (e != hipSuccess) return pfail(p, BEX_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  p->graph = g;
  BEX_PHIP(p, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  for (auto& k : p->cls) k.config_epoch = k.s->config_epoch;
  p->graph_valid = true;
  return BEX_OK;
}

}  // namespace

// Solve() of a pool class's own solver: that class alone, with the pool's current variant and
// order, on the solver's stream.
static int pool_class_solve(bex_solver* s) {
  bex_pool* p = s->pool;
  if (s->WarmStart || s->d_samples)
    return fail(s, BEX_E_UNSUPPORTED,
                "Solve: pool classes run the plain method (no warm start, no sample collection)");
  if (s->d.NDO > 0) return fail(s, BEX_E_UNSUPPORTED, "Solve: pool classes have no dense output");
  const int rc = launch_class(p, s->pool_class, s->stream);
  if (rc != BEX_OK) return fail(s, rc, p->err);
  return BEX_OK;
}

extern "C" {

int bex_pool_variant_count(int units) { return (int)pool_variants(units).size(); }

const char* bex_pool_variant_name(int units, int variant) {
  const auto& v = pool_variants(units);
  return (variant >= 0 && variant < (int)v.size()) ? v[variant].name : nullptr;
}

int bex_pool_create(int nclass, const int* units, const int* systems, int device, bex_pool** out) {
  if (!out || !units || !systems || nclass <= 0) return BEX_E_ARGUMENT;
  *out = nullptr;
  bex_pool* p = new bex_pool();
  p->device = device;
  auto bail = [&](int code, const std::string& msg) {
    bex_pool_destroy(p);
    return pfail(nullptr, code, msg);
  };
  if (hipSetDevice(device) != hipSuccess) return bail(BEX_E_HIP, "bex_pool_create: hipSetDevice failed");
  if (hipDeviceGetAttribute(&p->cu_count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return bail(BEX_E_HIP, "bex_pool_create: device query failed");
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->done, hipEventDisableTiming) != hipSuccess)
    return bail(BEX_E_HIP, "bex_pool_create: stream/event creation failed");
  bex_problem_desc dims{};
  bex_compiled_dims(&dims);
  for (int c = 0; c < nclass; c++) {
    const int N = units[c], NS = systems[c];
    if (pool_variants(N).empty() || NS <= 0)
      return bail(BEX_E_UNSUPPORTED,
                  "bex_pool_create: units per system must be 8, 16, 32, 64, 128 or 256");
    bex_problem_desc d = dims;
    d.NS = NS;
    d.UPS = N;
    d.TPB = N;
    d.SPB = 1;
    d.NC = 1;
    d.CBW = 0;
    d.CCI = 0;
    d.NDO = 0;
    d.Algorithm = BEX_RKCK45;
    bex_pool_class k;
    k.N = N;
    k.NS = NS;
    p->cls.push_back(k);
    bex_pool_class& kk = p->cls.back();
    const int rc = bex_create(&d, device, &kk.s);
    if (rc != BEX_OK) return bail(rc, g_create_error);
    kk.s->PerSystemCoupling = true;
    kk.s->pool = p;
    kk.s->pool_class = c;
    std::vector<int> order(NS);
    for (int i = 0; i < NS; i++) order[i] = i;
    if (hipMalloc((void**)&kk.s->d_CouplingSys, sizeof(double) * (size_t)NS * N * N) != hipSuccess ||
        hipMemset(kk.s->d_CouplingSys, 0, sizeof(double) * (size_t)NS * N * N) != hipSuccess ||
        hipMalloc((void**)&kk.d_order, sizeof(int) * (size_t)NS) != hipSuccess ||
        hipMemcpy(kk.d_order, order.data(), sizeof(int) * (size_t)NS, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void**)&kk.d_counter, sizeof(unsigned long long)) != hipSuccess)
      return bail(BEX_E_MEMORY, "bex_pool_create: not enough device memory");
    if (N > 128 && (hipMalloc((void**)&kk.s->d_CouplingSysT, sizeof(double) * (size_t)NS * N * N) != hipSuccess ||
                    hipMemset(kk.s->d_CouplingSysT, 0, sizeof(double) * (size_t)NS * N * N) != hipSuccess))
      return bail(BEX_E_MEMORY, "bex_pool_create: not enough device memory");
    hipStream_t st = nullptr;
    hipEvent_t j = nullptr, r = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&r, hipEventDisableTiming) != hipSuccess)
      return bail(BEX_E_HIP, "bex_pool_create: stream/event creation failed");
    p->side.push_back(st);
    p->join.push_back(j);
    p->ready.push_back(r);
    for (const auto& v : pool_variants(N))
      if (v.lds > 65536 &&
          hipFuncSetAttribute(reinterpret_cast<const void*>(v.fn),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)v.lds) != hipSuccess)
        return bail(BEX_E_HIP, "bex_pool_create: LDS attribute failed");
  }
  // The matrix clears and order uploads above ran on the null stream, which the pool's
  // non-blocking streams do not wait for.
  if (hipDeviceSynchronize() != hipSuccess) return bail(BEX_E_HIP, "bex_pool_create: synchronisation failed");
  *out = p;
  return BEX_OK;
}

int bex_pool_destroy(bex_pool* p) {
  if (!p) return BEX_E_ARGUMENT;
  hipSetDevice(p->device);
  if (p->stream) hipStreamSynchronize(p->stream);
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  for (auto& k : p->cls) {
    if (k.d_order) hipFree(k.d_order);
    if (k.d_counter) hipFree(k.d_counter);
    if (k.s) bex_destroy(k.s);
  }
  for (auto st : p->side) hipStreamDestroy(st);
  for (auto e : p->join) hipEventDestroy(e);
  for (auto e : p->ready) hipEventDestroy(e);
  if (p->fork) hipEventDestroy(p->fork);
  if (p->done) hipEventDestroy(p->done);
  if (p->stream) hipStreamDestroy(p->stream);
  delete p;
  return BEX_OK;
}

int bex_pool_class_solver(bex_pool* p, int c, bex_solver** out) {
  if (!p || !out || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  *out = p->cls[c].s;
  return BEX_OK;
}

int bex_pool_set_couplings(bex_pool* p, int c, int first, int count, const double* M) {
  if (!p || !M || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  bex_pool_class& k = p->cls[c];
  if (first < 0 || count < 0 || first + count > k.NS)
    return pfail(p, BEX_E_BOUND, "bex_pool_set_couplings: system range out of bounds");
  BEX_PHIP(p, hipSetDevice(p->device));
  BEX_PHIP(p, hipStreamSynchronize(k.s->stream));
  BEX_PHIP(p, hipStreamSynchronize(p->stream));
  const size_t nn = (size_t)k.N * k.N;
  // On the class solver's stream, complete before return (the pool's graph and the class's own
  // Solve() come after it).
  BEX_PHIP(p, hipMemcpyAsync(k.s->d_CouplingSys + (size_t)first * nn, M,
                             sizeof(double) * nn * count, hipMemcpyHostToDevice, k.s->stream));
  std::vector<double> t;
  if (k.s->d_CouplingSysT && count > 0) {  // transposes for the streamed layouts
    t.resize(nn * (size_t)count);
    for (int q = 0; q < count; q++) {
      const double* src = M + (size_t)q * nn;
      double* dst = t.data() + (size_t)q * nn;
      for (int r = 0; r < k.N; r++)
        for (int c2 = 0; c2 < k.N; c2++) dst[(size_t)c2 + (size_t)r * k.N] = src[(size_t)r + (size_t)c2 * k.N];
    }
    BEX_PHIP(p, hipMemcpyAsync(k.s->d_CouplingSysT + (size_t)first * nn, t.data(),
                               sizeof(double) * nn * count, hipMemcpyHostToDevice, k.s->stream));
  }
  BEX_PHIP(p, hipStreamSynchronize(k.s->stream));
  return BEX_OK;
}

int bex_pool_set_order(bex_pool* p, int c, const int* order) {
  if (!p || !order || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  bex_pool_class& k = p->cls[c];
  std::vector<char> seen(k.NS, 0);
  for (int i = 0; i < k.NS; i++) {
    if (order[i] < 0 || order[i] >= k.NS || seen[order[i]])
      return pfail(p, BEX_E_ARGUMENT, "bex_pool_set_order: not a permutation of the class's systems");
    seen[order[i]] = 1;
  }
  BEX_PHIP(p, hipSetDevice(p->device));
  BEX_PHIP(p, hipStreamSynchronize(p->stream));
  // On the pool's stream, which the next period's graph follows; complete before return.
  BEX_PHIP(p, hipMemcpyAsync(k.d_order, order, sizeof(int) * (size_t)k.NS, hipMemcpyHostToDevice,
                             p->stream));
  BEX_PHIP(p, hipStreamSynchronize(p->stream));
  return BEX_OK;
}

int bex_pool_get_order(bex_pool* p, int c, int* order) {
  if (!p || !order || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  bex_pool_class& k = p->cls[c];
  BEX_PHIP(p, hipSetDevice(p->device));
  BEX_PHIP(p, hipStreamSynchronize(p->stream));
  BEX_PHIP(p, hipMemcpy(order, k.d_order, sizeof(int) * (size_t)k.NS, hipMemcpyDeviceToHost));
  return BEX_OK;
}

// Orders class c's systems by the BiCG iterations they needed so far (their accumulated
// counters), most expensive first; ties keep the index order.
int bex_pool_order_by_cost(bex_pool* p, int c) {
  if (!p || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  bex_pool_class& k = p->cls[c];
  std::vector<long long> st((size_t)k.NS * BEX_NSTAT);
  const int rc = bex_get_stats(k.s, st.data());
  if (rc != BEX_OK) return pfail(p, rc, k.s->err);
  std::vector<int> order(k.NS);
  for (int i = 0; i < k.NS; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const long long cx = st[(size_t)x * BEX_NSTAT + BEX_STAT_BICG_ITERATIONS] +
                         st[(size_t)x * BEX_NSTAT + BEX_STAT_STAGES];
    const long long cy = st[(size_t)y * BEX_NSTAT + BEX_STAT_BICG_ITERATIONS] +
                         st[(size_t)y * BEX_NSTAT + BEX_STAT_STAGES];
    return cx > cy;
  });
  return bex_pool_set_order(p, c, order.data());
}

int bex_pool_set_variant(bex_pool* p, int c, int variant) {
  if (!p || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  if (variant < 0 || variant >= (int)pool_variants(p->cls[c].N).size())
    return pfail(p, BEX_E_ARGUMENT, "bex_pool_set_variant: no such variant");
  p->cls[c].variant = variant;
  p->graph_valid = false;
  return BEX_OK;
}

int bex_pool_get_variant(bex_pool* p, int c, int* variant) {
  if (!p || !variant || c < 0 || c >= (int)p->cls.size()) return BEX_E_ARGUMENT;
  *variant = p->cls[c].variant;
  return BEX_OK;
}

// One period (Solve) of every class: one graph launch, ordered after pending work of the class
// solvers' streams and before their later work.
int bex_pool_solve_periods(bex_pool* p, int periods) {
  if (!p || periods < 0) return BEX_E_ARGUMENT;
  BEX_PHIP(p, hipSetDevice(p->device));
  if (graph_stale(p)) {
    const int rc = capture(p);
    if (rc != BEX_OK) return rc;
  }
  for (size_t c = 0; c < p->cls.size(); c++) {
    BEX_PHIP(p, hipEventRecord(p->ready[c], p->cls[c].s->stream));
    BEX_PHIP(p, hipStreamWaitEvent(p->stream, p->ready[c], 0));
  }
  for (int k = 0; k < periods; k++) BEX_PHIP(p, hipGraphLaunch(p->exec, p->stream));
  BEX_PHIP(p, hipEventRecord(p->done, p->stream));
  for (auto& k : p->cls) BEX_PHIP(p, hipStreamWaitEvent(k.s->stream, p->done, 0));
  return BEX_OK;
}

int bex_pool_solve(bex_pool* p) { return bex_pool_solve_periods(p, 1); }

int bex_pool_synchronise(bex_pool* p) {
  if (!p) return BEX_E_ARGUMENT;
  BEX_PHIP(p, hipSetDevice(p->device));
  BEX_PHIP(p, hipStreamSynchronize(p->stream));
  for (auto& k : p->cls) BEX_PHIP(p, hipStreamSynchronize(k.s->stream));
  return BEX_OK;
}

// Times every launch variant of every class on the first `probe_fraction` of the current time
// domain and keeps the fastest; device state, counters and the time domain are restored.
int bex_pool_autotune(bex_pool* p, double probe_fraction, double* times_ms) {
  if (!p || !(probe_fraction > 0.0) || probe_fraction > 1.0) return BEX_E_ARGUMENT;
  BEX_PHIP(p, hipSetDevice(p->device));
  BEX_PHIP(p, hipDeviceSynchronize());
  int slot = 0;
  for (size_t c = 0; c < p->cls.size(); c++) {
    bex_pool_class& k = p->cls[c];
    bex_solver* s = k.s;
    Buf* keep[] = {&s->ActualState, &s->ActualTime, &s->UnitAccessories, &s->IntegerUnitAccessories,
                   &s->SystemAccessories, &s->IntegerSystemAccessories, &s->TimeDomain};
    const size_t nkeep = sizeof(keep) / sizeof(keep[0]);
    std::vector<void*> saved(nkeep, nullptr);
    long long* stats_saved = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    const size_t sbytes = sizeof(long long) * (size_t)k.NS * BEX_NSTAT;
    const int old = k.variant;
    int rc = BEX_OK;
    std::string msg;
    auto check = [&](hipError_t e, const char* what) {
      if (e != hipSuccess && rc == BEX_OK) {
        rc = BEX_E_HIP;
        msg = std::string(what) + ": " + hipGetErrorString(e);
      }
      return rc == BEX_OK;
    };
    // Copies on the pool's stream, which the probe launches follow, each complete on return
    // (the snapshot, the probe time domain and the counters are read or reused right after).
    auto copy = [&](void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
      hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, p->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
      return e;
    };
    // Snapshot of the class's device state and counters.
    for (size_t b = 0; b < nkeep && rc == BEX_OK; b++)
      if (check(hipMalloc(&saved[b], keep[b]->bytes()), "hipMalloc"))
        check(copy(saved[b], keep[b]->d, keep[b]->bytes(), hipMemcpyDeviceToDevice), "hipMemcpy");
    if (check(hipMalloc((void**)&stats_saved, sbytes), "hipMalloc"))
      check(copy(stats_saved, s->d_stats, sbytes, hipMemcpyDeviceToDevice), "hipMemcpy");
    // Probe time domain: [TD0, TD0 + f (TD1 - TD0)] from the device copy.
    std::vector<double> td(2 * (size_t)k.NS);
    check(copy(td.data(), s->TimeDomain.d, sizeof(double) * td.size(), hipMemcpyDeviceToHost),
          "hipMemcpy");
    for (int i = 0; i < k.NS; i++) td[k.NS + i] = td[i] + probe_fraction * (td[k.NS + i] - td[i]);
    check(hipEventCreate(&t0), "hipEventCreate");
    check(hipEventCreate(&t1), "hipEventCreate");
    const int nv = (int)pool_variants(k.N).size();
    int best = old;
    float best_ms = 0.0f;
    for (int v = 0; v < nv && rc == BEX_OK; v++) {
      for (size_t b = 0; b < nkeep; b++)
        check(copy(keep[b]->d, saved[b], keep[b]->bytes(), hipMemcpyDeviceToDevice), "hipMemcpy");
      check(copy(s->TimeDomain.d, td.data(), sizeof(double) * td.size(), hipMemcpyHostToDevice),
            "hipMemcpy");
      if (rc != BEX_OK) break;
      k.variant = v;
      check(hipEventRecord(t0, p->stream), "hipEventRecord");
      const int lrc = launch_class(p, (int)c, p->stream);
      check(hipEventRecord(t1, p->stream), "hipEventRecord");
      check(hipEventSynchronize(t1), "hipEventSynchronize");
      if (lrc != BEX_OK && rc == BEX_OK) {
        rc = lrc;
        msg = p->err;
      }
      float ms = 0.0f;
      if (check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime")) {
        if (times_ms) times_ms[slot + v] = ms;
        if (v == 0 || ms < best_ms) {
          best_ms = ms;
          best = v;
        }
      }
    }
    slot += nv;
    // Every probe did the same per-system work (results do not depend on the variant): order
    // the class's work queue by it, longest first (LPT), as bex_pool_order_by_cost does with
    // the counters of earlier periods.
    std::vector<int> order;
    if (rc == BEX_OK && stats_saved) {
      std::vector<long long> now((size_t)k.NS * BEX_NSTAT), before((size_t)k.NS * BEX_NSTAT);
      if (check(copy(now.data(), s->d_stats, sbytes, hipMemcpyDeviceToHost), "hipMemcpy") &&
          check(copy(before.data(), stats_saved, sbytes, hipMemcpyDeviceToHost), "hipMemcpy")) {
        std::vector<long long> work(k.NS);
        for (int i = 0; i < k.NS; i++) {
          const size_t b = (size_t)i * BEX_NSTAT;
          work[i] = (now[b + BEX_STAT_BICG_ITERATIONS] + now[b + BEX_STAT_STAGES]) -
                    (before[b + BEX_STAT_BICG_ITERATIONS] + before[b + BEX_STAT_STAGES]);
        }
        order.resize(k.NS);
        for (int i = 0; i < k.NS; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
      }
    }
    // Restore the snapshot whatever happened, then release it.
    for (size_t b = 0; b < nkeep; b++)
      if (saved[b]) check(copy(keep[b]->d, saved[b], keep[b]->bytes(), hipMemcpyDeviceToDevice),
                          "hipMemcpy");
    if (stats_saved) check(copy(s->d_stats, stats_saved, sbytes, hipMemcpyDeviceToDevice), "hipMemcpy");
    for (void* d : saved)
      if (d) hipFree(d);
    if (stats_saved) hipFree(stats_saved);
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    k.variant = (rc == BEX_OK) ? best : old;
    if (rc != BEX_OK) {
      p->graph_valid = false;
      return pfail(p, rc, "bex_pool_autotune: " + msg);
    }
    if (!order.empty()) {
      const int orc = bex_pool_set_order(p, (int)c, order.data());
      if (orc != BEX_OK) return orc;
    }
  }
  p->graph_valid = false;
  return BEX_OK;
}

const char* bex_pool_last_error(const bex_pool* p) { return p ? p->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
