/* C ABI of the MI355X coupled Keller-Miksis ensemble solver (libbex.so).
 *
 * One opaque handle replaces one reference `ProblemSolver<...>` object
 * (SourceCodes/CoupledSystems_PerBlock_Interface.cuh:135-248). The handle owns pinned host
 * and device buffers in the reference's structure-of-arrays layout; values are exchanged by
 * copy through the Set/Get calls or in bulk through host views, and moved between host and
 * device explicitly, in stream order, exactly as in the reference.
 *
 * Every call returns BEX_OK (0) or a negative BEX_E_* code; bex_last_error() returns the
 * message. The C++ header CoupledSystems_PerBlock_Interface.hpp turns these codes back into
 * the reference's print-and-exit behaviour.
 *
 * Enumerations keep the reference's numeric order:
 *   Algorithms          Interface.cuh:33
 *   ListOfVariables     Interface.cuh:35-53
 *   ListOfSolverOptions Interface.cuh:55-68 (extended after SharedCouplingMatrices)
 */
#ifndef BEX_CAPI_H
#define BEX_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define BEX_OK 0
#define BEX_E_ARGUMENT -1    /* null handle/pointer, bad enum value */
#define BEX_E_BOUND -2       /* index outside its range (Interface.cuh:718-733 BoundCheck) */
#define BEX_E_VARIABLE -3    /* variable not valid for this overload (Interface.cuh:903-908) */
#define BEX_E_OPTION -4      /* option not valid for this overload (Interface.cuh:811-817) */
#define BEX_E_MEMORY -5      /* not enough device memory (Interface.cuh:463-471) */
#define BEX_E_HIP -6         /* HIP runtime error (Interface.cuh:13-22 gpuErrCHK) */
#define BEX_E_UNSUPPORTED -7 /* shape or dimension this build does not support */
#define BEX_E_MISMATCH -8    /* dimensions differ from the compiled system definition */

/* Algorithms (Interface.cuh:33) */
#define BEX_RK4 0
#define BEX_RKCK45 1

/* ListOfVariables (Interface.cuh:35-53) */
enum bex_variable {
  BEX_All = 0,
  BEX_TimeDomain,
  BEX_ActualTime,
  BEX_ActualState,
  BEX_UnitParameters,
  BEX_SystemParameters,
  BEX_GlobalParameters,
  BEX_IntegerGlobalParameters,
  BEX_UnitAccessories,
  BEX_IntegerUnitAccessories,
  BEX_SystemAccessories,
  BEX_IntegerSystemAccessories,
  BEX_CouplingMatrix,
  BEX_CouplingStrength,
  BEX_CouplingIndex,
  BEX_DenseOutput,
  BEX_DenseIndex,
  BEX_DenseTime,
  BEX_DenseState
};

/* ListOfSolverOptions (Interface.cuh:55-68), then this library's additions. */
enum bex_option {
  BEX_InitialTimeStep = 0,
  BEX_ActiveSystems,
  BEX_MaximumTimeStep,
  BEX_MinimumTimeStep,
  BEX_TimeStepGrowLimit,
  BEX_TimeStepShrinkLimit,
  BEX_RelativeTolerance,
  BEX_AbsoluteTolerance,
  BEX_EventTolerance,
  BEX_EventDirection,
  BEX_DenseOutputSaveFrequency,
  BEX_DenseOutputMinimumTimeStep,
  BEX_SharedGlobalVariables,
  BEX_SharedCouplingMatrices,
  /* additions: the reference hard-codes these BiCG constants (Steppers.cuh:508,522,558) */
  BEX_BiCGRhoTolerance,
  BEX_BiCGResidualTolerance,
  BEX_BiCGMaxIterations,
  /* lane layout of the solver kernel: BEX_LAYOUT_* (no reference counterpart) */
  BEX_KernelLayout,
  /* accepted steps (per launch) after which a system's wave priority rises; 0 disables
   * (no reference counterpart; default 3000) */
  BEX_PriorityStepThreshold,
  /* 1: BiCG starts from the learned guess x0 = b * y(features) instead of the uncoupled
   * acceleration (Steppers.cuh:494-499), with a restart from the uncoupled acceleration when
   * that solve does not converge; needs bex_set_warm_start_network. 32 units only. */
  BEX_WarmStart,
  /* 1: dense output at t0 + k * DenseOutputMinimumTimeStep by cubic Hermite interpolation over
   * the accepted step containing each instant (from its end points and their derivatives),
   * instead of the reference's step-point sampling (DenseOutput.cuh:313-395); the index still
   * persists across Solve() calls. No reference counterpart. */
  BEX_DenseOutputInterpolation,
  /* > 0: a system that has made this many step attempts in one Solve() ends at its next check
   * (checks every 256 attempts), with its state and time written back as they stand; 0 (the
   * default) runs every system to the end of its time domain as the reference does. Bounds a
   * launch whose systems reach the Keller-Miksis Mach-1 singularity (millions of forced steps
   * at the minimum step size). No reference counterpart. */
  BEX_MaximumStepAttempts,
  /* 1 (the default): Solve() launches as many workgroups as fit on the device at once, and
   * each takes its next systems from a device work queue when it finishes its current ones
   * (hand-out order: bex_set_launch_order, else system index); 0: one workgroup per group of
   * systems, dispatched by the hardware in index order, as the reference launches
   * (Interface.cuh:1949-1956). The results are the same; the queue keeps every slot busy when
   * the systems' run times differ. No reference counterpart. */
  BEX_PersistentLaunch,
  /* > 0: at most this many workgroups in the persistent launch (0, the default: as many as fit
   * on the device). Leaves room on the device for other work that runs at the same time. No
   * reference counterpart. */
  BEX_PersistentWorkgroups
};

#define BEX_LAYOUT_AUTO 0      /* role pairs for 32, 64 and 128 units, lane groups otherwise */
#define BEX_LAYOUT_GROUP 1     /* lane = unit, 64/N systems per wave, M in LDS */
#define BEX_LAYOUT_ROLE_PAIR 2 /* 32 units: one system per wave, M / M^T rows in registers;
                                  64 and 128 units: two and eight waves per system, M / M^T in
                                  registers */
#define BEX_LAYOUT_SYSTEM_PAIR 3 /* 32 units: two systems per wave (lane = unit), both rows of
                                    M and M^T in registers; plain BiCG start only */

/* The reference's template parameters (Interface.cuh:135). */
typedef struct {
  int NS, UPS, UD, TPB, SPB, NC, CBW, CCI, NUP, NSP, NGP, NiGP, NUA, NiUA, NSA, NiSA, NE, NDO;
  int Algorithm;
} bex_problem_desc;

typedef struct bex_solver bex_solver;

/* Dimensions of the system definition compiled into this library (UD, NUP, NSP, NGP, NiGP,
 * NUA, NiUA, NSA, NiSA, NE, NC); the remaining fields are zero. No reference counterpart:
 * the reference compiles the system definition into the application instead. */
int bex_compiled_dims(bex_problem_desc* out);

/* ProblemSolver::ProblemSolver(int) (Interface.cuh:352-657) */
int bex_create(const bex_problem_desc* desc, int device, bex_solver** out);
/* ProblemSolver::~ProblemSolver() (Interface.cuh:660-714) */
int bex_destroy(bex_solver* s);

/* ProblemSolver::SolverOption(ListOfSolverOptions, T) (Interface.cuh:771-818) */
int bex_solver_option(bex_solver* s, int option, double value);
/* ProblemSolver::SolverOption(ListOfSolverOptions, int, T) (Interface.cuh:820-865) */
int bex_solver_option_indexed(bex_solver* s, int option, int index, double value);

/* SetHost overloads (Interface.cuh:867-1140); integer variables take the value cast to int. */
int bex_set_host_unit(bex_solver* s, int sid, int uid, int var, int k, double v);     /* :867-912 */
int bex_set_host_system(bex_solver* s, int sid, int var, int k, double v);            /* :914-960 */
int bex_set_host_global(bex_solver* s, int var, int k, double v);                     /* :962-996 */
int bex_set_host_coupling(bex_solver* s, int c, int var, int row, int col, double v); /* :998-1086 */
int bex_set_host_dense_index(bex_solver* s, int sid, int var, double v);              /* :1088-1107 */
int bex_set_host_dense_state(bex_solver* s, int sid, int uid, int var, int comp, int k,
                             double v);                                                /* :1109-1140 */

/* GetHost overloads (Interface.cuh:1338-1529) */
int bex_get_host_unit(bex_solver* s, int sid, int uid, int var, int k, double* v);
int bex_get_host_system(bex_solver* s, int sid, int var, int k, double* v);
int bex_get_host_global(bex_solver* s, int var, int k, double* v);
int bex_get_host_coupling(bex_solver* s, int c, int var, int row, int col, double* v);
int bex_get_host_dense_index(bex_solver* s, int sid, int var, double* v);
int bex_get_host_dense_state(bex_solver* s, int sid, int uid, int var, int comp, int k,
                             double* v);

/* Bulk access to a host buffer in the reference layout (no reference counterpart; the
 * reference only exposes element-wise Set/Get). *elem_bytes is 8 (double) or 4 (int). */
int bex_host_view(bex_solver* s, int var, void** ptr, long long* count, int* elem_bytes);

/* ProblemSolver::SynchroniseFromHostToDevice / FromDeviceToHost (Interface.cuh:1142-1336) */
int bex_sync_h2d(bex_solver* s, int var);
int bex_sync_d2h(bex_solver* s, int var);

/* ProblemSolver::Solve() (Interface.cuh:1949-1956): one asynchronous launch. */
int bex_solve(bex_solver* s);
/* ProblemSolver::SynchroniseDevice / InsertSynchronisationPoint / SynchroniseSolver
 * (Interface.cuh:1958-1983) */
int bex_synchronise_device(bex_solver* s);
int bex_insert_synchronisation_point(bex_solver* s);
int bex_synchronise_solver(bex_solver* s);

/* Device-resident checkpoint of everything one Solve() writes back (Solver.cuh:601-648):
 * TimeDomain, ActualState, ActualTime, UnitAccessories, IntegerUnitAccessories,
 * SystemAccessories, IntegerSystemAccessories and the dense-output index. bex_checkpoint_save
 * copies those device buffers into a device snapshot (allocated by the first save),
 * bex_checkpoint_restore copies the snapshot back; both are asynchronous device-to-device
 * copies on the solver's stream, ordered with Solve() like any other call. The reference
 * checkpoints the same variables through the host at a period boundary
 * (SynchroniseFromDeviceToHost(All), Interface.cuh:1310-1328, then SetHost and
 * SynchroniseFromHostToDevice); this keeps the copy in HBM. Restoring without a saved
 * checkpoint fails with BEX_E_ARGUMENT. */
int bex_checkpoint_save(bex_solver* s);
int bex_checkpoint_restore(bex_solver* s);

/* Cooperative stop (no reference counterpart): bex_request_stop sets a flag in host-mapped
 * memory that every running system of this solver reads every 256 step attempts; a system
 * that sees it set ends there, as if the user had terminated it, with its state, time and
 * accessories written back as they stand (ActualTime < TimeDomain[1] marks it). The flag stays
 * set, stopping later launches too, until bex_clear_stop. For time-bounded runs of workloads
 * whose stiffest systems need millions of steps. */
int bex_request_stop(bex_solver* s);
int bex_clear_stop(bex_solver* s);

/* Order in which the persistent launch hands out the systems (no reference counterpart):
 * order[0..n-1] lists every system index once (n = NS), longest expected run first for the
 * best balance; order = NULL or n = 0 returns to index order. On layouts with several systems
 * per wave, consecutive entries share a wave (for the system pair: entries 2k and 2k + 1), so
 * systems of similar cost should be neighbours. Results do not depend on it. */
int bex_set_launch_order(bex_solver* s, const int* order, int n);

/* Which systems take the learned BiCG start when WarmStart is on (no reference counterpart):
 * flags[0..n-1], n = NS, 0 = the plain start for that system (its results are then those of
 * the plain solve bit for bit); flags = NULL or n = 0: every system takes it. Lets one launch
 * run learned and plain systems side by side. */
int bex_set_warm_start_systems(bex_solver* s, const unsigned char* flags, int n);

/* Workgroups of the persistent launch that the next Solve() would start (as many as fit on the
 * device, or PersistentWorkgroups when that is smaller, and no more than the systems need; 0
 * without systems to solve), and the waves each has; BEX_E_UNSUPPORTED with PersistentLaunch 0
 * or per-system coupling matrices. No reference counterpart. */
int bex_persistent_grid(bex_solver* s, int* workgroups, int* waves_per_workgroup);

/* Per-system counters [NS][32] accumulated over Solve() calls (layout in bex_types.h);
 * copies device counters to `out` (synchronous). No reference counterpart. */
int bex_get_stats(bex_solver* s, long long* out);
int bex_reset_stats(bex_solver* s);

/* Kernel timing on the solver's own stream: bex_timer_start records a HIP event, bex_timer_stop
 * records a second one, waits for it and returns the elapsed milliseconds. No reference
 * counterpart (the reference times its drivers with clock()). */
int bex_timer_start(bex_solver* s);
int bex_timer_stop(bex_solver* s, double* elapsed_ms);

/* Per-system timeline of the most recent launch: when enabled, the kernel stores the device
 * wall-clock counter at the start and end of each system ([NS][2]); bex_get_timeline waits for
 * the solver's stream, copies it and returns the counter frequency. Enabling waits for the
 * whole device, then clears the timeline. No reference counterpart. */
int bex_enable_timeline(bex_solver* s, int on);
int bex_get_timeline(bex_solver* s, long long* out, double* ticks_per_second);
/* The timeline as it stands while a launch may still run (an end tick of 0 marks a system not
 * finished yet). The timeline lives in pinned host memory mapped into the device, so this
 * reads it without issuing any GPU command. For progress reports of long launches. */
int bex_peek_timeline(bex_solver* s, long long* out);

/* BiCG warm start (no reference counterpart). params: BEX_MLP_NPARAM (4609) floats in the
 * layout of bubbleensemble_implicitextension_amd/csrc/bex_mlp.h (W1[6][64], W2[64][64], b2[64],
 * w3[64], b3); copied, converted to the kernel's bf16 operand layout and uploaded. */
int bex_set_warm_start_network(bex_solver* s, const float* params);
/* Training-sample collection: while enabled (capacity > 0) and the warm start is off, every
 * `every`-th RK stage of each system appends 32 rows of 10 floats (5 warm-start features, the
 * target x/b clamped to [-5, 5], x, b, the uncoupled acceleration, 0) until `capacity` rows
 * are filled. capacity 0 disables. */
int bex_enable_sampling(bex_solver* s, long long capacity, int every);
/* Copies up to max_rows collected rows to out (host) and returns the number collected. */
int bex_get_samples(bex_solver* s, float* out, long long max_rows, long long* rows);

/* Warm-start network training on the device (no reference counterpart). All pointers are
 * device pointers; stream is a hipStream_t (NULL = default stream); calls are asynchronous.
 * rows: n samples of `stride` floats, features in columns 0-4, target in column 5.
 * bex_mlp_loss_grad writes the mean squared error to loss[0] and its gradient with respect
 * to the 4609 parameters to grad, using workspace (bex_mlp_workspace_floats(n) floats). */
long long bex_mlp_workspace_floats(long long n);
int bex_mlp_loss_grad(const float* rows, int stride, long long n, const float* params,
                      float* grad, float* loss, float* workspace, void* stream);
/* One Adam step (bias-corrected, step >= 1) on the 4609 parameters. */
int bex_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr,
                 float beta1, float beta2, float eps, void* stream);
/* Network output y for n rows (FP32 VALU, the training forward pass). */
int bex_mlp_predict(const float* rows, int stride, long long n, const float* params, float* out,
                    void* stream);
/* Network output y for n rows through the solver's BF16 MFMA inference path; params is a host
 * pointer (4609 floats), rows/out device pointers; synchronous. */
int bex_mlp_predict_mfma(const float* rows, int stride, long long n, const float* params,
                         float* out, void* stream);
/* The same inference without any allocation: no reference counterpart; for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
 * it). The packed network goes to a library-owned static device buffer of 9744 bytes per device
 * (one network) instead of a per-call hipMalloc; rows and out are the caller's device buffers
 * as above; rows beyond one launch's reach (2^31 - 1 blocks of 32) are rejected with
 * BEX_E_UNSUPPORTED. Every call on a device overwrites that device's buffer, so calls that
 * overlap on one device race. Synchronous. */
int bex_mlp_predict_mfma_noalloc(const float* rows, int stride, long long n, const float* params,
                                 float* out, void* stream);

/* Pools of uneven ensembles (no reference counterpart). A pool holds `nclass` size classes;
 * class c is a solver (bex_pool_class_solver: options, Set/Get, host views, synchronisation) of
 * systems[c] systems of units[c] units (8, 16, 32, 64 or 128), RKCK45, each system with its own
 * coupling matrix. bex_pool_solve runs one Solve() of every class as one hipGraph launch: one
 * persistent kernel per class whose waves take systems from a device work queue in the
 * class's order; bex_pool_solve_periods replays that graph `periods` times without host
 * synchronisation. Classes of one pool run concurrently. */
typedef struct bex_pool bex_pool;
int bex_pool_create(int nclass, const int* units, const int* systems, int device, bex_pool** out);
int bex_pool_destroy(bex_pool* p);
int bex_pool_class_solver(bex_pool* p, int c, bex_solver** out);
/* Coupling matrices of systems first..first+count-1 of class c: count consecutive N x N
 * column-major matrices (row + col*N), copied synchronously. */
int bex_pool_set_couplings(bex_pool* p, int c, int first, int count, const double* M);
/* Order in which class c's systems are handed out (a permutation of 0..NS-1). */
int bex_pool_set_order(bex_pool* p, int c, const int* order);
/* Orders class c by the BiCG work its systems have accumulated (counters), largest first. */
int bex_pool_order_by_cost(bex_pool* p, int c);
/* The order in which class c's systems are handed out (NS entries). */
int bex_pool_get_order(bex_pool* p, int c, int* order);
int bex_pool_solve(bex_pool* p);
int bex_pool_solve_periods(bex_pool* p, int periods);
int bex_pool_synchronise(bex_pool* p);
/* Launch variants (lane layout x waves per workgroup) available for a unit count. */
int bex_pool_variant_count(int units);
const char* bex_pool_variant_name(int units, int variant);
int bex_pool_set_variant(bex_pool* p, int c, int variant);
int bex_pool_get_variant(bex_pool* p, int c, int* variant);
/* Autotuner: times every variant of every class on the first probe_fraction of the current
 * time domain (state, counters and time domain restored afterwards) and selects the fastest.
 * It also orders each class's work queue by the per-system work of the probe, longest first
 * (call bex_pool_set_order afterwards to override). times_ms (optional) receives the times,
 * class by class, variant by variant. No reference counterpart. */
int bex_pool_autotune(bex_pool* p, double probe_fraction, double* times_ms);
const char* bex_pool_last_error(const bex_pool* p);

/* Message of the last failing call on this handle (or of the last bex_create failure when
 * s is NULL). */
const char* bex_last_error(const bex_solver* s);

/* Device properties that the reference's device helpers print (Interface.cuh:259-345). */
typedef struct {
  char name[256];                   /* device name */
  char arch[256];                   /* target id, e.g. "gfx950:sramecc+:xnack-" */
  long long total_global_mem;       /* bytes */
  long long shared_mem_per_block;   /* LDS bytes available to one workgroup */
  long long total_const_mem;        /* bytes */
  int regs_per_block;               /* 32-bit registers per workgroup */
  int multiprocessor_count;         /* compute units */
  int major, minor;                 /* architecture revision */
  int clock_rate_khz;
  int memory_clock_rate_khz;
  int memory_bus_width;             /* bits */
  int warp_size;                    /* wavefront size */
  int max_threads_per_multiprocessor;
  int max_threads_per_block;
  int max_threads_dim[3];
  int max_grid_size[3];
  int concurrent_copy;              /* copies overlap kernels */
  int concurrent_kernels;
  int ecc_enabled;
} bex_device_properties;

/* Number of visible devices (0, with BEX_OK, where there is none). Replaces the device loop
 * of ListCUDADevices (Interface.cuh:259-294). */
int bex_device_count(int* count);
/* Properties of one device (cudaGetDeviceProperties in Interface.cuh:264, :314). */
int bex_get_device_properties(int device, bex_device_properties* out);
/* The device whose revision is closest to major.minor (SelectDeviceByClosestRevision,
 * Interface.cuh:296-310). */
int bex_select_device_by_closest_revision(int major, int minor, int* device);

#ifdef __cplusplus
}
#endif
#endif
