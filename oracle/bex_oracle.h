/* CPU oracle for the coupled Keller-Miksis ensemble solver.
 *
 * TEST INFRASTRUCTURE ONLY. This is a host restatement of the reference's per-block solver
 * path (kernel time loop, RKCK45 stepper, in-block BiCG coupling solve, error controller,
 * event handling, dense output and the Keller-Miksis system definition). Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and only as the
 * checker. The product path never calls it.
 *
 * Two arithmetic modes:
 *   BEX_ORACLE_CANONICAL  deterministic elementary functions (bex_detmath.h) and the
 *                         association order the gfx950 kernel uses; results are bit-identical
 *                         to the GPU for identical inputs.
 *   BEX_ORACLE_LITERAL    glibc pow/sin/cos and the exact association of the reference source
 *                         (row accumulation, per-column tree reductions for A^T p); agrees with
 *                         the canonical mode to rounding, up to the chaos of the trajectory.
 *
 * Buffers use the reference's structure-of-arrays layout with one system per block and
 * TPB == UPS (CoupledSystems_PerBlock_Interface.cuh:875-879,918-919,1007,1125):
 *   unit scope   [k][NS*N]   index  sid*N + uid + k*NS*N
 *   system scope [k][NS]     index  sid + k*NS
 *   coupling     [c][N][N]   index  row + col*N + c*N*N
 *   dense state  [idx][comp][NS*N]
 */
#ifndef BEX_ORACLE_H
#define BEX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define BEX_ORACLE_CANONICAL 0
#define BEX_ORACLE_LITERAL 1

/* Per-system counters; the GPU library fills the same layout. */
#define BEX_STAT_ACCEPTED 0
#define BEX_STAT_REJECTED 1
#define BEX_STAT_STAGES 2
#define BEX_STAT_BICG_ITERATIONS 3
#define BEX_STAT_BICG_RHO_EXITS 4
#define BEX_STAT_BICG_MAXIT 5
#define BEX_STAT_EVENTS 6
#define BEX_STAT_FORCED_ACCEPTS 7
#define BEX_STAT_TERMINATED 8
#define BEX_STAT_EVENTS_OMITTED 9
#define BEX_STAT_ATTEMPTS 10
#define BEX_STAT_HIST0 11 /* 20 bins: stages that ran k = 0..19 BiCG iterations */
#define BEX_NSTAT 32

typedef struct {
  int NS, N, UD, NUP, NSP, NGP, NiGP, NUA, NiUA, NSA, NiSA, NE, NDO, NC;
} bex_oracle_dims;

typedef struct {
  double InitialTimeStep;
  int ActiveSystems;
  double MaximumTimeStep;
  double MinimumTimeStep;
  double TimeStepGrowLimit;
  double TimeStepShrinkLimit;
  int DenseOutputSaveFrequency;
  double DenseOutputMinimumTimeStep;
  double BiCGRhoTolerance;      /* reference: 1e-6 (Steppers.cuh:522) */
  double BiCGResidualTolerance; /* reference: 1e-6 (Steppers.cuh:558) */
  int BiCGMaxIterations;        /* reference: 19  (Steppers.cuh:508) */
  const double* RelativeTolerance; /* [UD] */
  const double* AbsoluteTolerance; /* [UD] */
  const double* EventTolerance;    /* [NE] */
  const int* EventDirection;       /* [NE] */
  int Algorithm;                   /* 0 = RK4 (fixed step), 1 = RKCK45 */
  int DenseOutputInterpolation;    /* 0 = step-point sampling (reference), 1 = Hermite grid */
} bex_oracle_options;

typedef struct {
  double* TimeDomain;
  double* ActualState;
  double* ActualTime;
  double* UnitParameters;
  double* SystemParameters;
  double* GlobalParameters;
  int* IntegerGlobalParameters;
  double* UnitAccessories;
  int* IntegerUnitAccessories;
  double* SystemAccessories;
  int* IntegerSystemAccessories;
  double* CouplingMatrix;
  double* DenseOutputTimeInstances;
  double* DenseOutputStates;
  int* DenseOutputIndex;
} bex_oracle_buffers;

/* One Solve(): integrates every active system over its time domain.
 * stats: NS*BEX_NSTAT int64 counters, accumulated. Returns 0 on success. */
int bex_oracle_solve(const bex_oracle_dims* dims, const bex_oracle_options* opt,
                     bex_oracle_buffers* buf, long long* stats, int mode, int nthreads);

/* Stand-alone BiCG coupling solve for one system (tests): given x1 = X[0] per unit, the
 * row factors rD, explicit coupling terms CPT and the uncoupled accelerations xu, returns
 * the accelerations in x and the iteration count. */
int bex_oracle_bicg(int N, const double* M, const double* x1, const double* rD,
                    const double* CPT, const double* xu, double rho_tol, double res_tol,
                    int max_iter, int mode, double* x);

/* Keller-Miksis right-hand side for one unit (tests). */
void bex_oracle_km_rhs(const double* X, double T, const double* uPAR, int mode, double* F,
                       double* CPT, double* CPF);

/* The oracle's RKCK45 / RK4 tableaux applied to y' = lambda*y for one step (tests). */
int bex_oracle_rkck45_linear(double lambda, double y0, double h, int mode, double* y5,
                             double* err);
int bex_oracle_rk4_linear(double lambda, double y0, double h, int mode, double* y);

/* The RKCK45 step-size controller after its reductions (tests). */
void bex_oracle_control(int updateStep, int isFinite, double relErr, double dt,
                        const bex_oracle_options* o, int mode, double* dtNew, int* updateOut,
                        int* terminate, int* forced);

/* Deterministic elementary functions, exposed so tests can measure their accuracy. */
double bex_oracle_pow(double x, double y, int mode);
void bex_oracle_sincospi(double x, int mode, double* s, double* c);
double bex_oracle_sin(double x, int mode);
double bex_oracle_cos(double x, int mode);

#ifdef __cplusplus
}
#endif
#endif
