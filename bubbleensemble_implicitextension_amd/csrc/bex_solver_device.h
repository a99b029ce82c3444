// gfx950 solver kernel: adaptive RKCK45 (or fixed-step RK4) integration of coupled
// Keller-Miksis ensembles with an in-wave BiCG solve of the implicit coupling.
//
// Work decomposition (MI355X, wave64):
//   * one lane = one unit (bubble), one group of N consecutive lanes = one system;
//     a wave holds 64/N systems, a 256-thread workgroup holds 4 waves;
//   * the coupling matrix (coupling 0, shared by all systems as in the reference) and its
//     transpose live in LDS, loaded once per workgroup;
//   * every system-level reduction (dot products, residual max, step acceptance, step-size
//     minimum, event step) is a butterfly inside the lane group: no workgroup barrier is
//     executed after the initial matrix load, so systems in one workgroup never wait for
//     each other;
//   * vectors that every unit needs (the BiCG search directions) are exchanged through a
//     per-system LDS slot, read back as broadcast ds_read_b128 of (u_j, w_j) pairs.
//
// Arithmetic order is fixed (explicit fma, -ffp-contract=off) and identical to the CPU
// oracle's canonical mode, so the kernel reproduces it bit for bit.
//
// Reference: Solver.cuh = SourceCodes/CoupledSystmes_PerBlock_Solver.cuh,
// Steppers.cuh = SourceCodes/CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh,
// ErrorControllers.cuh, EventHandling.cuh, DenseOutput.cuh (SourceCodes/CoupledSystems_PerBlock_*).
#pragma once

#include <hip/hip_runtime.h>

#include "bex_detmath.h"
#include "bex_types.h"

#ifndef BEX_SYSTEM_DEFINITION
#define BEX_SYSTEM_DEFINITION "km_system_definition.h"
#endif
#include BEX_SYSTEM_DEFINITION

namespace bexk {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

enum Algorithm { kRK4 = 0, kRKCK45 = 1 };

template <int N>
struct Geometry {
  static_assert(N == 8 || N == 16 || N == 32 || N == 64, "units per system: 8, 16, 32 or 64");
  static constexpr int SPW = kWave / N;               // systems per wave
  static constexpr int SPB = SPW * kWavesPerBlock;    // systems per workgroup
};

// System-scope values the user hooks receive by reference (the reference keeps them in
// shared memory, Solver.cuh:51-74).
template <class D>
struct SysShared {
  double T;
  double dT;
  double TD[2];
  double sPAR[D::NSP > 0 ? D::NSP : 1];
  double sACC[D::NSA > 0 ? D::NSA : 1];
  int isACC[D::NiSA > 0 ? D::NiSA : 1];
  int UDT;
  int DOIDX;
};

template <int N>
__host__ __device__ constexpr size_t lds_bytes() {
  using D = BexSystemDims;
  return sizeof(double) * 2 * N * N + sizeof(double2) * Geometry<N>::SPB * N +
         sizeof(SysShared<D>) * Geometry<N>::SPB + sizeof(int) * Geometry<N>::SPB * BEX_NSTAT;
}

// Orders LDS traffic between the lanes of one wave: LDS operations of a wave complete in
// issue order, so a compiler-level barrier is all that is needed.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the N lanes of a group. The butterfly leaves every lane with
// ((v0+v[N/2]) + ...) in exactly the association of the reference's halving tree
// (Steppers.cuh:733-742); IEEE addition is commutative, so all lanes agree bitwise.
template <int N>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int h = N / 2; h >= 1; h >>= 1) v = v + __shfl_xor(v, h, kWave);
  return v;
}

// Maximum with the C fmax NaN rule (Steppers.cuh:776-805 uses fmax).
template <int N>
__device__ __forceinline__ double group_fmax(double v) {
#pragma unroll
  for (int h = N / 2; h >= 1; h >>= 1) v = bex::fmax_d(v, __shfl_xor(v, h, kWave));
  return v;
}

// Minimum where a NaN candidate never replaces a value (MathFunction.cuh:54-64 atomicFMIN).
template <int N>
__device__ __forceinline__ double group_min_keep(double v) {
#pragma unroll
  for (int h = N / 2; h >= 1; h >>= 1) {
    double o = __shfl_xor(v, h, kWave);
    v = (o < v) ? o : v;
  }
  return v;
}

template <int N>
__device__ __forceinline__ unsigned long long group_mask(int grp) {
  return (N == 64) ? ~0ull : (((1ull << N) - 1ull) << (grp * N));
}

template <int N>
__device__ __forceinline__ bool group_all(bool p, int grp) {
  unsigned long long m = group_mask<N>(grp);
  return (__ballot(p) & m) == m;
}

template <int N>
__device__ __forceinline__ bool group_any(bool p, int grp) {
  return (__ballot(p) & group_mask<N>(grp)) != 0ull;
}

struct BicgOut {
  int iterations;
  bool rho_exit;
  bool maxit;
};

// One right-hand-side evaluation with the implicit coupling solve for unit uid
// (Steppers.cuh:431-562). On return F[1] holds the coupled acceleration.
//   A x = b,  A = I + diag(rD) M diag(x1^2),  b = xu + rD * (M CPT),  x0 = xu.
template <int N>
__device__ __forceinline__ BicgOut stage_rhs(double* F, double* Xs, double Ts, int sid, int uid,
                                             double* upar, SysShared<BexSystemDims>& S,
                                             double* uacc, int* iuacc, const BexKernelArgs& a,
                                             const double* sM,
                                             const double* sMT, double2* vec) {
  double cpt, cpf;
  CoupledSystems_PerBlock_OdeFunction<double>(sid, uid, F, Xs, Ts, upar, S.sPAR,
                                              a.GlobalParameters, a.IntegerGlobalParameters,
                                              uacc, iuacc, S.sACC, S.isACC, &cpt, &cpf);
  const double rD = cpf;
  const double x1sq = Xs[0] * Xs[0];  // Steppers.cuh:491
  const double xu = F[1];             // Steppers.cuh:492

  // b = rD * (M CPT) + xu   (Steppers.cuh:497-498, :576-584)
  wave_sync();
  vec[uid].x = cpt;
  wave_sync();
  double acc = 0.0;
#pragma unroll 8
  for (int j = 0; j < N; j++) acc = fma(sM[uid + j * N], vec[j].x, acc);
  const double b = fma(rD, acc, xu);

  // r0 = b - A xu   (Steppers.cuh:502-504, :647-666)
  wave_sync();
  vec[uid].x = x1sq * xu;
  wave_sync();
  acc = 0.0;
#pragma unroll 8
  for (int j = 0; j < N; j++) acc = fma(sM[uid + j * N], vec[j].x, acc);
  const double ax = fma(rD, acc, xu);
  double r = -ax + b;
  double rt = r;

  double x = xu, p = 0.0, pt = 0.0, rho2 = 0.0;
  BicgOut out{0, false, true};
  for (int iter = 1; iter <= a.BiCGMaxIterations; iter++) {  // Steppers.cuh:508
    const double rho1 = group_sum<N>(r * rt);                 // :511-513
    if (iter == 1) {                                          // :517-523
      p = r;
      pt = rt;
      if (bex::fabs_d(rho1) < a.BiCGRhoTolerance) {
        out.rho_exit = true;
        out.maxit = false;
        break;
      }
    } else {                                                  // :526-531
      const double beta = rho1 / rho2;
      p = r + beta * p;
      pt = rt + beta * pt;
    }
    // q = A p, qt = A^T pt   (:536, :669-706) as scaled products with M and M^T.
    wave_sync();
    vec[uid] = make_double2(x1sq * p, rD * pt);
    wave_sync();
    double a1 = 0.0, a2 = 0.0;
#pragma unroll 8
    for (int j = 0; j < N; j++) {
      const double2 v = vec[j];
      a1 = fma(sM[uid + j * N], v.x, a1);
      a2 = fma(sMT[uid + j * N], v.y, a2);
    }
    const double q = fma(rD, a1, p);
    const double qt = fma(x1sq, a2, pt);
    out.iterations++;
    const double alpha = rho1 / group_sum<N>(pt * q);        // :539-541
    x = x + alpha * p;                                        // :545-547
    r = r - alpha * q;
    rt = rt - alpha * qt;
    rho2 = rho1;                                              // :551-552
    if (group_fmax<N>(bex::fabs_d(r)) < a.BiCGResidualTolerance) {  // :556-559
      out.maxit = false;
      break;
    }
  }
  F[1] = x;
  return out;
}

template <int N, int ALG>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void solve_kernel(const BexKernelArgs a) {
  using D = BexSystemDims;
  using G = Geometry<N>;
  constexpr int UD = D::UD;
  constexpr int NUP = D::NUP;
  constexpr int NUA = D::NUA;
  constexpr int NiUA = D::NiUA;
  constexpr int NE = D::NE;
  static_assert(UD == 2, "the implicit coupling acts on component 1 of a two-component unit");

  extern __shared__ __align__(16) unsigned char smem[];
  double* sM = reinterpret_cast<double*>(smem);
  double* sMT = sM + N * N;
  double2* sVec = reinterpret_cast<double2*>(sMT + N * N);
  SysShared<D>* sSys = reinterpret_cast<SysShared<D>*>(sVec + G::SPB * N);
  int* sStat = reinterpret_cast<int*>(sSys + G::SPB);

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int grp = lane / N;
  const int uid = lane % N;
  const int lsys = wave * G::SPW + grp;
  const int sid = a.first_system + blockIdx.x * G::SPB + lsys;

  // Coupling matrix and its transpose into LDS (the only workgroup barrier).
  for (int k = tid; k < N * N; k += kBlock) {
    const double v = a.CouplingMatrix[k];
    sM[k] = v;
    sMT[(k / N) + (k % N) * N] = v;
  }
  for (int k = tid; k < G::SPB * BEX_NSTAT; k += kBlock) sStat[k] = 0;
  __syncthreads();

  const bool active = (sid < a.NS) && (sid < a.ActiveSystems);  // Solver.cuh:82,132
  if (!active) return;

  SysShared<D>& S = sSys[lsys];
  double2* vec = sVec + lsys * N;
  int* st = sStat + lsys * BEX_NSTAT;
  const long long TLT = a.TLT;
  const long long gtid = (long long)(sid / a.ref_spb) * a.ref_block_stride +
                         (long long)(sid % a.ref_spb) * N + uid;  // Interface.cuh:875-879
  const int NS = a.NS;

  // System scope (Solver.cuh:84-129).
  if (uid == 0) {
    S.TD[0] = a.TimeDomain[sid];
    S.TD[1] = a.TimeDomain[sid + NS];
    S.T = a.ActualTime[sid];
    S.dT = a.InitialTimeStep;
    for (int i = 0; i < D::NSP; i++) S.sPAR[i] = a.SystemParameters[sid + i * NS];
    for (int i = 0; i < D::NSA; i++) S.sACC[i] = a.SystemAccessories[sid + i * NS];
    for (int i = 0; i < D::NiSA; i++) S.isACC[i] = a.IntegerSystemAccessories[sid + i * NS];
    S.UDT = 0;
    S.DOIDX = (a.NDO > 0) ? a.DenseOutputIndex[sid] : 0;
  }

  // Unit scope into registers / private memory (Solver.cuh:234-257).
  double x[UD], xn[UD], err[UD];
  double upar[NUP > 0 ? NUP : 1];
  double uacc[NUA > 0 ? NUA : 1];
  int iuacc[NiUA > 0 ? NiUA : 1];
  double ea[NE > 0 ? NE : 1], en[NE > 0 ? NE : 1];
  double rtol[UD], atol[UD], etol[NE > 0 ? NE : 1];
  int edir[NE > 0 ? NE : 1];
#pragma unroll
  for (int i = 0; i < UD; i++) {
    x[i] = a.ActualState[gtid + i * TLT];
    rtol[i] = a.RelativeTolerance[i];
    atol[i] = a.AbsoluteTolerance[i];
  }
#pragma unroll
  for (int i = 0; i < NUP; i++) upar[i] = a.UnitParameters[gtid + i * TLT];
#pragma nounroll
  for (int i = 0; i < NUA; i++) uacc[i] = a.UnitAccessories[gtid + i * TLT];
#pragma unroll
  for (int i = 0; i < NiUA; i++) iuacc[i] = a.IntegerUnitAccessories[gtid + i * TLT];
#pragma unroll
  for (int i = 0; i < NE; i++) {
    etol[i] = a.EventTolerance[i];
    edir[i] = a.EventDirection[i];
  }
  wave_sync();

  // Initialization hook (Solver.cuh:262-281).
  CoupledSystems_PerBlock_Initialization<double>(sid, uid, S.DOIDX, S.T, S.dT, S.TD, x, upar,
                                                 S.sPAR, a.GlobalParameters,
                                                 a.IntegerGlobalParameters, uacc, iuacc, S.sACC,
                                                 S.isACC);
  wave_sync();
  double t = S.T;
  double dt = a.InitialTimeStep;
  double dtNew = a.InitialTimeStep;
  const double TD1 = S.TD[1];

  if (NE > 0)  // Solver.cuh:283-305
    CoupledSystems_PerBlock_EventFunction<double>(sid, uid, ea, t, dt, S.TD, x, upar, S.sPAR,
                                                  a.GlobalParameters,
                                                  a.IntegerGlobalParameters, uacc, iuacc,
                                                  S.sACC, S.isACC);

  // Dense output state (DenseOutput.cuh:313-395); the index lives in S.DOIDX.
  int updateDO = 0, skipped = 0;
  double doTime = 0.0;
  auto store_dense = [&](int updateStep) {
    if (updateDO == 1) {
      const int idx = S.DOIDX;
#pragma unroll
      for (int i = 0; i < UD; i++)
        a.DenseOutputStates[gtid + i * TLT + (long long)idx * TLT * UD] = x[i];
      if (uid == 0) a.DenseOutputTimeInstances[sid + (long long)idx * NS] = t;
      wave_sync();
      if (uid == 0) S.DOIDX = idx + 1;
      wave_sync();
      skipped = 0;
      doTime = bex::fmin_d(t + a.DenseOutputMinimumTimeStep, TD1);
    }
    if ((updateDO == 0) && (updateStep == 1)) skipped++;
  };
  if (a.NDO > 0) {
    updateDO = (S.DOIDX < a.NDO) ? 1 : 0;
    store_dense(0);
  }

  int terminate = 0;
  while (terminate == 0) {  // Solver.cuh:329-572
    int updateStep = 1, endReached = 0;
    dt = dtNew;
    if (dt > (TD1 - t)) {
      dt = TD1 - t;
      endReached = 1;
    }
    if (uid == 0) st[BEX_STAT_ATTEMPTS]++;

    bool finiteLane = true;
    double k1[UD], k2[UD], k3[UD], k4[UD], k5[UD], k6[UD];
    constexpr int NSTAGE = (ALG == kRKCK45) ? 6 : 4;
    // One copy of the right-hand side + BiCG body: the stage loop is not unrolled and each
    // stage's input state is formed in the reference's association (Steppers.cuh:153-400).
#pragma nounroll
    for (int s = 0; s < NSTAGE; s++) {
      double xs[UD], F[UD];
      double Ts;
#pragma unroll
      for (int i = 0; i < UD; i++) {
        if (ALG == kRKCK45) {
          switch (s) {
            case 0: xs[i] = x[i]; break;
            case 1: xs[i] = x[i] + (k1[i] * (1.0 / 5.0)) * dt; break;
            case 2: xs[i] = x[i] + (k1[i] * (3.0 / 40.0) + k2[i] * (9.0 / 40.0)) * dt; break;
            case 3:
              xs[i] = x[i] + (k1[i] * (3.0 / 10.0) + k2[i] * (-9.0 / 10.0) + k3[i] * (6.0 / 5.0)) * dt;
              break;
            case 4:
              xs[i] = x[i] + (k1[i] * (-11.0 / 54.0) + k2[i] * (5.0 / 2.0) +
                              k3[i] * (-70.0 / 27.0) + k4[i] * (35.0 / 27.0)) *
                                 dt;
              break;
            default:
              xs[i] = x[i] + (k1[i] * (1631.0 / 55296.0) + k2[i] * (175.0 / 512.0) +
                              k3[i] * (575.0 / 13824.0) + k4[i] * (44275.0 / 110592.0) +
                              k5[i] * (253.0 / 4096.0)) *
                                 dt;
              break;
          }
        } else {
          switch (s) {
            case 0: xs[i] = x[i]; break;
            case 1: xs[i] = x[i] + k1[i] * (0.5 * dt); break;
            case 2: xs[i] = x[i] + k2[i] * (0.5 * dt); break;
            default: xs[i] = x[i] + k3[i] * dt; break;
          }
        }
      }
      if (ALG == kRKCK45) {
        switch (s) {
          case 0: Ts = t; break;
          case 1: Ts = t + (1.0 / 5.0) * dt; break;
          case 2: Ts = t + (3.0 / 10.0) * dt; break;
          case 3: Ts = t + (3.0 / 5.0) * dt; break;
          case 4: Ts = t + dt; break;
          default: Ts = t + (7.0 / 8.0) * dt; break;
        }
      } else {
        Ts = (s == 0) ? t : ((s == 3) ? t + dt : t + 0.5 * dt);
      }
      const BicgOut o = stage_rhs<N>(F, xs, Ts, sid, uid, upar, S, uacc, iuacc, a, sM, sMT, vec);
      if (uid == 0) {
        st[BEX_STAT_STAGES]++;
        st[BEX_STAT_BICG_ITERATIONS] += o.iterations;
        st[BEX_STAT_BICG_RHO_EXITS] += o.rho_exit ? 1 : 0;
        st[BEX_STAT_BICG_MAXIT] += o.maxit ? 1 : 0;
        st[BEX_STAT_HIST0 + (o.iterations < 20 ? o.iterations : 19)]++;
      }
#pragma unroll
      for (int i = 0; i < UD; i++) {
        switch (s) {
          case 0: k1[i] = F[i]; break;
          case 1: k2[i] = F[i]; break;
          case 2: k3[i] = F[i]; break;
          case 3: k4[i] = F[i]; break;
          case 4: k5[i] = F[i]; break;
          default: k6[i] = F[i]; break;
        }
      }
    }

    if (ALG == kRKCK45) {
#pragma unroll
      for (int i = 0; i < UD; i++) {
        xn[i] = x[i] + (k1[i] * (37.0 / 378.0) + k3[i] * (250.0 / 621.0) +
                        k4[i] * (125.0 / 594.0) + k6[i] * (512.0 / 1771.0)) *
                           dt;
        double e = k1[i] * (37.0 / 378.0 - 2825.0 / 27648.0) +
                   k3[i] * (250.0 / 621.0 - 18575.0 / 48384.0) +
                   k4[i] * (125.0 / 594.0 - 13525.0 / 55296.0) +
                   k5[i] * (0.0 - 277.0 / 14336.0) + k6[i] * (512.0 / 1771.0 - 1.0 / 4.0);
        err[i] = dt * bex::fabs_d(e) + 1e-18;
        if (!bex::is_finite(xn[i]) || !bex::is_finite(err[i])) finiteLane = false;
      }
    } else {
#pragma unroll
      for (int i = 0; i < UD; i++) {
        xn[i] = x[i] + (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) * (dt * (1.0 / 6.0));
        err[i] = 0.0;
        if (!bex::is_finite(xn[i])) finiteLane = false;
      }
    }
    int isFinite = group_all<N>(finiteLane, grp) ? 1 : 0;

    if (ALG == kRKCK45) {
      // Error controller (ErrorControllers.cuh:453-542).
      double rr = 1e30;
      bool ru = true;
#pragma unroll
      for (int i = 0; i < UD; i++) {
        const double tol =
            bex::fmax_d(rtol[i] * bex::fmax_d(bex::fabs_d(xn[i]), bex::fabs_d(x[i])), atol[i]);
        ru = ru && (err[i] < tol);
        rr = bex::fmin_d(rr, tol / err[i]);
      }
      updateStep = group_all<N>(ru, grp) ? 1 : 0;
      const double relErr = group_min_keep<N>(rr);
      double mult = (updateStep == 1) ? 0.8 * bex::pow(relErr, 1.0 / 5.0)
                                      : 0.8 * bex::pow(relErr, 1.0 / 4.0);
      if (!bex::is_finite(mult)) isFinite = 0;
      if (isFinite == 0) {
        mult = a.TimeStepShrinkLimit;
        updateStep = 0;
        if (dt < (a.MinimumTimeStep * 1.01)) {
          terminate = 1;
          if (uid == 0) st[BEX_STAT_TERMINATED]++;
        }
      } else if (dt < (a.MinimumTimeStep * 1.01)) {
        updateStep = 1;
        if (uid == 0) st[BEX_STAT_FORCED_ACCEPTS]++;
      }
      mult = bex::fmin_d(mult, a.TimeStepGrowLimit);
      mult = bex::fmax_d(mult, a.TimeStepShrinkLimit);
      dtNew = dt * mult;
      dtNew = bex::fmin_d(dtNew, a.MaximumTimeStep);
      dtNew = bex::fmax_d(dtNew, a.MinimumTimeStep);
    } else {
      // RK4 controller (ErrorControllers.cuh:424-450).
      if (isFinite == 0) {
        terminate = 1;
        updateStep = 0;
        if (uid == 0) st[BEX_STAT_TERMINATED]++;
      }
      dtNew = a.InitialTimeStep;
    }

    if (NE > 0) {  // Solver.cuh:393-428, EventHandling.cuh:215-268
      if (terminate == 0)
        CoupledSystems_PerBlock_EventFunction<double>(sid, uid, en, t + dt, dtNew, S.TD, xn, upar,
                                                      S.sPAR, a.GlobalParameters,
                                                      a.IntegerGlobalParameters, uacc, iuacc,
                                                      S.sACC, S.isACC);
      double evLane = dt;
      bool corrLane = false;
      if (updateStep == 1 && terminate == 0) {
#pragma unroll
        for (int i = 0; i < NE; i++) {
          if (((ea[i] > etol[i]) && (en[i] < -etol[i]) && (edir[i] <= 0)) ||
              ((ea[i] < -etol[i]) && (en[i] > etol[i]) && (edir[i] >= 0))) {
            const double cand = -ea[i] / (en[i] - ea[i]) * dt;
            evLane = (cand < evLane) ? cand : evLane;
            corrLane = true;
          }
        }
      }
      if (group_any<N>(corrLane, grp)) {
        const double evStep = group_min_keep<N>(evLane);
        if (evStep < a.MinimumTimeStep) {
          if (uid == 0) st[BEX_STAT_EVENTS_OMITTED]++;
        } else {
          dtNew = evStep;
          updateStep = 0;
        }
      }
    }

    if (updateStep == 1) {  // Solver.cuh:433-525
      t += dt;
#pragma unroll
      for (int i = 0; i < UD; i++) x[i] = xn[i];
      if (uid == 0) {
        st[BEX_STAT_ACCEPTED]++;
        S.T = t;
        S.dT = dt;
      }
      wave_sync();
      CoupledSystems_PerBlock_ActionAfterSuccessfulTimeStep<double>(
          sid, uid, S.UDT, S.T, S.dT, S.TD, x, upar, S.sPAR, a.GlobalParameters,
          a.IntegerGlobalParameters, uacc, iuacc, S.sACC, S.isACC);
      wave_sync();
      if (NE > 0) {
#pragma unroll
        for (int i = 0; i < NE; i++) {
          if (((ea[i] > etol[i]) && (bex::fabs_d(en[i]) < etol[i]) && (edir[i] <= 0)) ||
              ((ea[i] < -etol[i]) && (bex::fabs_d(en[i]) < etol[i]) && (edir[i] >= 0))) {
            CoupledSystems_PerBlock_ActionAfterEventDetection<double>(
                sid, uid, i, S.UDT, S.T, S.dT, S.TD, x, upar, S.sPAR, a.GlobalParameters,
                a.IntegerGlobalParameters, uacc, iuacc, S.sACC, S.isACC);
            atomicAdd(&st[BEX_STAT_EVENTS], 1);
          }
        }
        wave_sync();
        CoupledSystems_PerBlock_EventFunction<double>(sid, uid, en, S.T, dtNew, S.TD, x, upar,
                                                      S.sPAR, a.GlobalParameters,
                                                      a.IntegerGlobalParameters, uacc, iuacc,
                                                      S.sACC, S.isACC);
#pragma unroll
        for (int i = 0; i < NE; i++) ea[i] = en[i];
      }
      wave_sync();
      t = S.T;
    } else if (uid == 0) {
      st[BEX_STAT_REJECTED]++;
    }
    const int udt = S.UDT;

    if (a.NDO > 0) {  // Solver.cuh:528-557
      if (updateStep == 1) {
        const int idx = S.DOIDX;
        updateDO = ((idx < a.NDO) && (doTime < t) && (skipped >= (a.DenseOutputSaveFrequency - 1)))
                       ? 1
                       : 0;
        if ((idx < a.NDO) && ((endReached == 1) || (udt == 1))) updateDO = 1;
      } else {
        updateDO = 0;
      }
      store_dense(updateStep);
    }

    if (updateStep == 1 && (endReached == 1 || udt == 1)) terminate = 1;  // :560-569
  }

  // Finalization hook and write back (Solver.cuh:577-647).
  if (uid == 0) S.T = t;
  wave_sync();
  CoupledSystems_PerBlock_Finalization<double>(sid, uid, S.DOIDX, S.T, S.dT, S.TD, x, upar,
                                               S.sPAR, a.GlobalParameters,
                                               a.IntegerGlobalParameters, uacc, iuacc, S.sACC,
                                               S.isACC);
  wave_sync();
  // Recompute the write-back addresses here instead of keeping the load-time ones alive
  // across the time loop (they would occupy ~2 registers per accessory).
  long long wb = gtid;
  asm volatile("" : "+v"(wb));
#pragma unroll
  for (int i = 0; i < UD; i++) a.ActualState[wb + i * TLT] = x[i];
#pragma nounroll
  for (int i = 0; i < NUA; i++) a.UnitAccessories[wb + i * TLT] = uacc[i];
#pragma unroll
  for (int i = 0; i < NiUA; i++) a.IntegerUnitAccessories[wb + i * TLT] = iuacc[i];
  if (uid == 0) {
    a.ActualTime[sid] = S.T;
    a.TimeDomain[sid] = S.TD[0];
    a.TimeDomain[sid + NS] = S.TD[1];
    for (int i = 0; i < D::NSA; i++) a.SystemAccessories[sid + i * NS] = S.sACC[i];
    for (int i = 0; i < D::NiSA; i++) a.IntegerSystemAccessories[sid + i * NS] = S.isACC[i];
    if (a.NDO > 0) a.DenseOutputIndex[sid] = S.DOIDX;
  }
  for (int k = uid; k < BEX_NSTAT; k += N) a.Stats[(long long)sid * BEX_NSTAT + k] += st[k];
}

}  // namespace bexk
