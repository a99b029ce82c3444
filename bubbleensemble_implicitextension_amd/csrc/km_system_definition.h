// Keller-Miksis bubble-ensemble system definition: the user hooks of the per-block solver.
//
// Same hook names, argument lists and semantics as the reference's system-definition header
// (BubbleEnsemble_SystemDefinition.cuh:7-126). Differences:
//   * the elementary functions come from bex_detmath.h (bit-identical on host and device)
//     instead of the device math library;
//   * the header also declares the model's dimensions (BexSystemDims), because the solver
//     library is compiled for one system definition at a time and keeps unit parameters and
//     accessories in registers and private memory sized at compile time.
//
// To solve a different model, write a header with the same hooks and BexSystemDims and build
// the library with -DBEX_SYSTEM_DEFINITION='"your_header.h"'.
#pragma once

#include "bex_detmath.h"

#ifndef PI
#define PI 3.14159265358979323846
#endif

#ifndef NCOLL
#define NCOLL 32  // recorded collapses per bubble (CollapseStrength.cu:22)
#endif

struct BexSystemDims {
  static constexpr int UD = 2;            // unit dimension: x1 = R/R_E, x2 = dx1/dtau
  static constexpr int NUP = 21;          // C0..C20 (CollapseStrength.cu:337-357)
  static constexpr int NSP = 0;
  static constexpr int NGP = 0;
  static constexpr int NiGP = 0;
  static constexpr int NUA = 2 * NCOLL;   // NCOLL maxima followed by NCOLL minima
  static constexpr int NiUA = 1;          // event counter
  static constexpr int NSA = 0;
  static constexpr int NiSA = 0;
  static constexpr int NE = 1;            // dR/dtau = 0
  static constexpr int NC = 1;
  // ActionAfterSuccessfulTimeStep and ActionAfterEventDetection write nothing but the unit's
  // accessories, and no other hook reads the accessories: the solver may then run the two
  // action hooks only on the one lane per unit whose accessories it writes back.
  static constexpr bool kActionsWriteOnlyAccessories = true;
};

// SYSTEM (BubbleEnsemble_SystemDefinition.cuh:7-49)
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_OdeFunction(
    int sid, int uid, Precision* F, Precision* X, Precision T, Precision* uPAR,
    Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC, int* iuACC,
    Precision* sACC, int* isACC, Precision* CPT, Precision* CPF) {
  Precision rx1 = 1.0 / X[0];
  Precision p = bex::pow(rx1, uPAR[10]);

  Precision s1;
  Precision c1;
  bex::sincospi(2.0 * T, &s1, &c1);

  // Second forcing frequency. bex::sincos(+-0) is (+0, 1) exactly, so the call is skipped for a
  // zero argument (f2 = theta = 0, single-frequency forcing) with the same result.
  Precision s2 = 0.0;
  Precision c2 = 1.0;
  const Precision a2 = 2.0 * uPAR[11] * PI * T + uPAR[12];
  if (a2 != 0.0) bex::sincos(a2, &s2, &c2);

  Precision N = (uPAR[0] + uPAR[1] * X[1]) * p - uPAR[2] * (1.0 + uPAR[9] * X[1]) -
                uPAR[3] * rx1 - uPAR[4] * X[1] * rx1 -
                1.5 * (1.0 - uPAR[9] * X[1] / 3.0) * X[1] * X[1] -
                (uPAR[5] * s1 + uPAR[6] * s2) * (1.0 + uPAR[9] * X[1]) -
                X[0] * (uPAR[7] * c1 + uPAR[8] * c2);
  Precision D = X[0] - uPAR[9] * X[0] * X[1] + uPAR[4] * uPAR[9];
  Precision rD = 1.0 / D;

  // Uncoupled part; F[1] is the unknown of the implicit coupling solve.
  F[0] = X[1];
  F[1] = N * rD;

  // Explicit part of the coupling.
  CPT[0] = -2 * X[0] * X[1] * X[1];

  // Row factor of the implicit coupling matrix.
  CPF[0] = rD;
}

// EVENTS (BubbleEnsemble_SystemDefinition.cuh:68-76): local extrema of R.
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_EventFunction(
    int sid, int uid, Precision* EF, Precision T, Precision dT, Precision* TD, Precision* X,
    Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC,
    int* iuACC, Precision* sACC, int* isACC) {
  EF[0] = X[1];
}

// Called after ActionAfterSuccessfulTimeStep (BubbleEnsemble_SystemDefinition.cuh:78-92):
// store the maximum and start tracking the following minimum.
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_ActionAfterEventDetection(
    int sid, int uid, int IDX, int& UDT, Precision& T, Precision& dT, Precision* TD,
    Precision* X, Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR,
    Precision* uACC, int* iuACC, Precision* sACC, int* isACC) {
  iuACC[0]++;

  if (iuACC[0] < NCOLL) {
    uACC[iuACC[0]] = X[0];
    uACC[iuACC[0] + NCOLL] = X[0];
  }
}

// ACCESSORIES (BubbleEnsemble_SystemDefinition.cuh:96-105): running minimum after a maximum.
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_ActionAfterSuccessfulTimeStep(
    int sid, int uid, int& UDT, Precision& T, Precision& dT, Precision* TD, Precision* X,
    Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC,
    int* iuACC, Precision* sACC, int* isACC) {
  if ((iuACC[0] < NCOLL) && (X[0] < uACC[iuACC[0] + NCOLL])) uACC[iuACC[0] + NCOLL] = X[0];
}

// BubbleEnsemble_SystemDefinition.cuh:107-116: every Solve() restarts at TD[0].
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_Initialization(
    int sid, int uid, int& DOIDX, Precision& T, Precision& dT, Precision* TD, Precision* X,
    Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC,
    int* iuACC, Precision* sACC, int* isACC) {
  if (uid == 0) T = TD[0];
}

// Warm-start features of one unit for the learned BiCG initial guess (no reference
// counterpart; only called when the warm start or training-sample collection is enabled).
// CRS is the row sum of the implicit operator's coupling, sum_j M_ij X_j[0]^2, supplied by the
// solver. Each feature is scaled to O(1) over the solver's operating range:
//   WF[0] = log(R/R_E), WF[1] = 10 Rdot/c_L, WF[2] = forcing pressure / C0 (both frequencies),
//   WF[3] = log(1 + CRS), WF[4] = log(dT) / 10.
// The network reads its inputs in bfloat16 (8 significant bits), so the features are evaluated
// in single precision with the hardware's log2 and sine (the sine's argument in revolutions);
// the solution itself never depends on them beyond the starting guess.
namespace bexfeat {
__forceinline__ __device__ __host__ float log_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
#else
  return logf(x);
#endif
}
// sin(2 pi x)
__forceinline__ __device__ __host__ float sin_rev(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sinf(x - rintf(x));
#else
  return sinf(6.283185307179586f * (x - rintf(x)));
#endif
}
}  // namespace bexfeat

template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_WarmStartFeatures(
    int sid, int uid, Precision* WF, Precision T, Precision dT, Precision* X, Precision* uPAR,
    Precision* sPAR, Precision* gPAR, int* igPAR, Precision CRS) {
  const float s1 = bexfeat::sin_rev((float)T);
  const float s2 = bexfeat::sin_rev((float)(uPAR[11] * T + uPAR[12] * (0.5 / PI)));
  WF[0] = bexfeat::log_f((float)X[0]);
  WF[1] = 10.0 * uPAR[9] * X[1];
  WF[2] = ((float)uPAR[5] * s1 + (float)uPAR[6] * s2) / (float)uPAR[0];
  WF[3] = bexfeat::log_f((float)(1.0 + CRS));
  WF[4] = 0.1f * bexfeat::log_f((float)dT);
}

// BubbleEnsemble_SystemDefinition.cuh:118-126
template <class Precision>
__forceinline__ __device__ __host__ void CoupledSystems_PerBlock_Finalization(
    int sid, int uid, int& DOIDX, Precision& T, Precision& dT, Precision* TD, Precision* X,
    Precision* uPAR, Precision* sPAR, Precision* gPAR, int* igPAR, Precision* uACC,
    int* iuACC, Precision* sACC, int* isACC) {}
