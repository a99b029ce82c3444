// gfx950 solver kernels: adaptive RKCK45 (or fixed-step RK4) integration of coupled
// Keller-Miksis ensembles with an in-wave BiCG solve of the implicit coupling.
//
// The time loop (Solver.cuh:329-572 semantics) is written once and parametrised by a lane
// layout that owns the coupling matrix and implements the BiCG solve and the system-level
// reductions:
//
//   RolePairLayout32  N = 32, one system per wave. Lanes 0-31 ("primal") hold unit i and row i
//                     of M; lanes 32-63 ("shadow") hold the same unit and row i of M^T. One
//                     BiCG matrix-vector step computes A p on the primal half and A^T p~ on the
//                     shadow half at once: 32 v_fmac_f64 with a DPP row_newbcast operand and
//                     two v_permlane16_swap, no LDS traffic. Dot products pair the halves with
//                     v_permlane32_swap and reduce with the reference's halving tree through
//                     permlane16/row_shl DPP; every system-level decision is wave-uniform.
//   PairLayout32      N = 32, two systems per wave, lane = unit; every lane holds row i of M and
//                     of M^T, so the products run as in the role pair for both systems at once
//                     and everything else is issued once for the two.
//   GroupLayout<N>    N = 8, 16, 64: lane = unit, a wave holds 64/N systems, M and M^T live in
//                     LDS (one copy per 4-wave workgroup) and the search directions are
//                     exchanged through per-system LDS slots.
//
// Arithmetic order is fixed (explicit fma, -ffp-contract=off) and identical to the CPU oracle's
// canonical mode: every matrix-vector row is four fma chains over the columns j = q (mod 4)
// combined as (s0 + s1) + (s2 + s3), every dot product follows the reference's halving tree.
// The kernels reproduce the oracle bit for bit.
//
// Reference: Solver.cuh = SourceCodes/CoupledSystmes_PerBlock_Solver.cuh,
// Steppers.cuh = SourceCodes/CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh,
// ErrorControllers.cuh, EventHandling.cuh, DenseOutput.cuh (SourceCodes/CoupledSystems_PerBlock_*).
#pragma once

#include <hip/hip_runtime.h>

#include "bex_detmath.h"
#include "bex_mlp.h"
#include "bex_types.h"

#ifndef BEX_SYSTEM_DEFINITION
#define BEX_SYSTEM_DEFINITION "km_system_definition.h"
#endif
#include BEX_SYSTEM_DEFINITION

// Occupancy target of the role-pair kernel: 3 waves per SIMD leaves 168 VGPRs per lane for the
// 64 that hold the coupling-matrix row.
#ifndef BEX_ROLEPAIR_WAVES_PER_EU
#define BEX_ROLEPAIR_WAVES_PER_EU 3
#endif

namespace bexk {

constexpr int kWave = 64;

enum Algorithm { kRK4 = 0, kRKCK45 = 1 };

// BiCG initial guess and sample collection of a kernel instantiation.
enum WarmStartMode {
  kPlain = 0,    // x0 = uncoupled acceleration (the reference's choice)
  kSampling = 1, // plain, and every SampleEvery-th stage records (features, x / b, x, b)
  kLearned = 2   // x0 = b * network(features); plain restart if that solve does not converge
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
  // Interpolated dense output: the pending step's end times, the output grid and flags.
  double tA, tB, gridT0;
  long long kq;
  int pending, flushOnly;
};
using Sys = SysShared<BexSystemDims>;

// Orders LDS traffic between the lanes of one wave: LDS operations of a wave complete in
// issue order, so a compiler-level barrier is all that is needed.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double uniform(double v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double((int)hi, (int)lo);
}

struct BicgOut {
  int iterations;
  bool rho_exit;
  bool maxit;
  double rmax;  // max |r| after the last iteration (role-pair layout; +inf before the first)
};

// =========================================================================================
// Cross-lane building blocks (DPP, permlane swaps) shared by the layouts.

// acc[j % 4] = fma(z(lane j of the lane's 16-lane row), m[j], acc[j % 4]) for j = 0..15, in
// order: the canonical four-partial-sum order of every matrix-vector product. The s_nop covers
// the DPP read-after-VALU-write hazard of z (the compiler does not look inside the asm).
__device__ __forceinline__ void fmac_row16(double* acc, double z, const double* m) {
  asm volatile(
      "s_nop 1\n"
      "v_fmac_f64_dpp %0, %4, %5 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %4, %6 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %4, %7 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %4, %8 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %4, %9 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %4, %10 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %4, %11 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %4, %12 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %4, %13 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %4, %14 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %4, %15 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %4, %16 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %4, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %4, %18 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %4, %19 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %4, %20 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
      : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])
      : "v"(z), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]),
        "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]),
        "v"(m[14]), "v"(m[15]));
}

// A second register copy of v. A permlane swap overwrites both of its operands; handing it v
// and this copy costs one 64-bit move.
__device__ __forceinline__ double twin(double v) {
  asm("" : "+v"(v));
  return v;
}

// Row exchange inside each 32-lane half: returns the value of the half's even row and of its
// odd row at the lane's position in the row, i.e. z[c] and z[16 + c] of the lane's half.
__device__ __forceinline__ void rows_of_half(double z, double* even, double* odd) {
  const double c = twin(z);
  const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(z), __double2loint(c), false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(z), __double2hiint(c), false, false);
  *even = __hiloint2double((int)h[0], (int)l[0]);
  *odd = __hiloint2double((int)h[1], (int)l[1]);
}

// Value of the lower half and of the upper half at lanes i and i + 32.
__device__ __forceinline__ void halves(double z, double* lower, double* upper) {
  const double c = twin(z);
  const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(z), __double2loint(c), false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(z), __double2hiint(c), false, false);
  *lower = __hiloint2double((int)h[0], (int)l[0]);
  *upper = __hiloint2double((int)h[1], (int)l[1]);
}

// DPP row_ror:k (k = 8, 4, 2, 1): the value of lane (c - k) mod 16 of the lane's 16-lane row.
template <int K>
__device__ __forceinline__ double row_ror(double v) {
  return __builtin_amdgcn_mov_dpp(v, 0x120 + K, 0xf, 0xf, true);
}

// Halving tree over each 32-lane half (Steppers.cuh:733-742): s[i] = op(s[i], s[i + h]) for
// h = 16, 8, 4, 2, 1, evaluated as a butterfly so that every lane ends with its half's result
// (no broadcast). At each level a lane combines its value with that of a lane whose index
// differs by h modulo 2h; after the level both lanes of such a pair hold the same value, so the
// lane that computed op(upper, lower) holds op(lower, upper), bit for bit, when op is
// commutative (IEEE addition; VmaxOp). Lane 0 (lane 32) computes exactly the reference's tree.
template <class Op>
__device__ __forceinline__ double tree32(double s, Op op) {
  double even, odd;
  rows_of_half(s, &even, &odd);
  s = op(even, odd);
  s = op(s, row_ror<8>(s));
  s = op(s, row_ror<4>(s));
  s = op(s, row_ror<2>(s));
  s = op(s, row_ror<1>(s));
  return s;
}

// Two independent trees (as tree32) evaluated in lockstep, so that each one's cross-lane moves
// fill the other's dependency stalls.
template <class OpA, class OpB>
__device__ __forceinline__ void tree32x2(double& a, double& b, OpA opa, OpB opb) {
  double ea, oa, eb, ob;
  rows_of_half(a, &ea, &oa);
  rows_of_half(b, &eb, &ob);
  a = opa(ea, oa);
  b = opb(eb, ob);
  a = opa(a, row_ror<8>(a));
  b = opb(b, row_ror<8>(b));
  a = opa(a, row_ror<4>(a));
  b = opb(b, row_ror<4>(b));
  a = opa(a, row_ror<2>(a));
  b = opb(b, row_ror<2>(b));
  a = opa(a, row_ror<1>(a));
  b = opb(b, row_ror<1>(b));
}

// The levels h = 8, 4, 2, 1 of two trees (as tree32x2), in lockstep.
template <class OpA, class OpB>
__device__ __forceinline__ void tree16x2(double& a, double& b, OpA opa, OpB opb) {
  a = opa(a, row_ror<8>(a));
  b = opb(b, row_ror<8>(b));
  a = opa(a, row_ror<4>(a));
  b = opb(b, row_ror<4>(b));
  a = opa(a, row_ror<2>(a));
  b = opb(b, row_ror<2>(b));
  a = opa(a, row_ror<1>(a));
  b = opb(b, row_ror<1>(b));
}

// max(|a|, |b|) in one v_max_f64 with absolute-value source modifiers (as VmaxOp on |a|, |b|).
struct VmaxAbsOp {
  __device__ __forceinline__ double operator()(double a, double b) const {
    double r;
    asm("v_max_f64 %0, |%1|, |%2|" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
};

// max(a, b) for operands that are >= +0 or a quiet NaN, in one v_max_f64 (IEEE maxNum: a NaN
// operand yields the other one). On such operands it equals bex::fmax_d (the reference's fmax,
// MathFunction.cuh:9-38) and is commutative. Written as asm because the compiler's maxnum
// first re-quiets both operands (two more instructions) for signalling NaNs, which arithmetic
// never produces.
struct VmaxOp {
  __device__ __forceinline__ double operator()(double a, double b) const {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
};

// min(a, b) in one v_min_f64 (IEEE minNum), for operands that are > 0 and not NaN: equal to
// MinKeepOp there, and commutative.
struct VminOp {
  __device__ __forceinline__ double operator()(double a, double b) const {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
  }
};

struct AddOp {
  __device__ __forceinline__ double operator()(double a, double b) const { return a + b; }
};
struct FmaxOp {
  __device__ __forceinline__ double operator()(double a, double b) const {
    return bex::fmax_d(a, b);
  }
};
// Minimum where a NaN candidate never replaces a value (MathFunction.cuh:54-64 atomicFMIN).
struct MinKeepOp {
  __device__ __forceinline__ double operator()(double a, double b) const { return (b < a) ? b : a; }
};

// =========================================================================================
// Lane-per-unit layouts: GroupLayout<N, W, PS> (N <= 64, 64/N systems per wave, W waves per
// workgroup) and WideLayout<N> (N = 128, two waves per system). PS = each system has its own
// coupling matrix (pools of uneven ensembles); otherwise all systems share coupling 0.
//
// The coupling matrix lives in LDS. Shared matrices keep M and M^T column-major with stride N;
// a per-system matrix keeps one copy, column-major with column stride N + 1, which serves both
// products without bank conflicts: (M v)_i reads M[i + j (N+1)], (M^T v)_i reads M[j + i (N+1)].
// Either way every row product is the canonical four-partial-sum order.

// A x = b with A = I + diag(rD) M diag(x1^2), b = xu + rD (M CPT), x0 = xu
// (Steppers.cuh:489-561), for the lane-per-unit layouts. Returns x in F[1].
template <class L>
__device__ __forceinline__ BicgOut group_coupling(const L& lay, double* F, double x1, double cpt,
                                                  double rD, const BexKernelArgs& a) {
  const double x1sq = x1 * x1;
  const double xu = F[1];
  double2* vec = lay.vec;
  const int u = lay.uid();
  lay.sync();
  vec[u] = make_double2(cpt, x1sq * xu);
  lay.sync();
  double mc, mx;
  lay.mv_two(&mc, &mx);  // M CPT and M (x1^2 x0)
  const double b = fma(rD, mc, xu);
  const double ax = fma(rD, mx, xu);
  double r = -ax + b;
  double rt = r;
  double x = xu, p = 0.0, pt = 0.0, rho2 = 0.0;
  double rho1 = lay.sum(r * rt);
  BicgOut out{0, false, true, 0.0};
  for (int iter = 1; iter <= a.BiCGMaxIterations; iter++) {
    if (iter == 1) {
      p = r;
      pt = rt;
      if (bex::fabs_d(rho1) < a.BiCGRhoTolerance) {
        out.rho_exit = true;
        out.maxit = false;
        break;
      }
    } else {
      const double beta = rho1 / rho2;
      p = fma(beta, p, r);
      pt = fma(beta, pt, rt);
    }
    lay.sync();
    vec[u] = make_double2(x1sq * p, rD * pt);
    lay.sync();
    double a1, a2;
    lay.mv_mtv(&a1, &a2);
    const double q = fma(rD, a1, p);
    const double qt = fma(x1sq, a2, pt);
    out.iterations++;
    const double alpha = rho1 / lay.sum(pt * q);
    x = fma(alpha, p, x);
    r = fma(-alpha, q, r);
    rt = fma(-alpha, qt, rt);
    rho2 = rho1;
    const double rmax = lay.fmax_all(bex::fabs_d(r));
    rho1 = lay.sum(r * rt);
    if (rmax < a.BiCGResidualTolerance) {
      out.maxit = false;
      break;
    }
  }
  F[1] = x;
  return out;
}

// Row products in the canonical order: four partial sums over columns j = q (mod 4), each a
// sequential fma chain, combined as (s0 + s1) + (s2 + s3). m(j) is the row's j-th element,
// z(j) the vector's.
// Columns are fetched C at a time into registers before their FMAs, so C LDS reads are in
// flight at once instead of one per FMA.
template <int N, int C, class MF, class ZF>
__device__ __forceinline__ double row_dot4(MF m, ZF z) {
  static_assert(C % 4 == 0 && N % C == 0, "chunk of whole column quadruples");
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int j0 = 0; j0 < N; j0 += C) {
    double mm[C], zz[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
      mm[k] = m(j0 + k);
      zz[k] = z(j0 + k);
    }
#pragma unroll
    for (int k = 0; k < C; k++) s[k & 3] = fma(mm[k], zz[k], s[k & 3]);
  }
  return (s[0] + s[1]) + (s[2] + s[3]);
}

// Copies system sid's coupling matrix (column-major, stride N, in global memory) into a padded
// LDS copy (column stride N + 1); lanes u = 0..N-1 of the system's group cooperate.
template <int N>
__device__ __forceinline__ void load_padded(double* dst, const double* src, int u, int nthreads) {
  for (int k = u; k < N * N; k += nthreads) {
    const int row = k % N, col = k / N;
    dst[row + col * (N + 1)] = src[k];
  }
}

// Halving tree (Steppers.cuh:733-742) over the N lanes of each group of a wave, N = 8..64:
// s[i] = op(s[i], s[i + h]) for h = N/2, ..., 1, evaluated with cross-lane moves that keep
// that pairing (permlane32 swap for h = 32, permlane16 swap for h = 16, DPP row_shl for
// h <= 8); the result is the group's first lane's, broadcast to the group.
template <int N, class Op>
__device__ __forceinline__ double group_tree(double v, Op op, int grp) {
  if (N == 64) {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v = op(v, __hiloint2double((int)h[1], (int)l[1]));
  }
  if (N >= 32) {
    double even, odd;
    rows_of_half(v, &even, &odd);
    v = op(v, odd);
  }
  if (N >= 16) v = op(v, __builtin_amdgcn_mov_dpp(v, 0x108, 0xf, 0xf, true));  // row_shl:8
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x104, 0xf, 0xf, true));              // row_shl:4
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x102, 0xf, 0xf, true));              // row_shl:2
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x101, 0xf, 0xf, true));              // row_shl:1
  if (N == 64) return uniform(v);
  if (N == 32) {  // lanes 0 and 32 hold the two groups' results
    const int alo = __builtin_amdgcn_readlane((int)__double2loint(v), 0);
    const int ahi = __builtin_amdgcn_readlane((int)__double2hiint(v), 0);
    const int blo = __builtin_amdgcn_readlane((int)__double2loint(v), 32);
    const int bhi = __builtin_amdgcn_readlane((int)__double2hiint(v), 32);
    return grp == 0 ? __hiloint2double(ahi, alo) : __hiloint2double(bhi, blo);
  }
  if (N == 16) {  // row_newbcast:0: lane 0 of each 16-lane row to the row
    const int lo = __builtin_amdgcn_mov_dpp((int)__double2loint(v), 0x150, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)__double2hiint(v), 0x150, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
  }
  return __shfl(v, grp * N, kWave);  // N = 8
}

template <int N, int W = 4, bool PS = false>
struct GroupLayout {
  static_assert(N == 8 || N == 16 || N == 32 || N == 64, "units per system: 8, 16, 32 or 64");
  static_assert(W >= 1 && W <= 4, "1 to 4 waves per workgroup");
  static constexpr int kUnits = N;
  static constexpr int kSystemsPerWave = kWave / N;
  static constexpr int kWaves = W;
  static constexpr int kBlock = kWave * kWaves;
  static constexpr int kSystemsPerBlock = kSystemsPerWave * kWaves;
  // Per-system matrices bound the occupancy through LDS (N = 64: 33 KB per wave, at most four
  // waves per CU); give those kernels the registers instead.
  static constexpr int kWavesPerEU = (PS && N == 64) ? 1 : (PS && N == 32) ? 2 : 4;
  static constexpr int kChunk = (kWavesPerEU <= 2) ? 16 : 8;
  static constexpr int kMode = kPlain;
  static constexpr bool kPerSystemM = PS;
  static constexpr int kItemsPerBlock = W;  // persistent kernels: each wave pops its own item
  static constexpr int kSystemsPerItem = kSystemsPerWave;
  static constexpr int kLd = PS ? N + 1 : N;

  static constexpr size_t m_bytes() {
    return PS ? sizeof(double) * kSystemsPerBlock * N * (N + 1) : sizeof(double) * 2 * N * N;
  }
  static constexpr size_t lds_bytes() {
    return m_bytes() + sizeof(double2) * kSystemsPerBlock * N + sizeof(Sys) * kSystemsPerBlock +
           sizeof(int) * kSystemsPerBlock * BEX_NSTAT;
  }

  int lane, grp, uid_, lsys, wave, sid_;
  const double* sM;   // shared: M; per-system: this system's padded copy
  const double* sMT;  // shared: M^T
  double* myM;
  double2* vec;
  Sys* sys;
  int* stat;

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs& a) {
    const int tid = threadIdx.x;
    lane = tid & (kWave - 1);
    wave = tid / kWave;
    grp = lane / N;
    uid_ = lane % N;
    lsys = wave * kSystemsPerWave + grp;
    double* m = reinterpret_cast<double*>(smem);
    double2* v = reinterpret_cast<double2*>(smem + m_bytes());
    Sys* s = reinterpret_cast<Sys*>(v + kSystemsPerBlock * N);
    int* st = reinterpret_cast<int*>(s + kSystemsPerBlock);
    if (PS) {
      myM = m + (size_t)lsys * N * (N + 1);
      sM = myM;
      sMT = myM;
    } else {
      double* mt = m + N * N;
      for (int k = tid; k < N * N; k += kBlock) {
        m[k] = a.CouplingMatrix[k];
        mt[k] = a.CouplingMatrixT[k];
      }
      __syncthreads();  // the only workgroup barrier
      sM = m;
      sMT = mt;
      myM = nullptr;
    }
    vec = v + lsys * N;
    sys = s + lsys;
    stat = st + lsys * BEX_NSTAT;
  }

  // Assigns this lane's system for work item `item` (one item = the systems of one wave):
  // order == nullptr -> systems first_system + item * kSystemsPerWave + grp; otherwise the
  // order array's entries. Loads a per-system matrix and clears the group's counters.
  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    const long long pos = item * kSystemsPerWave + grp;
    int s = -1;
    if (order) {
      if (pos < count) s = order[pos];
    } else if (pos < count) {
      s = a.first_system + (int)pos;
    }
    sid_ = s;
    for (int k = uid_; k < BEX_NSTAT; k += N) stat[k] = 0;
    if (PS && s >= 0 && s < a.NS)
      load_padded<N>(myM, a.CouplingMatrix + (size_t)s * N * N, uid_, N);
    wave_sync();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return uid_; }
  __device__ __forceinline__ bool primal() const { return true; }
  __device__ __forceinline__ void sync() const { wave_sync(); }

  __device__ __forceinline__ unsigned long long mask() const {
    return (N == 64) ? ~0ull : (((1ull << N) - 1ull) << (grp * N));
  }
  __device__ __forceinline__ bool all(bool p) const { return (__ballot(p) & mask()) == mask(); }
  __device__ __forceinline__ bool any(bool p) const { return (__ballot(p) & mask()) != 0ull; }

  // Halving-tree reductions (Steppers.cuh:733-742) of the group, broadcast to the group.
  __device__ __forceinline__ double sum(double v) const { return group_tree<N>(v, AddOp(), grp); }
  __device__ __forceinline__ double fmax_all(double v) const {
    return group_tree<N>(v, FmaxOp(), grp);
  }
  // Minimum where a NaN candidate never replaces a value (MathFunction.cuh:54-64 atomicFMIN).
  __device__ __forceinline__ double min_keep(double v) const {
    return group_tree<N>(v, MinKeepOp(), grp);
  }

  __device__ __forceinline__ double mv() const {
    const double* m = sM;
    const double2* z = vec;
    const int i = uid_;
    return row_dot4<N, kChunk>([&](int j) { return m[i + j * kLd]; },
                               [&](int j) { return z[j].x; });
  }
  __device__ __forceinline__ double mtv() const {
    const double2* z = vec;
    const int i = uid_;
    if (PS) {
      const double* m = sM;
      return row_dot4<N, kChunk>([&](int j) { return m[j + i * kLd]; },
                                 [&](int j) { return z[j].y; });
    }
    const double* mt = sMT;
    return row_dot4<N, kChunk>([&](int j) { return mt[i + j * N]; },
                               [&](int j) { return z[j].y; });
  }

  // The two products of a BiCG iteration (M z on vec[].x, M^T w on vec[].y).
  __device__ __forceinline__ void mv_mtv(double* a1, double* a2) const {
    *a1 = mv();
    *a2 = mtv();
  }
  // M z1 and M z2 for z1 on vec[].x and z2 on vec[].y.
  __device__ __forceinline__ void mv_two(double* a1, double* a2) const {
    const double* m = sM;
    const double2* z = vec;
    const int i = uid_;
    row_dot4_two<N, kChunk>([&](int j) { return m[i + j * kLd]; },
                            [&](int j) { return z[j].x; }, [&](int j) { return z[j].y; }, a1,
                            a2);
  }

  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&&, bool) const {
    return group_coupling(*this, F, x1, cpt, rD, a);
  }
};

// Row product split over the two lanes of a unit (lane half h = 0, 1): half h runs the
// canonical chains j = 2h and j = 2h + 1 (mod 4), each sequential in j; s(2h) + s(2h+1) of the
// two halves are exchanged with a DPP quad swap and added, giving (s0 + s1) + (s2 + s3) on both
// lanes (IEEE addition is commutative). C columns per chunk are loaded before their FMAs.
template <int N, int C, class MF, class ZF>
__device__ __forceinline__ double row_dot4_pair(MF m, ZF z, int h) {
  static_assert(C % 4 == 0 && N % C == 0, "chunk of whole column quadruples");
  double sa = 0.0, sb = 0.0;
#pragma unroll 1
  for (int j0 = 0; j0 < N; j0 += C) {
    double mm[C / 2], zz[C / 2];
#pragma unroll
    for (int k = 0; k < C / 2; k++) {
      const int j = j0 + 4 * (k >> 1) + 2 * h + (k & 1);
      mm[k] = m(j);
      zz[k] = z(j);
    }
#pragma unroll
    for (int k = 0; k < C / 2; k += 2) {
      sa = fma(mm[k], zz[k], sa);
      sb = fma(mm[k + 1], zz[k + 1], sb);
    }
  }
  const double pair = sa + sb;
  // quad_perm [1, 0, 3, 2]: the partner lane of the unit
  const double other = __builtin_amdgcn_mov_dpp(pair, 0xB1, 0xf, 0xf, false);
  return h == 0 ? pair + other : other + pair;
}

// Two row products over the same matrix row, m(j) z1(j) and m(j) z2(j), each in row_dot4's
// canonical order (bit-identical to two row_dot4 calls), reading each m(j) once.
template <int N, int C, class MF, class ZF1, class ZF2>
__device__ __forceinline__ void row_dot4_two(MF m, ZF1 z1, ZF2 z2, double* r1, double* r2) {
  static_assert(C % 4 == 0 && N % C == 0, "chunk of whole column quadruples");
  double s[4] = {0.0, 0.0, 0.0, 0.0}, t[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int j0 = 0; j0 < N; j0 += C) {
    double mm[C];
#pragma unroll
    for (int k = 0; k < C; k++) mm[k] = m(j0 + k);
#pragma unroll
    for (int k = 0; k < C; k++) {
      s[k & 3] = fma(mm[k], z1(j0 + k), s[k & 3]);
      t[k & 3] = fma(mm[k], z2(j0 + k), t[k & 3]);
    }
  }
  *r1 = (s[0] + s[1]) + (s[2] + s[3]);
  *r2 = (t[0] + t[1]) + (t[2] + t[3]);
}

// The same for two lanes per unit (as row_dot4_pair).
template <int N, int C, class MF, class ZF1, class ZF2>
__device__ __forceinline__ void row_dot4_pair_two(MF m, ZF1 z1, ZF2 z2, int h, double* r1,
                                                  double* r2) {
  static_assert(C % 4 == 0 && N % C == 0, "chunk of whole column quadruples");
  double sa = 0.0, sb = 0.0, ta = 0.0, tb = 0.0;
#pragma unroll 1
  for (int j0 = 0; j0 < N; j0 += C) {
    double mm[C / 2];
#pragma unroll
    for (int k = 0; k < C / 2; k++) mm[k] = m(j0 + 4 * (k >> 1) + 2 * h + (k & 1));
#pragma unroll
    for (int k = 0; k < C / 2; k += 2) {
      const int j = j0 + 4 * (k >> 1) + 2 * h;
      sa = fma(mm[k], z1(j), sa);
      sb = fma(mm[k + 1], z1(j + 1), sb);
      ta = fma(mm[k], z2(j), ta);
      tb = fma(mm[k + 1], z2(j + 1), tb);
    }
  }
  const double p1 = sa + sb, p2 = ta + tb;
  const double o1 = __builtin_amdgcn_mov_dpp(p1, 0xB1, 0xf, 0xf, false);
  const double o2 = __builtin_amdgcn_mov_dpp(p2, 0xB1, 0xf, 0xf, false);
  *r1 = h == 0 ? p1 + o1 : o1 + p1;
  *r2 = h == 0 ? p2 + o2 : o2 + p2;
}

// Levels of the reference's reduction fold (Steppers.cuh:745-805) above 64 entries, for N a
// multiple of 32 from 96 to 256: the 32-unit warps beyond the largest power of two, nP2, fold
// onto the first ones (s[i] op s[i + 32 nP2] for i < 32 (N/32 - nP2)), then the halving levels
// run down to 64 entries. Returns entry l < 64 of that partial tree from the N values in red.
template <int N, class Op>
__device__ __forceinline__ double fold_to_64(const double* red, int l, Op op) {
  constexpr int nFull = N / 32;
  constexpr int nP2 = nFull >= 8 ? 8 : (nFull >= 4 ? 4 : 2);
  constexpr int lim = 32 * nP2;
  constexpr int folded = 32 * (nFull - nP2);
  auto t = [&](int i) { return i < folded ? op(red[i], red[i + lim]) : red[i]; };
  if (lim == 64) return t(l);
  if (lim == 128) return op(t(l), t(l + 64));
  return op(op(t(l), t(l + 128)), op(t(l + 64), t(l + 192)));
}

// N a multiple of 32 from 96 to 256: one system per workgroup, LPU lanes per unit
// (N * LPU / 64 waves, whole waves only). Reductions combine the waves through LDS (the fold's
// levels above 64 entries, fold_to_64) and finish with an in-wave tree that every wave
// evaluates identically; system-scope synchronisation is the workgroup barrier. N <= 128 keeps
// the system's matrix in LDS (132 KB for N = 128, one system per CU); larger matrices
// (a 256 x 256 FP64 matrix is 512 KB) exceed the CU's LDS, so they stream M and M^T
// (column-major) from global memory through L2 and MALL. With two lanes per unit, each lane
// runs half of a row's FMA chains (row_dot4_pair) and all four SIMDs of the CU take part; the
// per-unit work is duplicated on the unit's lanes and the first lane ("primal") writes results.
//
// The two products of a BiCG iteration read M once between them (mv_mtv): M is cut into
// kNB x kNB blocks of kBS x kBS (kNB = 8 / LPU, so that the 8 N row and column FMA chains of the
// pair, four per row of M z and four per row of M^T w, number kNB per thread), and the blocks
// are staged through LDS one anti-diagonal I + J = t at a time, t = 0 .. 2 kNB - 2. Each thread
// owns one chain per block index b: threads of the first half the chain c of row b kBS + rho of
// M z, which meets block (b, t - b) at step t, i.e. its columns in ascending order; threads of
// the second half the chain c of column b kBS + rho of M (row of M^T w), which meets block
// (t - b, b), its rows in ascending order. So every chain is the canonical sequential chain
// over the indices = c (mod 4), every thread runs the same number of chain segments at every
// step, and the four chain sums combine as (s0 + s1) + (s2 + s3) as in row_dot4. The chain
// accumulators live in LDS ([2][4][N], indexed by row and chain), an anti-diagonal is staged
// kSlots blocks at a time, and the matrix traffic per iteration halves against reading M and
// M^T separately.
template <int N, bool PS = true, int LPU = 1>
struct WideLayout {
  static_assert(N % 32 == 0 && N >= 96 && N <= 256, "multi-wave layout: 96 to 256 units");
  static_assert(LPU == 1 || LPU == 2, "one or two lanes per unit");
  static_assert((N * LPU) % 64 == 0, "whole waves only");
  static constexpr bool kLdsM = (N <= 128);
  static constexpr int kUnits = N;
  static constexpr int kBlock = N * LPU;
  static constexpr int kWaves = kBlock / kWave;
  static constexpr int kSystemsPerBlock = 1;
  // N = 128: LDS holds one system per CU; N = 96 (74 KB of matrix) two.
  static constexpr int kWavesPerEU = (N == 128) ? 1 : 2;
  static constexpr int kChunk = 16;
  static constexpr int kMode = kPlain;
  static constexpr bool kPerSystemM = PS;
  static constexpr int kItemsPerBlock = 1;
  static constexpr int kSystemsPerItem = 1;
  static constexpr int kLd = N + 1;
  static constexpr int kBS = N * LPU / 8;
  static constexpr int kNB = 8 / LPU;
  static constexpr int kBLd = kBS + 1;
  static constexpr int kSlots = (LPU == 1) ? 4 : 2;
  static_assert(kBS % 4 == 0 && kNB * kBS == N, "blocks of whole column quadruples");
  static constexpr size_t kStageDoubles = kLdsM ? 0 : (size_t)kSlots * kBS * kBLd + 8 * N;

  static constexpr size_t lds_bytes() {
    return (kLdsM ? sizeof(double) * N * (N + 1) : 0) + sizeof(double) * kStageDoubles +
           sizeof(double2) * N + sizeof(double) * N + sizeof(Sys) + sizeof(int) * BEX_NSTAT +
           sizeof(int) * 4;
  }

  int tid, u, h, sid_;
  double* sM;         // N <= 128: padded LDS copy
  const double* gM;   // N > 128: the system's M, column-major, global memory
  const double* gMT;  // N > 128: its transpose
  double* stage;      // N > 128: staged blocks of M, then the chain accumulators of mv_mtv
  double2* vec;
  double* red;
  Sys* sys;
  int* stat;
  int* flags;

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs& a) {
    tid = threadIdx.x;
    u = tid / LPU;
    h = tid % LPU;
    unsigned char* p = smem;
    sM = reinterpret_cast<double*>(p);
    if (kLdsM) p += sizeof(double) * N * (N + 1);
    stage = reinterpret_cast<double*>(p);
    p += sizeof(double) * kStageDoubles;
    vec = reinterpret_cast<double2*>(p);
    red = reinterpret_cast<double*>(vec + N);
    sys = reinterpret_cast<Sys*>(red + N);
    stat = reinterpret_cast<int*>(sys + 1);
    flags = stat + BEX_NSTAT;
    gM = a.CouplingMatrix;
    gMT = a.CouplingMatrixT;
    if (!PS && kLdsM) {
      load_padded<N>(sM, a.CouplingMatrix, tid, kBlock);
      __syncthreads();
    }
  }

  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    int s = -1;
    if (item < count) s = order ? order[item] : a.first_system + (int)item;
    sid_ = s;
    if (tid < BEX_NSTAT) stat[tid] = 0;
    if (PS && s >= 0 && s < a.NS) {
      if (kLdsM) {
        load_padded<N>(sM, a.CouplingMatrix + (size_t)s * N * N, tid, kBlock);
      } else {
        gM = a.CouplingMatrix + (size_t)s * N * N;
        gMT = a.CouplingMatrixT + (size_t)s * N * N;
      }
    }
    __syncthreads();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return u; }
  __device__ __forceinline__ bool primal() const { return h == 0; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }

  template <class Op>
  __device__ __forceinline__ double reduce(double v, Op op) const {
    if (h == 0) red[u] = v;
    __syncthreads();
    const int l = tid & (kWave - 1);
    const double w = group_tree<kWave>(fold_to_64<N>(red, l, op), op, 0);
    __syncthreads();
    return w;
  }
  __device__ __forceinline__ double sum(double v) const { return reduce(v, AddOp()); }
  __device__ __forceinline__ double fmax_all(double v) const { return reduce(v, FmaxOp()); }
  __device__ __forceinline__ double min_keep(double v) const { return reduce(v, MinKeepOp()); }
  __device__ __forceinline__ bool all(bool p) const {
    if (tid == 0) flags[0] = 0;
    __syncthreads();
    if (!p) flags[0] = 1;
    __syncthreads();
    const bool r = flags[0] == 0;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool any(bool p) const {
    if (tid == 0) flags[1] = 0;
    __syncthreads();
    if (p) flags[1] = 1;
    __syncthreads();
    const bool r = flags[1] != 0;
    __syncthreads();
    return r;
  }

  template <class MF, class ZF>
  __device__ __forceinline__ double row(MF m, ZF z) const {
    if (LPU == 2) return row_dot4_pair<N, kChunk>(m, z, h);
    return row_dot4<N, kChunk>(m, z);
  }
  __device__ __forceinline__ double mv() const {
    const double2* z = vec;
    const int i = u;
    if (kLdsM) {
      const double* m = sM;
      return row([&](int j) { return m[i + j * kLd]; }, [&](int j) { return z[j].x; });
    }
    const double* m = gM;
    return row([&](int j) { return m[i + (size_t)j * N]; }, [&](int j) { return z[j].x; });
  }
  __device__ __forceinline__ double mtv() const {
    const double2* z = vec;
    const int i = u;
    if (kLdsM) {
      const double* m = sM;
      return row([&](int j) { return m[j + i * kLd]; }, [&](int j) { return z[j].y; });
    }
    const double* mt = gMT;
    return row([&](int j) { return mt[i + (size_t)j * N]; }, [&](int j) { return z[j].y; });
  }
  // M z1 and M z2 for z1 on vec[].x and z2 on vec[].y, each row of M read once.
  __device__ __forceinline__ void mv_two(double* a1, double* a2) const {
    const double2* z = vec;
    const int i = u;
    auto z1 = [&](int j) { return z[j].x; };
    auto z2 = [&](int j) { return z[j].y; };
    if (kLdsM) {
      const double* m = sM;
      auto mf = [&](int j) { return m[i + j * kLd]; };
      if (LPU == 2) row_dot4_pair_two<N, kChunk>(mf, z1, z2, h, a1, a2);
      else row_dot4_two<N, kChunk>(mf, z1, z2, a1, a2);
      return;
    }
    const double* m = gM;
    auto mf = [&](int j) { return m[i + (size_t)j * N]; };
    if (LPU == 2) row_dot4_pair_two<N, kChunk>(mf, z1, z2, h, a1, a2);
    else row_dot4_two<N, kChunk>(mf, z1, z2, a1, a2);
  }

  // Sub-steps of the anti-diagonal pass: anti-diagonal t, staged row blocks i0 .. i1.
  __device__ __forceinline__ static int diag_lo(int t) { return t < kNB ? 0 : t - kNB + 1; }
  __device__ __forceinline__ static int diag_hi(int t) { return t < kNB ? t : kNB - 1; }
  static constexpr int kPer = kBS * kBS / 2;  // double2 per block
  static constexpr int kLoads = (kSlots * kPer + kBlock - 1) / kBlock;

  // Issues this thread's loads of the blocks (i, t - i), i = i0 .. i1, into registers. Slots
  // past the staged blocks load the last element again (same address as another lane's load)
  // instead of branching, so that all loads are in flight together.
  __device__ __forceinline__ void stage_load(int t, int i0, int i1, double2* v) const {
    const int n = (i1 - i0 + 1) * kPer;
#pragma unroll
    for (int r = 0; r < kLoads; r++) {
      const int q = min(tid + r * kBlock, n - 1);
      const int sl = q / kPer;
      const int e = q - sl * kPer;
      const int col = e / (kBS / 2);
      const int row = 2 * (e - col * (kBS / 2));
      const int I = i0 + sl;
      v[r] = *reinterpret_cast<const double2*>(gM + (size_t)((t - I) * kBS + col) * N +
                                               I * kBS + row);
    }
  }
  // Writes them to the LDS slots (padded column stride kBLd).
  __device__ __forceinline__ void stage_store(int i0, int i1, const double2* v) const {
    const int n = (i1 - i0 + 1) * kPer;
#pragma unroll
    for (int r = 0; r < kLoads; r++) {
      const int q = tid + r * kBlock;
      if (q < n) {
        const int sl = q / kPer;
        const int e = q - sl * kPer;
        const int col = e / (kBS / 2);
        const int row = 2 * (e - col * (kBS / 2));
        double* d = stage + (size_t)sl * kBS * kBLd + row + kBLd * col;
        d[0] = v[r].x;
        d[1] = v[r].y;
      }
    }
  }

  __device__ __forceinline__ void mv_mtv(double* a1, double* a2) const {
    if (kLdsM) {
      *a1 = mv();
      *a2 = mtv();
      return;
    }
    constexpr int kHalf = kBlock / 2;
    const bool rows = tid < kHalf;  // chains of M z; otherwise of M^T w
    const int k = rows ? tid : tid - kHalf;
    const int rho = k % kBS;
    const int c = k / kBS;
    double* acc = stage + (size_t)kSlots * kBS * kBLd + (rows ? 0 : 4 * N) + c * N + rho;
    for (int b = 0; b < kNB; b++) acc[b * kBS] = 0.0;
    // The next sub-step's blocks are loaded into registers while the current one is computed.
    double2 v[kLoads];
    int t = 0, i0 = 0;
    stage_load(0, 0, 0, v);
    while (t < 2 * kNB - 1) {
      const int i1 = min(diag_hi(t), i0 + kSlots - 1);
      __syncthreads();  // the previous blocks are read
      stage_store(i0, i1, v);
      __syncthreads();
      int tn = t, in = i1 + 1;
      if (in > diag_hi(t)) {
        tn = t + 1;
        in = diag_lo(tn);
      }
      if (tn < 2 * kNB - 1) stage_load(tn, in, min(diag_hi(tn), in + kSlots - 1), v);
      // Row chains meet block (b, t - b), staged for i0 <= b <= i1; column chains meet block
      // (t - b, b), staged for i0 <= t - b <= i1.
      const int blo = rows ? i0 : t - i1;
      const int bhi = rows ? i1 : t - i0;
      for (int b = blo; b <= bhi; b++) {
        const int I = rows ? b : t - b;
        const double* s = stage + (size_t)(I - i0) * kBS * kBLd;
        double sum = acc[b * kBS];
        if (rows) {
          const double2* z = vec + (t - b) * kBS;
#pragma unroll
          for (int m = 0; m < kBS / 4; m++) {
            const int j = c + 4 * m;
            sum = fma(s[rho + kBLd * j], z[j].x, sum);
          }
        } else {
          const double2* w = vec + I * kBS;
#pragma unroll
          for (int m = 0; m < kBS / 4; m++) {
            const int i = c + 4 * m;
            sum = fma(s[i + kBLd * rho], w[i].y, sum);
          }
        }
        acc[b * kBS] = sum;
      }
      t = tn;
      i0 = in;
    }
    __syncthreads();
    const double* pm = stage + (size_t)kSlots * kBS * kBLd + u;
    *a1 = (pm[0] + pm[N]) + (pm[2 * N] + pm[3 * N]);
    *a2 = (pm[4 * N] + pm[5 * N]) + (pm[6 * N] + pm[7 * N]);
    __syncthreads();  // every lane has read its sums before the accumulators are reused
  }

  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&&, bool) const {
    return group_coupling(*this, F, x1, cpt, rD, a);
  }
};

// =========================================================================================
// RolePairLayout32: one N = 32 system per wave, primal and shadow halves.


typedef short bex_v8s __attribute__((ext_vector_type(8)));
typedef __bf16 bex_v8bf __attribute__((ext_vector_type(8)));
typedef float bex_v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bex_v8bf pack_bf16x8(const float* v) {
  bex_v8s r;
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = (short)bex_bf16_bits(v[e]);
  return __builtin_bit_cast(bex_v8bf, r);
}
__device__ __forceinline__ bex_v8bf load_bf16x8(const uint64_t* w) {
  const uint4 q = *reinterpret_cast<const uint4*>(w);
  return __builtin_bit_cast(bex_v8bf, q);
}
__device__ __forceinline__ float relu(float z) { return z > 0.0f ? z : 0.0f; }

// Guess factor y of the warm-start network (definition in bex_mlp.h) for the lane's unit, on
// both halves of a role-pair wave. The primal half supplies the layer-1 inputs (five features,
// the bias input 1, two zeros), the shadow half zeros. Hidden layers run on
// v_mfma_f32_32x32x16_bf16 with the layer-1 accumulators feeding layer 2's B operand in place
// (2 + 8 MFMAs); the output layer is a VALU dot over the lane's 32 hidden values plus the
// partner half's 32.
__device__ __forceinline__ float mlp_factor(const BexMlpDevice* net, const double* wf, int lane) {
  const int h = lane >> 5;
  float in[8];
  in[0] = (float)wf[0];
  in[1] = (float)wf[1];
  in[2] = (float)wf[2];
  in[3] = (float)wf[3];
  in[4] = (float)wf[4];
  in[5] = 1.0f;
  in[6] = 0.0f;
  in[7] = 0.0f;
  if (h != 0) {
#pragma unroll
    for (int e = 0; e < 8; e++) in[e] = 0.0f;
  }
  const bex_v8bf bin = pack_bf16x8(in);
  bex_v16f zero;
#pragma unroll
  for (int r = 0; r < 16; r++) zero[r] = 0.0f;
  const bex_v8bf nil = __builtin_bit_cast(bex_v8bf, make_uint4(0u, 0u, 0u, 0u));
  bex_v16f d1[2], d2[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const bex_v8bf a1 = (h == 0) ? load_bf16x8(net->a1[t][lane & 31]) : nil;
    d1[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bin, zero, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; r++) d2[t][r] = net->b2[32 * t + 8 * (r >> 2) + 4 * h + (r & 3)];
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const bex_v16f& src = d1[s >> 1];
    const int q = 8 * (s & 1);
    float act[8];
#pragma unroll
    for (int e = 0; e < 8; e++) act[e] = relu(src[q + e]);
    const bex_v8bf bs = pack_bf16x8(act);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const bex_v8bf a2 = load_bf16x8(net->a2[t][s][lane]);
      d2[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, bs, d2[t], 0, 0, 0);
    }
  }
  float part = 0.0f;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int r = 0; r < 16; r++)
      part = fmaf(net->w3[32 * t + 8 * (r >> 2) + 4 * h + (r & 3)],
                  bex_bf16_round(relu(d2[t][r])), part);
  const float other = __shfl_xor(part, 32, kWave);
  const float p0 = (h == 0) ? part : other;
  const float p1 = (h == 0) ? other : part;
  return net->b3 + (p0 + p1);
}

// BiCG of the role-pair layouts, in which the primal half of a wave carries r, p, q, x of its
// units and the shadow half r~, p~, q~ in the same registers (one instruction stream serves
// A p and A^T p~). The layout L provides matvec(z) (M z on the primal half, M^T z on the shadow
// half, for a per-unit z), dot_halves(v) (sum over the units of primal(v) * shadow(v), on
// every lane), res_test(r, tol, &r.r~) (whether max |r| < tol, with the reference's NaN-ignoring
// max), res_max(r), uniform_any(p) and primal_half().

// Initial residual r0 = b - A x0 of the primal half, on both halves.
template <class L>
__device__ __forceinline__ double rp_residual(const L& lay, double b, double x0, double x1sq,
                                              double rD) {
  const double ax = fma(rD, lay.matvec(x1sq * x0), x0);
  double lo, up;
  halves(-ax + b, &lo, &up);
  return lo;
}

// BiCG for A x = b from x0 (Steppers.cuh:494-561 with x0 in place of xu): x on return is
// meaningful on the primal half. The loop is rotated: the reference's iteration-1 branch (the
// rho exit, p = r) runs before it and the direction update p = r + beta p of iteration k + 1
// closes iteration k, which leaves the arithmetic and its order unchanged. The vector updates
// x + alpha p, r - alpha q and r + beta p are single fmas (one rounding, as nvcc contracts
// them in the reference's kernel), in the oracle's canonical mode too. Scalars (rho,
// alpha, beta, max |r|) are identical in every lane, so the branches on them are uniform.
template <class L>
__device__ __forceinline__ BicgOut rp_bicg(const L& lay, double res0, double x0, double x1sq,
                                           double rD, const BexKernelArgs& a, double* xout) {
  const bool primal = lay.primal_half();
  double res = res0;                     // r on the primal half, r~ = r on the shadow half
  const double s1 = primal ? x1sq : rD;  // A p    = p  + rD   (M   (x1^2 p))
  const double s2 = primal ? rD : x1sq;  // A^T p~ = p~ + x1^2 (M^T (rD   p~))
  double x = x0;
  BicgOut out{0, false, true, bex::infinity()};
  const int maxit = a.BiCGMaxIterations;
  if (maxit >= 1) {
    double rho1 = lay.dot_halves(res);  // r . r~
    if (lay.uniform_any(bex::fabs_d(rho1) < a.BiCGRhoTolerance)) {
      out.rho_exit = true;
      out.maxit = false;
    } else {
      double dir = res;  // p (primal), p~ (shadow)
      for (int iter = 1;; iter++) {
        const double qv = fma(s2, lay.matvec(s1 * dir), dir);  // q (primal), q~ (shadow)
        out.iterations++;
        const double alpha = rho1 / lay.dot_halves(primal ? qv : dir);  // p~ . q
        x = fma(alpha, dir, x);
        res = fma(-alpha, qv, res);
        const double rho2 = rho1;
        // The residual test (max |r| < tolerance) and the next r . r~.
        if (lay.res_test(res, a.BiCGResidualTolerance, &rho1)) {
          out.maxit = false;
          break;
        }
        if (iter >= maxit) {
          out.rmax = lay.res_max(res);  // consulted by the learned start's restart rule
          break;
        }
        const double beta = rho1 / rho2;
        dir = fma(beta, dir, res);
      }
    }
  }
  *xout = x;
  return out;
}

template <int MODE, bool PS = false>
struct RolePairLayout32 {
  static constexpr int kUnits = 32;
  static constexpr int kSystemsPerWave = 1;
  static constexpr int kWaves = 1;
  static constexpr int kBlock = kWave;
  static constexpr int kSystemsPerBlock = 1;
  // The learned kernel needs ~70 more registers for the MFMA tiles: 2 waves per SIMD.
  static constexpr int kWavesPerEU = (MODE == kLearned) ? 2 : BEX_ROLEPAIR_WAVES_PER_EU;
  static constexpr int kMode = MODE;
  static constexpr bool kPerSystemM = PS;
  static constexpr int kItemsPerBlock = 1;
  static constexpr int kSystemsPerItem = 1;

  // The plain kernel keeps the unit parameters in LDS ([unit][NUP], one copy for both halves,
  // 5.4 KB per wave with the Keller-Miksis model): the right-hand side reads each once per
  // stage, and the ~42 registers they would hold across the time loop are what the stage
  // state otherwise spills for at 3 waves per SIMD.
  static constexpr bool kUnitParamsInLds = (MODE == kPlain);
  static constexpr int kNup = BexSystemDims::NUP > 0 ? BexSystemDims::NUP : 1;

  static constexpr size_t lds_bytes() {
    return (MODE == kLearned ? sizeof(BexMlpDevice) : 0) +
           (kUnitParamsInLds ? sizeof(double) * 32 * kNup : 0) + sizeof(Sys) +
           sizeof(int) * BEX_NSTAT;
  }

  int lane, sid_;
  bool primal_;
  bool learned_;    // learned kernel: this system takes the learned start (wave-uniform)
  double mrow[32];  // row i of M (primal) or of M^T (shadow)
  Sys* sys;
  int* stat;
  double* upar_lds;  // [32][NUP] when kUnitParamsInLds
  const BexMlpDevice* net;

  __device__ __forceinline__ double* unit_params() const { return upar_lds + (lane & 31) * kNup; }

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs& a) {
    lane = threadIdx.x;
    primal_ = lane < 32;
    unsigned char* p = smem;
    if (MODE == kLearned) {
      BexMlpDevice* n = reinterpret_cast<BexMlpDevice*>(p);
      const uint4* from = reinterpret_cast<const uint4*>(a.WarmStart);
      uint4* to = reinterpret_cast<uint4*>(n);
      for (int k = lane; k < (int)(sizeof(BexMlpDevice) / sizeof(uint4)); k += kWave) to[k] = from[k];
      net = n;
      p += sizeof(BexMlpDevice);
    }
    upar_lds = reinterpret_cast<double*>(p);
    if (kUnitParamsInLds) p += sizeof(double) * 32 * kNup;
    sys = reinterpret_cast<Sys*>(p);
    stat = reinterpret_cast<int*>(sys + 1);
  }

  // One work item = one system: first_system + item, or order[item]. Loads the lane's matrix
  // row (shared coupling 0 and its transpose, or the system's own matrix) into registers.
  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    int s = -1;
    if (item < count) s = order ? order[item] : a.first_system + (int)item;
    sid_ = s;
    const int take = (MODE == kLearned && s >= 0 && s < a.NS &&
                      (!a.WarmStartSystems || a.WarmStartSystems[s] != 0)) ? 1 : 0;
    learned_ = __builtin_amdgcn_readfirstlane(take) != 0;
    if (lane < BEX_NSTAT) stat[lane] = 0;
    const int i = lane & 31;
    if (s >= 0 && s < a.NS) {
      if (PS) {
        const double* m = a.CouplingMatrix + (size_t)s * 1024;
        if (primal_) {
#pragma unroll
          for (int j = 0; j < 32; j++) mrow[j] = m[i + j * 32];
        } else {
#pragma unroll
          for (int j = 0; j < 32; j++) mrow[j] = m[j + i * 32];
        }
      } else {
        const double* src = primal_ ? a.CouplingMatrix : a.CouplingMatrixT;
#pragma unroll
        for (int j = 0; j < 32; j++) mrow[j] = src[i + j * 32];
      }
    }
    wave_sync();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return lane & 31; }
  __device__ __forceinline__ bool primal() const { return primal_; }
  __device__ __forceinline__ bool primal_half() const { return primal_; }
  __device__ __forceinline__ void sync() const { wave_sync(); }

  // Both halves hold identical per-unit values, so wave-wide AND/OR/min equal the system's.
  __device__ __forceinline__ bool all(bool p) const { return __ballot(!p) == 0ull; }
  __device__ __forceinline__ bool any(bool p) const { return __ballot(p) != 0ull; }
  // Minimum over the system's units, broadcast as a wave-uniform value. Its operands (the
  // controller's tol/err ratios, the event step candidates) are positive and never NaN, where
  // v_min_f64 and the keep-on-NaN minimum agree.
  __device__ __forceinline__ double min_keep(double v) const {
    return uniform(tree32(v, VminOp()));
  }

  // acc_i = sum_j mrow[j] * z_j over j = 0..31 in the canonical order (four partial sums),
  // z taken from the lane's half.
  __device__ __forceinline__ double matvec(double z) const {
    double even, odd;
    rows_of_half(z, &even, &odd);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    fmac_row16(acc, even, &mrow[0]);
    fmac_row16(acc, odd, &mrow[16]);
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }

  // sum_i a_i b_i (tree order) for a on the primal half and b on the shadow half of v, in every
  // lane.
  __device__ __forceinline__ double dot_halves(double v) const {
    double lo, up;
    halves(v, &lo, &up);
    return tree32(lo * up, AddOp());
  }

  // r . r~ on every lane (the product of the two halves, then the tree, as p~ . q) and the
  // residual test max |r| < tol over the primal half. With the reference's NaN-ignoring max the
  // test holds exactly when no non-NaN |r_i| fails |r_i| < tol and some r_i is not NaN: two
  // ballots instead of a max tree.
  __device__ __forceinline__ bool res_test(double res, double tol, double* rho) const {
    *rho = dot_halves(res);
    // Lane masks straight from the compares (11: unordered or >=, 7: ordered); the primal half
    // is the low 32 lanes.
    const unsigned long long numbers = __builtin_amdgcn_fcmp(res, res, 7);
    const unsigned long long failing = __builtin_amdgcn_fcmp(bex::fabs_d(res), tol, 11) & numbers;
    return (unsigned)failing == 0u && (unsigned)numbers != 0u;
  }
  // max |r| over the primal half (NaN ignored), on every lane.
  __device__ __forceinline__ double res_max(double res) const {
    double lo2, up2;
    halves(res, &lo2, &up2);
    double re, ro;
    rows_of_half(lo2, &re, &ro);
    double m = VmaxAbsOp()(re, ro);
    m = VmaxOp()(m, row_ror<8>(m));
    m = VmaxOp()(m, row_ror<4>(m));
    m = VmaxOp()(m, row_ror<2>(m));
    m = VmaxOp()(m, row_ror<1>(m));
    return m;
  }
  // Branch on a predicate that has the same value in every lane.
  __device__ __forceinline__ bool uniform_any(bool p) const { return any(p); }

  __device__ __forceinline__ double residual(double b, double x0, double x1sq, double rD) const {
    return rp_residual(*this, b, x0, x1sq, rD);
  }
  __device__ __forceinline__ BicgOut bicg(double res0, double x0, double x1sq, double rD,
                                          const BexKernelArgs& a, double* xout) const {
    return rp_bicg(*this, res0, x0, x1sq, rD, a, xout);
  }

  // A x = b with A = I + diag(rD) M diag(x1^2), b = xu + rD (M CPT) (Steppers.cuh:489-561).
  // Returns x in F[1] on both halves. feat(crs, wf) evaluates the warm-start features of the
  // lane's unit given the coupling row sum crs (learned and sampling kernels only).
  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&& feat,
                                              bool sample_now) const {
    const double x1sq = x1 * x1;
    const double xu = F[1];
    const double b = fma(rD, matvec(cpt), xu);  // meaningful on the primal half
    double lo, up;
    double wf[BEX_MLP_IN];
    double bl = 0.0;
    const bool need_features = (MODE == kLearned && learned_) || (MODE == kSampling && sample_now);
    if (need_features) {
      halves(matvec(x1sq), &lo, &up);  // sum_j M_ij x1_j^2 of the lane's unit, both halves
      feat(lo, wf);
      halves(b, &bl, &up);             // b_i on both halves
    }
    double x0 = xu;
    double r0 = residual(b, xu, x1sq, rD);
    bool learned_start = false;  // wave-uniform
    double mu = 0.0;             // max |r0| of the plain start (learned kernel)
    if (MODE == kLearned && learned_) {
      // The learned start is used only when its initial residual is smaller (max norm) than
      // the plain start's, so BiCG never starts further from the solution than the reference.
      const float y = mlp_factor(net, wf, lane);
      double xl = xu;
      if (bex::is_finite((double)y) && y <= BEX_MLP_TARGET_CLAMP && y >= -BEX_MLP_TARGET_CLAMP)
        xl = bl * (double)y;
      const double rl = residual(b, xl, x1sq, rD);
      double ml = bex::fabs_d(rl);
      mu = bex::fabs_d(r0);
      tree32x2(mu, ml, VmaxOp(), VmaxOp());
      if (any(ml < mu)) {
        learned_start = true;
        x0 = xl;
        r0 = rl;
      }
    }
    double x;
    BicgOut out = bicg(r0, x0, x1sq, rD, a, &x);
    // A learned start that ends unconverged (iteration cap) and no closer than the plain start
    // began, or non-finite, is discarded: plain restart from the uncoupled acceleration.
    if (MODE == kLearned && learned_start && out.maxit && any(!(out.rmax < mu))) {
      const BicgOut again = bicg(residual(b, xu, x1sq, rD), xu, x1sq, rD, a, &x);
      if (lane == 0) stat[BEX_STAT_WARM_FALLBACKS]++;
      out.iterations += again.iterations;
      out.rho_exit = again.rho_exit;
      out.maxit = again.maxit;
    }
    halves(x, &lo, &up);
    F[1] = lo;
    if (MODE == kSampling && sample_now) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(a.SampleCount, 32ull);
      base = __builtin_amdgcn_readfirstlane((unsigned)base) |
             ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32)) << 32);
      if (primal_ && base + 32 <= (unsigned long long)a.SampleCapacity) {
        float* row = a.Samples + (base + (unsigned long long)lane) * BEX_SAMPLE_FLOATS;
        double t = lo / bl;
        t = (t > BEX_MLP_TARGET_CLAMP) ? BEX_MLP_TARGET_CLAMP : t;
        t = (t < -BEX_MLP_TARGET_CLAMP) ? -BEX_MLP_TARGET_CLAMP : t;
#pragma unroll
        for (int k = 0; k < BEX_MLP_IN; k++) row[k] = (float)wf[k];
        row[5] = bex::is_finite(t) ? (float)t : 1.0f;
        row[6] = (float)lo;
        row[7] = (float)bl;
        row[8] = (float)xu;
        row[9] = 0.0f;
      }
    }
    return out;
  }
};

// =========================================================================================
// PairLayout32: two N = 32 systems per wave, lane = unit. Lanes 0-31 serve one system, lanes
// 32-63 the next, and every lane holds row i of M and row i of M^T in registers (128 VGPRs; the
// coupling matrix all systems share, as in the reference). A product M z runs as in the role
// pair (two v_permlane16_swap and 32 v_fmac_f64 with a DPP row_newbcast operand, which never
// cross a 32-lane half), so one pass computes M z for both systems and a second pass M^T w; the
// halving trees, the divisions, the right-hand side, the controller and the event logic run once
// for the two systems, where the role pair runs them once per system. The two systems step
// independently: every system-level decision is taken per half and their branches diverge
// through the exec mask, so a wave costs the slower of its two systems (systems at adjacent
// grid points cost within about 1 % of each other). The BiCG arithmetic is group_coupling's,
// i.e. the canonical order; the unit parameters live in LDS ([lane][NUP]).
struct PairLayout32 {
  static constexpr int kUnits = 32;
  static constexpr int kSystemsPerWave = 2;
  static constexpr int kWaves = 1;
  static constexpr int kBlock = kWave;
  static constexpr int kSystemsPerBlock = 2;
  static constexpr int kWavesPerEU = 2;  // 256 VGPRs: the two matrix rows take half of them
  static constexpr int kMode = kPlain;
  static constexpr bool kPerSystemM = false;
  static constexpr int kItemsPerBlock = 1;
  static constexpr int kSystemsPerItem = 2;
  static constexpr bool kUnitParamsInLds = true;
  static constexpr int kNup = BexSystemDims::NUP > 0 ? BexSystemDims::NUP : 1;

  static constexpr size_t lds_bytes() {
    return sizeof(double) * kWave * kNup + 2 * sizeof(Sys) + 2 * sizeof(int) * BEX_NSTAT;
  }

  int lane, grp, sid_;
  double mrow[32];   // row i of M
  double mtrow[32];  // row i of M^T
  Sys* sys;
  int* stat;
  double* upar_lds;

  __device__ __forceinline__ double* unit_params() const { return upar_lds + lane * kNup; }

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs&) {
    lane = threadIdx.x;
    grp = lane >> 5;
    upar_lds = reinterpret_cast<double*>(smem);
    Sys* s = reinterpret_cast<Sys*>(smem + sizeof(double) * kWave * kNup);
    sys = s + grp;
    stat = reinterpret_cast<int*>(s + 2) + grp * BEX_NSTAT;
  }

  // One work item = two systems: first_system + 2 item + half, or the order array's entries
  // 2 item and 2 item + 1. Loads the lane's rows of M and M^T and clears the half's counters.
  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    const long long pos = item * 2 + grp;
    int s = -1;
    if (pos < count) s = order ? order[pos] : a.first_system + (int)pos;
    sid_ = s;
    const int i = lane & 31;
    if (i < BEX_NSTAT) stat[i] = 0;
    if (s >= 0 && s < a.NS) {
#pragma unroll
      for (int j = 0; j < 32; j++) {
        mrow[j] = a.CouplingMatrix[i + j * 32];
        mtrow[j] = a.CouplingMatrixT[i + j * 32];
      }
    }
    wave_sync();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return lane & 31; }
  __device__ __forceinline__ bool primal() const { return true; }
  __device__ __forceinline__ void sync() const { wave_sync(); }

  // Every lane of a half gets its half's value of a per-half condition: lo_true / hi_true are
  // the two halves' decisions, taken on the scalar unit from lane masks, and the resulting lane
  // mask is the predicate itself (no vector instruction).
  __device__ __forceinline__ static bool per_half(bool lo_true, bool hi_true) {
    const unsigned long long m = (lo_true ? 0x00000000ffffffffull : 0ull) |
                                 (hi_true ? 0xffffffff00000000ull : 0ull);
    return __builtin_amdgcn_inverse_ballot_w64(m);
  }
  // The two 32-bit halves of a lane mask as scalar words (the empty asm keeps the compiler from
  // turning a test of the upper word into a 64-bit vector compare).
  __device__ __forceinline__ static void words(unsigned long long m, unsigned* lo, unsigned* hi) {
    unsigned l = (unsigned)m, h = (unsigned)(m >> 32);
    asm("" : "+s"(l), "+s"(h));
    *lo = l;
    *hi = h;
  }
  // A half is either wholly active or wholly inactive, so the ballot over its 32 lanes is the
  // system's.
  __device__ __forceinline__ bool all(bool p) const {
    unsigned lo, hi;
    words(__ballot(p), &lo, &hi);
    return per_half(lo == 0xffffffffu, hi == 0xffffffffu);
  }
  __device__ __forceinline__ bool any(bool p) const {
    unsigned lo, hi;
    words(__ballot(p), &lo, &hi);
    return per_half(lo != 0u, hi != 0u);
  }
  // Halving-tree reductions of the half (Steppers.cuh:733-742) as butterflies: every lane of a
  // half ends with its system's value.
  __device__ __forceinline__ double sum(double v) const { return tree32(v, AddOp()); }
  // Minimum over the system's units (positive, non-NaN operands: the controller's tol/err
  // ratios and the event step candidates, as in the role pair).
  __device__ __forceinline__ double min_keep(double v) const { return tree32(v, VminOp()); }

  // sum_j row[j] z_j over the lane's half in the canonical order (four partial sums).
  __device__ __forceinline__ static double row_product(double z, const double* row) {
    double even, odd;
    rows_of_half(z, &even, &odd);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    fmac_row16(acc, even, &row[0]);
    fmac_row16(acc, odd, &row[16]);
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }

  // max |r| < tol over the half, NaN ignored (the reference's NaN-ignoring max), from two
  // compare masks as in the role pair.
  __device__ __forceinline__ bool res_test(double res, double tol) const {
    const unsigned long long numbers = __builtin_amdgcn_fcmp(res, res, 7);
    const unsigned long long failing = __builtin_amdgcn_fcmp(bex::fabs_d(res), tol, 11) & numbers;
    unsigned flo, fhi, nlo, nhi;
    words(failing, &flo, &fhi);
    words(numbers, &nlo, &nhi);
    return per_half(flo == 0u && nlo != 0u, fhi == 0u && nhi != 0u);
  }

  // A x = b with A = I + diag(rD) M diag(x1^2), b = xu + rD (M CPT), x0 = xu
  // (Steppers.cuh:489-561), group_coupling's arithmetic with the loop rotated as in rp_bicg.
  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&&, bool) const {
    const double x1sq = x1 * x1;
    const double xu = F[1];
    const double mc = row_product(cpt, mrow);
    const double mx = row_product(x1sq * xu, mrow);
    const double b = fma(rD, mc, xu);
    const double ax = fma(rD, mx, xu);
    double r = -ax + b;
    double rt = r;
    double x = xu;
    BicgOut out{0, false, true, 0.0};
    double rho1 = sum(r * rt);
    const int maxit = a.BiCGMaxIterations;
    if (maxit >= 1) {
      if (bex::fabs_d(rho1) < a.BiCGRhoTolerance) {
        out.rho_exit = true;
        out.maxit = false;
      } else {
        double p = r, pt = rt;
        for (int iter = 1;; iter++) {
          const double a1 = row_product(x1sq * p, mrow);
          const double a2 = row_product(rD * pt, mtrow);
          const double q = fma(rD, a1, p);
          const double qt = fma(x1sq, a2, pt);
          out.iterations++;
          const double alpha = rho1 / sum(pt * q);
          x = fma(alpha, p, x);
          r = fma(-alpha, q, r);
          rt = fma(-alpha, qt, rt);
          const double rho2 = rho1;
          const bool converged = res_test(r, a.BiCGResidualTolerance);
          rho1 = sum(r * rt);
          if (converged) {
            out.maxit = false;
            break;
          }
          if (iter >= maxit) break;
          const double beta = rho1 / rho2;
          p = fma(beta, p, r);
          pt = fma(beta, pt, rt);
        }
      }
    }
    F[1] = x;
    return out;
  }
};

// =========================================================================================
// ChainLayout128: one N = 128 system per workgroup of eight waves, the coupling matrix in
// registers. Wave w serves the 32 units of block B = w / 2 as a role pair (primal half: M and
// r, p, q, x; shadow half: M^T and r~, p~, q~) and runs two of the four FMA chains of every row
// product, chains 2 (w % 2) and 2 (w % 2) + 1, over all 128 columns: lane k of the half holds
// row u = 32 B + k of M (or M^T) at the chains' 64 columns, 128 VGPRs. A product's operand is
// redistributed once through LDS so that lane k of every half holds z_(4k + q) of its two
// chains; each chain then runs as 32 v_fmac_f64 with that operand broadcast by DPP
// row_newbcast from lane t, in column order. The two chain pairs of a row meet in LDS and are
// added as (s0 + s1) + (s2 + s3), the canonical order; reductions over the 128 units take their
// first two halving levels (h = 64, 32) from LDS and the rest in-wave. Every wave evaluates the
// BiCG scalars identically, so its branches agree across the workgroup without a barrier.
// Compared with M in LDS (WideLayout<128>), a BiCG iteration reads ~12 KB of LDS instead of
// ~512 KB.

__device__ __forceinline__ void fmac2_row8_lo(double& sa, double& sb, double za, double zb,
                                             const double* ma, const double* mb) {
  asm volatile(
      "s_nop 1\n"
      "v_fmac_f64_dpp %0, %2, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %12 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %13 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %14 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %15 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %16 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %10 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %18 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %11 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %19 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
      : "+v"(sa), "+v"(sb)
      : "v"(za), "v"(zb), "v"(ma[0]), "v"(ma[1]), "v"(ma[2]), "v"(ma[3]), "v"(ma[4]), "v"(ma[5]), "v"(ma[6]), "v"(ma[7]),
        "v"(mb[0]), "v"(mb[1]), "v"(mb[2]), "v"(mb[3]), "v"(mb[4]), "v"(mb[5]), "v"(mb[6]), "v"(mb[7]));
}

__device__ __forceinline__ void fmac2_row8_hi(double& sa, double& sb, double za, double zb,
                                             const double* ma, const double* mb) {
  asm volatile(
      "s_nop 1\n"
      "v_fmac_f64_dpp %0, %2, %4 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %12 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %5 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %13 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %6 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %14 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %7 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %15 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %8 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %16 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %9 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %10 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %18 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %0, %2, %11 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %3, %19 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
      : "+v"(sa), "+v"(sb)
      : "v"(za), "v"(zb), "v"(ma[0]), "v"(ma[1]), "v"(ma[2]), "v"(ma[3]), "v"(ma[4]), "v"(ma[5]), "v"(ma[6]), "v"(ma[7]),
        "v"(mb[0]), "v"(mb[1]), "v"(mb[2]), "v"(mb[3]), "v"(mb[4]), "v"(mb[5]), "v"(mb[6]), "v"(mb[7]));
}

template <bool PS = true>
struct ChainLayout128 {
  static constexpr int N = 128;
  static constexpr int kUnits = N;
  static constexpr int kWaves = 8;
  static constexpr int kBlock = kWaves * kWave;
  static constexpr int kSystemsPerBlock = 1;
  static constexpr int kWavesPerEU = 2;  // 128 VGPRs of matrix per lane: 256-register budget
  static constexpr int kMode = kPlain;
  static constexpr bool kPerSystemM = PS;
  static constexpr int kItemsPerBlock = 1;
  static constexpr int kSystemsPerItem = 1;

  static constexpr size_t lds_bytes() {
    return sizeof(double) * (2 * N + 8 * 32 * 2 + N + N) + sizeof(double2) * N + sizeof(Sys) +
           sizeof(int) * BEX_NSTAT + sizeof(int) * 4;
  }

  int tid, hh, k, blk, qp, u, sid_;
  double ma[32], mb[32];  // row u of M (primal) or of M^T (shadow) at columns 4t + 2 qp (+ 1)
  double* zbuf;           // [2][N]: product operands, primal then shadow
  double* cc;             // [4 blocks][2 halves][32][2 chain pairs]
  double* red;            // [32][4 blocks]
  double2* red2;          // [32][4 blocks]
  double* redm;           // [32][4 blocks]
  Sys* sys;
  int* stat;
  int* flags;

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs&) {
    tid = threadIdx.x;
    const int w = tid / kWave, l = tid & (kWave - 1);
    hh = l >> 5;
    k = l & 31;
    blk = w >> 1;
    qp = w & 1;
    u = 32 * blk + k;
    double* p = reinterpret_cast<double*>(smem);
    zbuf = p;
    cc = zbuf + 2 * N;
    red = cc + 8 * 32 * 2;
    redm = red + N;
    red2 = reinterpret_cast<double2*>(redm + N);
    sys = reinterpret_cast<Sys*>(red2 + N);
    stat = reinterpret_cast<int*>(sys + 1);
    flags = stat + BEX_NSTAT;
  }

  // Work item = one system. Loads the lane's 64 matrix entries, element (u, j) of M (primal) or
  // of M^T (shadow). A shared matrix comes with its transpose (both column-major: coalesced
  // across the lanes of a half); per-system matrices come without one, so the shadow half reads
  // column u of M itself (strided, once per system).
  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    int s = -1;
    if (item < count) s = order ? order[item] : a.first_system + (int)item;
    sid_ = s;
    if (tid < BEX_NSTAT) stat[tid] = 0;
    if (s >= 0 && s < a.NS) {
      const double* m;
      size_t su, sj;  // strides of u and j
      if (PS) {
        m = a.CouplingMatrix + (size_t)s * N * N;
        su = hh == 0 ? 1 : N;
        sj = hh == 0 ? N : 1;
      } else {
        m = hh == 0 ? a.CouplingMatrix : a.CouplingMatrixT;
        su = 1;
        sj = N;
      }
      const double* src = m + (size_t)u * su;
#pragma unroll
      for (int t = 0; t < 32; t++) {
        ma[t] = src[(size_t)(4 * t + 2 * qp) * sj];
        mb[t] = src[(size_t)(4 * t + 2 * qp + 1) * sj];
      }
    }
    __syncthreads();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return u; }
  __device__ __forceinline__ bool primal() const { return hh == 0 && qp == 0; }
  __device__ __forceinline__ bool primal_half() const { return hh == 0; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }

  __device__ __forceinline__ bool all(bool p) const {
    if (tid == 0) flags[0] = 0;
    __syncthreads();
    if (!p) flags[0] = 1;
    __syncthreads();
    const bool r = flags[0] == 0;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool any(bool p) const {
    if (tid == 0) flags[1] = 0;
    __syncthreads();
    if (p) flags[1] = 1;
    __syncthreads();
    const bool r = flags[1] != 0;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool uniform_any(bool p) const { return __ballot(p) != 0ull; }

  // Minimum over the units (positive, non-NaN operands: see RolePairLayout32::min_keep).
  __device__ __forceinline__ double min_keep(double v) const {
    if (primal()) redm[k * 4 + blk] = v;
    __syncthreads();
    const double y = VminOp()(VminOp()(redm[k * 4], redm[k * 4 + 2]),
                              VminOp()(redm[k * 4 + 1], redm[k * 4 + 3]));
    const double r = uniform(tree32(y, VminOp()));
    __syncthreads();
    return r;
  }

  // M z (primal half) and M^T z (shadow half) for the per-unit operand z, canonical order.
  __device__ __forceinline__ double matvec(double z) const {
    if (qp == 0) zbuf[hh * N + u] = z;
    __syncthreads();
    const double2 zz = reinterpret_cast<const double2*>(zbuf + hh * N)[2 * k + qp];
    double ea, oa, eb, ob;
    rows_of_half(zz.x, &ea, &oa);  // z_(4t + 2qp) from lane t of the half
    rows_of_half(zz.y, &eb, &ob);  // z_(4t + 2qp + 1)
    double sa = 0.0, sb = 0.0;
    fmac2_row8_lo(sa, sb, ea, eb, &ma[0], &mb[0]);
    fmac2_row8_hi(sa, sb, ea, eb, &ma[8], &mb[8]);
    fmac2_row8_lo(sa, sb, oa, ob, &ma[16], &mb[16]);
    fmac2_row8_hi(sa, sb, oa, ob, &ma[24], &mb[24]);
    cc[((blk * 2 + hh) * 32 + k) * 2 + qp] = sa + sb;  // s0 + s1, or s2 + s3
    __syncthreads();
    const double2 both = reinterpret_cast<const double2*>(cc)[(blk * 2 + hh) * 32 + k];
    return both.x + both.y;
  }

  // Sum over the 128 units of primal(v) * shadow(v): levels h = 64, 32 from the four blocks'
  // products in LDS, then the in-wave tree; on every lane.
  __device__ __forceinline__ double dot_halves(double v) const {
    double lo, up;
    halves(v, &lo, &up);
    if (primal()) red[k * 4 + blk] = lo * up;
    __syncthreads();
    const double2 v01 = reinterpret_cast<const double2*>(red)[2 * k];
    const double2 v23 = reinterpret_cast<const double2*>(red)[2 * k + 1];
    return tree32((v01.x + v23.x) + (v01.y + v23.y), AddOp());
  }

  // max |r| and r . r~ over the 128 units, on every lane.
  __device__ __forceinline__ void res_trees(double res, double* rmax, double* rho) const {
    double lo2, up2;
    halves(res, &lo2, &up2);
    if (primal()) red2[k * 4 + blk] = make_double2(bex::fabs_d(lo2), lo2 * up2);
    __syncthreads();
    const double2 w0 = red2[k * 4], w1 = red2[k * 4 + 1], w2 = red2[k * 4 + 2],
                  w3 = red2[k * 4 + 3];
    double m = VmaxOp()(VmaxOp()(w0.x, w2.x), VmaxOp()(w1.x, w3.x));
    double p = (w0.y + w2.y) + (w1.y + w3.y);
    tree32x2(m, p, VmaxOp(), AddOp());
    *rmax = m;
    *rho = p;
  }
  __device__ __forceinline__ bool res_test(double res, double tol, double* rho) const {
    double m;
    res_trees(res, &m, rho);
    return uniform_any(m < tol);
  }
  __device__ __forceinline__ double res_max(double res) const {
    double m, p;
    res_trees(res, &m, &p);
    return m;
  }

  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&&, bool) const {
    const double x1sq = x1 * x1;
    const double xu = F[1];
    const double b = fma(rD, matvec(cpt), xu);  // meaningful on the primal half
    const double r0 = rp_residual(*this, b, xu, x1sq, rD);
    double x;
    const BicgOut out = rp_bicg(*this, r0, xu, x1sq, rD, a, &x);
    double lo, up;
    halves(x, &lo, &up);
    F[1] = lo;
    return out;
  }
};

// =========================================================================================
// RolePairLayout64: one N = 64 system per workgroup of two waves, the coupling matrix in
// registers. Wave B serves units 32 B .. 32 B + 31 as a role pair (primal half: row u of M and
// r, p, q, x; shadow half: row u of M^T and r~, p~, q~), each lane holding its row's 64
// entries (128 VGPRs) and running all four FMA chains of its row product like
// RolePairLayout32. A product's operand passes through LDS once, so that lane k of each half
// holds z_k and z_(32 + k) for the DPP row_newbcast broadcasts (two LDS buffers used in turn:
// the barrier of one product orders the next product's stores after the previous reads);
// reductions over the 64 units take their first halving level (h = 32, across the two waves)
// from LDS and the rest in-wave.
template <bool PS = true>
struct RolePairLayout64 {
  static constexpr int N = 64;
  static constexpr int kUnits = N;
  static constexpr int kWaves = 2;
  static constexpr int kBlock = kWaves * kWave;
  static constexpr int kSystemsPerBlock = 1;
  static constexpr int kWavesPerEU = 2;  // 128 VGPRs of matrix per lane: 256-register budget
  static constexpr int kMode = kPlain;
  static constexpr bool kPerSystemM = PS;
  static constexpr int kItemsPerBlock = 1;
  static constexpr int kSystemsPerItem = 1;

  static constexpr size_t lds_bytes() {
    return sizeof(double) * (2 * 2 * N + N + N) + sizeof(double2) * N + sizeof(Sys) +
           sizeof(int) * BEX_NSTAT + sizeof(int) * 4;
  }

  int tid, hh, k, blk, u, sid_;
  mutable int zsel;       // LDS operand buffer of the next product
  double mrow[64];        // row u of M (primal) or of M^T (shadow)
  double* zbuf;           // [2 buffers][2 halves][32][2]: (z_k, z_(32+k))
  double* red;            // [32][2 blocks]
  double* redm;           // [32][2 blocks]
  double2* red2;          // [32][2 blocks]
  Sys* sys;
  int* stat;
  int* flags;

  __device__ __forceinline__ void init(unsigned char* smem, const BexKernelArgs&) {
    tid = threadIdx.x;
    const int l = tid & (kWave - 1);
    blk = tid / kWave;
    hh = l >> 5;
    k = l & 31;
    u = 32 * blk + k;
    zsel = 0;
    double* p = reinterpret_cast<double*>(smem);
    zbuf = p;
    red = zbuf + 2 * 2 * N;
    redm = red + N;
    red2 = reinterpret_cast<double2*>(redm + N);
    sys = reinterpret_cast<Sys*>(red2 + N);
    stat = reinterpret_cast<int*>(sys + 1);
    flags = stat + BEX_NSTAT;
  }

  // Work item = one system: loads element (u, j) of M (primal) or M^T (shadow), j = 0..63; a
  // per-system matrix has no stored transpose, so the shadow half reads column u of M.
  __device__ __forceinline__ void bind(const BexKernelArgs& a, const int* order, long long item,
                                       long long count) {
    int s = -1;
    if (item < count) s = order ? order[item] : a.first_system + (int)item;
    sid_ = s;
    if (tid < BEX_NSTAT) stat[tid] = 0;
    if (s >= 0 && s < a.NS) {
      const double* m;
      size_t su, sj;  // strides of u and j
      if (PS) {
        m = a.CouplingMatrix + (size_t)s * N * N;
        su = hh == 0 ? 1 : N;
        sj = hh == 0 ? N : 1;
      } else {
        m = hh == 0 ? a.CouplingMatrix : a.CouplingMatrixT;
        su = 1;
        sj = N;
      }
      const double* src = m + (size_t)u * su;
#pragma unroll
      for (int j = 0; j < 64; j++) mrow[j] = src[(size_t)j * sj];
    }
    __syncthreads();
  }

  __device__ __forceinline__ int sid() const { return sid_; }
  __device__ __forceinline__ int uid() const { return u; }
  __device__ __forceinline__ bool primal() const { return hh == 0; }
  __device__ __forceinline__ bool primal_half() const { return hh == 0; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }

  __device__ __forceinline__ bool all(bool p) const {
    if (tid == 0) flags[0] = 0;
    __syncthreads();
    if (!p) flags[0] = 1;
    __syncthreads();
    const bool r = flags[0] == 0;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool any(bool p) const {
    if (tid == 0) flags[1] = 0;
    __syncthreads();
    if (p) flags[1] = 1;
    __syncthreads();
    const bool r = flags[1] != 0;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool uniform_any(bool p) const { return __ballot(p) != 0ull; }

  // Minimum over the units (positive, non-NaN operands: see RolePairLayout32::min_keep).
  __device__ __forceinline__ double min_keep(double v) const {
    if (primal()) redm[k * 2 + blk] = v;
    __syncthreads();
    const double r = uniform(tree32(VminOp()(redm[k * 2], redm[k * 2 + 1]), VminOp()));
    __syncthreads();
    return r;
  }

  // M z (primal half) and M^T z (shadow half) for the per-unit operand z, canonical order.
  __device__ __forceinline__ double matvec(double z) const {
    double* zb = zbuf + zsel * (2 * N);
    zsel ^= 1;
    zb[(hh * 32 + k) * 2 + blk] = z;
    __syncthreads();
    const double2 zz = reinterpret_cast<const double2*>(zb)[hh * 32 + k];
    double e0, o0, e1, o1;
    rows_of_half(zz.x, &e0, &o0);  // z_c, z_(16+c)
    rows_of_half(zz.y, &e1, &o1);  // z_(32+c), z_(48+c)
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    fmac_row16(acc, e0, &mrow[0]);
    fmac_row16(acc, o0, &mrow[16]);
    fmac_row16(acc, e1, &mrow[32]);
    fmac_row16(acc, o1, &mrow[48]);
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }

  // Sum over the 64 units of primal(v) * shadow(v): level h = 32 across the two waves in LDS,
  // then the in-wave tree; on every lane.
  __device__ __forceinline__ double dot_halves(double v) const {
    double lo, up;
    halves(v, &lo, &up);
    if (primal()) red[k * 2 + blk] = lo * up;
    __syncthreads();
    const double2 v01 = reinterpret_cast<const double2*>(red)[k];
    return tree32(v01.x + v01.y, AddOp());
  }

  // max |r| and r . r~ over the 64 units, on every lane.
  __device__ __forceinline__ void res_trees(double res, double* rmax, double* rho) const {
    double lo2, up2;
    halves(res, &lo2, &up2);
    if (primal()) red2[k * 2 + blk] = make_double2(bex::fabs_d(lo2), lo2 * up2);
    __syncthreads();
    const double2 w0 = red2[k * 2], w1 = red2[k * 2 + 1];
    double m = VmaxOp()(w0.x, w1.x);
    double p = w0.y + w1.y;
    tree32x2(m, p, VmaxOp(), AddOp());
    *rmax = m;
    *rho = p;
  }
  __device__ __forceinline__ bool res_test(double res, double tol, double* rho) const {
    double m;
    res_trees(res, &m, rho);
    return uniform_any(m < tol);
  }
  __device__ __forceinline__ double res_max(double res) const {
    double m, p;
    res_trees(res, &m, &p);
    return m;
  }

  template <class Feat>
  __device__ __forceinline__ BicgOut coupling(double* F, double x1, double cpt, double rD,
                                              const BexKernelArgs& a, Feat&&, bool) const {
    const double x1sq = x1 * x1;
    const double xu = F[1];
    const double b = fma(rD, matvec(cpt), xu);  // meaningful on the primal half
    const double r0 = rp_residual(*this, b, xu, x1sq, rD);
    double x;
    const BicgOut out = rp_bicg(*this, r0, xu, x1sq, rD, a, &x);
    double lo, up;
    halves(x, &lo, &up);
    F[1] = lo;
    return out;
  }
};

// Cash-Karp tableau (Steppers.cuh:197-370): row s holds a_s1..a_s5 and c_s. Read through
// rkck45_row(s, dep): the row's address passes through an empty asm tied to dep, so the scalar
// loads happen at the stage (not hoisted out of the time loop into scalar registers, where the
// ~20 constants would compete with the solver's uniform values).
static __constant__ const double kRkck45[6][8] = {
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {1.0 / 5.0, 0.0, 0.0, 0.0, 0.0, 1.0 / 5.0, 0.0, 0.0},
    {3.0 / 40.0, 9.0 / 40.0, 0.0, 0.0, 0.0, 3.0 / 10.0, 0.0, 0.0},
    {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0.0, 0.0, 3.0 / 5.0, 0.0, 0.0},
    {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0.0, 1.0, 0.0, 0.0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0,
     7.0 / 8.0, 0.0, 0.0}};

typedef const __attribute__((address_space(4))) double* ConstRow;

__device__ __forceinline__ ConstRow rkck45_row(int s, double dep) {
  ConstRow p = (ConstRow)&kRkck45[0][0];
  asm("" : "+s"(p) : "v"(dep));
  return p + 8 * s;
}

template <int UD>
__device__ __forceinline__ void store_stage(double (&k)[UD], const double* F) {
#pragma unroll
  for (int i = 0; i < UD; i++) k[i] = F[i];
}

// Adds v to the system's counter k in LDS. An LDS atomic whose result is unused needs no wait,
// where a read-modify-write would stall the wave for the LDS round trip.
__device__ __forceinline__ void count(int* st, int k, int v) { atomicAdd(&st[k], v); }

// =========================================================================================
// Time loop of one system (Solver.cuh:11-651), shared by the layouts.

template <class L, class = void>
struct UnitParamsInLds {
  static constexpr bool value = false;
};
template <class L>
struct UnitParamsInLds<L, decltype(void(L::kUnitParamsInLds))> {
  static constexpr bool value = L::kUnitParamsInLds;
};
template <class L>
__device__ __forceinline__ constexpr bool unit_params_in_lds() {
  return UnitParamsInLds<L>::value;
}

// Whether the system definition declares that its two action hooks write only the unit's
// accessories and no other hook reads them (BexSystemDims::kActionsWriteOnlyAccessories; absent
// means no).
template <class D, class = void>
struct ActionsWriteOnlyAccessories {
  static constexpr bool value = false;
};
template <class D>
struct ActionsWriteOnlyAccessories<D, decltype(void(D::kActionsWriteOnlyAccessories))> {
  static constexpr bool value = D::kActionsWriteOnlyAccessories;
};

// Dense-output variants, compiled as separate kernels so that each carries only its own mode's
// registers: none (NDO = 0), the reference's step-point sampling, Hermite interpolation.
enum DenseMode { kDenseNone = 0, kDenseSteps = 1, kDenseHermite = 2 };

template <class L, int ALG, int DENSE>
__device__ __forceinline__ void integrate(L& lay, const BexKernelArgs& a) {
  using D = BexSystemDims;
  constexpr int N = L::kUnits;
  constexpr int UD = D::UD;
  constexpr int NUP = D::NUP;
  constexpr int NUA = D::NUA;
  constexpr int NiUA = D::NiUA;
  constexpr int NE = D::NE;
  static_assert(UD == 2, "the implicit coupling acts on component 1 of a two-component unit");

  const int sid = lay.sid();
  const bool active = (sid >= 0) && (sid < a.NS) && (sid < a.ActiveSystems);  // Solver.cuh:82,132
  if (!active) return;

  Sys& S = *lay.sys;
  int* st = lay.stat;
  const int uid = lay.uid();
  const bool primal = lay.primal();
  const bool leader = primal && uid == 0;
  // Lanes that duplicate another lane's unit (the shadow half of a role pair, the second lane of
  // a lane pair) skip the action hooks when those only write accessories nobody else reads:
  // their accessory copies are never written back, and each skipped hook saves a private-memory
  // access of the unit's accessory array.
  const bool run_actions = primal || !ActionsWriteOnlyAccessories<D>::value;
  const long long TLT = a.TLT;
  const long long gtid = (long long)(sid / a.ref_spb) * a.ref_block_stride +
                         (long long)(sid % a.ref_spb) * N + uid;  // Interface.cuh:875-879
  const int NS = a.NS;
  if (leader && a.Timeline) a.Timeline[2 * (long long)sid] = (long long)wall_clock64();

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
    // The host keeps the index in [0, NDO]; anything outside is treated as a full buffer (no
    // store), never used as an offset.
    const int di = (DENSE != kDenseNone && a.NDO > 0) ? a.DenseOutputIndex[sid] : 0;
    S.DOIDX = (di < 0 || di > a.NDO) ? a.NDO : di;
  }

  // Unit scope into registers / private memory (Solver.cuh:234-257); layouts that set
  // kUnitParamsInLds keep the unit parameters in LDS instead.
  double x[UD], xn[UD], err[UD];
  double upar_regs[unit_params_in_lds<L>() ? 1 : (NUP > 0 ? NUP : 1)];
  double* upar = upar_regs;
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
  if constexpr (unit_params_in_lds<L>()) {
    upar = lay.unit_params();
    if (primal) {
#pragma unroll
      for (int i = 0; i < NUP; i++) upar[i] = a.UnitParameters[gtid + i * TLT];
    }
  } else {
#pragma unroll
    for (int i = 0; i < NUP; i++) upar[i] = a.UnitParameters[gtid + i * TLT];
  }
#pragma nounroll
  for (int i = 0; i < NUA; i++) uacc[i] = a.UnitAccessories[gtid + i * TLT];
#pragma unroll
  for (int i = 0; i < NiUA; i++) iuacc[i] = a.IntegerUnitAccessories[gtid + i * TLT];
#pragma unroll
  for (int i = 0; i < NE; i++) {
    etol[i] = a.EventTolerance[i];
    edir[i] = a.EventDirection[i];
  }
  lay.sync();

  // Initialization hook (Solver.cuh:262-281).
  CoupledSystems_PerBlock_Initialization<double>(sid, uid, S.DOIDX, S.T, S.dT, S.TD, x, upar,
                                                 S.sPAR, a.GlobalParameters,
                                                 a.IntegerGlobalParameters, uacc, iuacc, S.sACC,
                                                 S.isACC);
  lay.sync();
  double t = S.T;
  double dt = a.InitialTimeStep;
  double dtNew = a.InitialTimeStep;
  const double TD1 = S.TD[1];

  if (NE > 0)  // Solver.cuh:283-305
    CoupledSystems_PerBlock_EventFunction<double>(sid, uid, ea, t, dt, S.TD, x, upar, S.sPAR,
                                                  a.GlobalParameters,
                                                  a.IntegerGlobalParameters, uacc, iuacc,
                                                  S.sACC, S.isACC);

  // Dense output (DenseOutput.cuh:313-395); the index lives in S.DOIDX.
  int updateDO = 0, skipped = 0;
  double doTime = 0.0;
  auto store_dense = [&](int updateStep) {
    if (updateDO == 1) {
      const int idx = S.DOIDX;
      if (primal) {
#pragma unroll
        for (int i = 0; i < UD; i++)
          a.DenseOutputStates[gtid + i * TLT + (long long)idx * TLT * UD] = x[i];
        if (uid == 0) a.DenseOutputTimeInstances[sid + (long long)idx * NS] = t;
      }
      lay.sync();
      if (uid == 0) S.DOIDX = idx + 1;
      lay.sync();
      skipped = 0;
      doTime = bex::fmin_d(t + a.DenseOutputMinimumTimeStep, TD1);
    }
    if ((updateDO == 0) && (updateStep == 1)) skipped++;
  };
  if (DENSE != kDenseNone && a.NDO > 0) {
    updateDO = (S.DOIDX < a.NDO) ? 1 : 0;
    store_dense(0);
  }

  // Interpolated dense output (no reference counterpart): outputs at t0 + k * DenseOutput-
  // MinimumTimeStep, each by cubic Hermite interpolation over the accepted step that contains it,
  // from the step's end points and their derivatives. The end point's derivative is the next
  // attempt's first stage, so an interval is flushed at that stage; after the last step one
  // extra first-stage evaluation flushes the final interval. The step's start point and
  // derivative wait in DenseAux ([2 UD][TLT]).
  // The interval and grid bookkeeping lives in the system's LDS record (uniform values; kept
  // out of the registers that hold the stage vectors).
  const bool interp = (DENSE == kDenseHermite) && (a.NDO > 0);
  if (uid == 0) {
    S.pending = 0;
    S.flushOnly = 0;
    S.gridT0 = t;
    S.kq = 1;
  }
  lay.sync();
  auto dense_flush = [&](const double* fB) {
    for (;;) {
      const int idx = S.DOIDX;
      if (idx >= a.NDO) break;
      const double tA = S.tA, tB = S.tB;
      const int kq = S.kq;  // read by every thread before the barrier, like idx
      const double tq = S.gridT0 + (double)kq * a.DenseOutputMinimumTimeStep;
      if (!(tq <= tB)) break;
      const double h = tB - tA;
      const double th = (tq - tA) / h;
      const double sm = 1.0 - th;
      const double h00 = ((1.0 + 2.0 * th) * sm) * sm;
      const double h10 = (th * sm) * sm;
      const double h01 = (th * th) * (3.0 - 2.0 * th);
      const double h11 = (th * th) * (th - 1.0);
      if (primal) {
#pragma unroll
        for (int i = 0; i < UD; i++) {
          const double yA = a.DenseAux[gtid + i * TLT];
          const double fA = a.DenseAux[gtid + (UD + i) * TLT];
          const double y = ((h00 * yA + (h10 * h) * fA) + h01 * x[i]) + (h11 * h) * fB[i];
          a.DenseOutputStates[gtid + i * TLT + (long long)idx * TLT * UD] = y;
        }
        if (uid == 0) a.DenseOutputTimeInstances[sid + (long long)idx * NS] = tq;
      }
      lay.sync();
      // Unit 0 may be served by threads of several waves: they store values read before the
      // barrier, never a read-modify-write of the record.
      if (uid == 0) {
        S.DOIDX = idx + 1;
        S.kq = kq + 1;
      }
      lay.sync();
    }
  };

  int terminate = 0;
  int accepted = 0;
  // Issue priority follows this system's accepted steps (raised below); a persistent wave that
  // takes its next system from a queue starts it at the base level again.
  __builtin_amdgcn_s_setprio(0);
  long long stage_count = 0;  // RK stages of this launch (sample selection)
  long long polls = 0;
  while (terminate == 0) {  // Solver.cuh:329-572
    // Cooperative stop and attempt cap (no reference counterpart): every 256 attempts the
    // system reads the host's stop flag and compares its attempts in this launch with
    // MaxStepAttempts; a set flag or a reached cap ends the system here, as a user termination
    // would, with its state and time written back as they stand. Both decisions are taken for
    // the whole system (the attempt count is the same on every lane).
    if ((++polls & 255) == 0) {
      const bool capped = a.MaxStepAttempts > 0 && polls >= a.MaxStepAttempts;
      const int req =
          a.StopFlag != nullptr ? *reinterpret_cast<const volatile int*>(a.StopFlag) : 0;
      if (capped || lay.any(req != 0)) {
        terminate = 1;
        continue;
      }
    }
    int updateStep = 1, endReached = 0;
    dt = dtNew;
    if (dt > (TD1 - t)) {
      dt = TD1 - t;
      endReached = 1;
    }
    if (leader && S.flushOnly == 0) count(st, BEX_STAT_ATTEMPTS, 1);

    bool finiteLane = true;
    double k1[UD], k2[UD], k3[UD], k4[UD], k5[UD], k6[UD];
    constexpr int NSTAGE = (ALG == kRKCK45) ? 6 : 4;
    // The final pass of interpolated dense output evaluates the first stage only.
    const int nstage = (interp && S.flushOnly != 0) ? 1 : NSTAGE;
    // One copy of the right-hand side + BiCG body: the stage loop is not unrolled. Each stage's
    // input is the reference's sum (Steppers.cuh:153-400) contracted as nvcc contracts it: a
    // multiply, one fma per further term, and x + sum * dt as one fma; likewise the stage time.
#pragma nounroll
    for (int s = 0; s < nstage; s++) {
      double xs[UD], F[UD];
      double Ts;
      if (ALG == kRKCK45) {
        // One branch per stage for both components and the stage time; the tableau row is read
        // from constant memory with scalar loads.
        const ConstRow A = rkck45_row(s, t);
        switch (s) {
          case 0:
#pragma unroll
            for (int i = 0; i < UD; i++) xs[i] = x[i];
            Ts = t;
            break;
          case 1:
#pragma unroll
            for (int i = 0; i < UD; i++) xs[i] = fma(k1[i] * A[0], dt, x[i]);
            Ts = fma(A[5], dt, t);
            break;
          case 2:
#pragma unroll
            for (int i = 0; i < UD; i++) xs[i] = fma(fma(k2[i], A[1], k1[i] * A[0]), dt, x[i]);
            Ts = fma(A[5], dt, t);
            break;
          case 3:
#pragma unroll
            for (int i = 0; i < UD; i++)
              xs[i] = fma(fma(k3[i], A[2], fma(k2[i], A[1], k1[i] * A[0])), dt, x[i]);
            Ts = fma(A[5], dt, t);
            break;
          case 4:
#pragma unroll
            for (int i = 0; i < UD; i++)
              xs[i] = fma(fma(k4[i], A[3], fma(k3[i], A[2], fma(k2[i], A[1], k1[i] * A[0]))),
                          dt, x[i]);
            Ts = t + dt;
            break;
          default:
#pragma unroll
            for (int i = 0; i < UD; i++)
              xs[i] = fma(fma(k5[i], A[4],
                              fma(k4[i], A[3], fma(k3[i], A[2], fma(k2[i], A[1], k1[i] * A[0])))),
                          dt, x[i]);
            Ts = fma(A[5], dt, t);
            break;
        }
      } else {
#pragma unroll
        for (int i = 0; i < UD; i++) {
          switch (s) {
            case 0: xs[i] = x[i]; break;
            case 1: xs[i] = fma(k1[i], 0.5 * dt, x[i]); break;
            case 2: xs[i] = fma(k2[i], 0.5 * dt, x[i]); break;
            default: xs[i] = fma(k3[i], dt, x[i]); break;
          }
        }
        Ts = (s == 0) ? t : ((s == 3) ? t + dt : t + 0.5 * dt);
      }
      // Right-hand side and coupling solve (Steppers.cuh:431-562).
      double cpt, cpf;
      CoupledSystems_PerBlock_OdeFunction<double>(sid, uid, F, xs, Ts, upar, S.sPAR,
                                                  a.GlobalParameters,
                                                  a.IntegerGlobalParameters, uacc, iuacc,
                                                  S.sACC, S.isACC, &cpt, &cpf);
      bool sample_now = false;
      if (L::kMode == kSampling) {
        sample_now = ((stage_count + sid) % a.SampleEvery) == 0;
        stage_count++;
      }
      auto features = [&](double crs, double* wf) {
        CoupledSystems_PerBlock_WarmStartFeatures<double>(sid, uid, wf, Ts, dt, xs, upar, S.sPAR,
                                                          a.GlobalParameters,
                                                          a.IntegerGlobalParameters, crs);
      };
      const BicgOut o = lay.coupling(F, xs[0], cpt, cpf, a, features, sample_now);
      if (leader) {
        count(st, BEX_STAT_STAGES, 1);
        count(st, BEX_STAT_BICG_ITERATIONS, o.iterations);
        count(st, BEX_STAT_BICG_RHO_EXITS, o.rho_exit ? 1 : 0);
        count(st, BEX_STAT_BICG_MAXIT, o.maxit ? 1 : 0);
        count(st, BEX_STAT_HIST0 + (o.iterations < 20 ? o.iterations : 19), 1);
      }
      switch (s) {
        case 0: store_stage(k1, F); break;
        case 1: store_stage(k2, F); break;
        case 2: store_stage(k3, F); break;
        case 3: store_stage(k4, F); break;
        case 4: store_stage(k5, F); break;
        default: store_stage(k6, F); break;
      }
    }
    // k1 = f(t, x) is the pending step's end-point derivative.
    if (interp && S.pending != 0) {
      dense_flush(k1);
      lay.sync();  // every wave has read the flag (an empty flush has no barrier of its own)
      if (uid == 0) S.pending = 0;
      lay.sync();
    }
    if (interp && S.flushOnly != 0) {
      terminate = 1;
      continue;
    }

    if (ALG == kRKCK45) {
#pragma unroll
      for (int i = 0; i < UD; i++) {
        // Steppers.cuh:408-418, contracted like the stage inputs.
        double sum = k1[i] * (37.0 / 378.0);
        sum = fma(k3[i], 250.0 / 621.0, sum);
        sum = fma(k4[i], 125.0 / 594.0, sum);
        sum = fma(k6[i], 512.0 / 1771.0, sum);
        xn[i] = fma(sum, dt, x[i]);
        double e = k1[i] * (37.0 / 378.0 - 2825.0 / 27648.0);
        e = fma(k3[i], 250.0 / 621.0 - 18575.0 / 48384.0, e);
        e = fma(k4[i], 125.0 / 594.0 - 13525.0 / 55296.0, e);
        e = fma(k5[i], 0.0 - 277.0 / 14336.0, e);
        e = fma(k6[i], 512.0 / 1771.0 - 1.0 / 4.0, e);
        err[i] = fma(dt, bex::fabs_d(e), 1e-18);
        if (!bex::is_finite(xn[i]) || !bex::is_finite(err[i])) finiteLane = false;
      }
    } else {
#pragma unroll
      for (int i = 0; i < UD; i++) {
        xn[i] = fma(k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i], dt * (1.0 / 6.0), x[i]);
        err[i] = 0.0;
        if (!bex::is_finite(xn[i])) finiteLane = false;
      }
    }
    int isFinite = lay.all(finiteLane) ? 1 : 0;

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
      updateStep = lay.all(ru) ? 1 : 0;
      const double relErr = lay.min_keep(rr);
      double mult = (updateStep == 1) ? 0.8 * bex::pow(relErr, 1.0 / 5.0)
                                      : 0.8 * bex::pow(relErr, 1.0 / 4.0);
      if (!bex::is_finite(mult)) isFinite = 0;
      if (isFinite == 0) {
        mult = a.TimeStepShrinkLimit;
        updateStep = 0;
        if (dt < (a.MinimumTimeStep * 1.01)) {
          terminate = 1;
          if (leader) st[BEX_STAT_TERMINATED]++;
        }
      } else if (dt < (a.MinimumTimeStep * 1.01)) {
        updateStep = 1;
        if (leader) st[BEX_STAT_FORCED_ACCEPTS]++;
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
        if (leader) st[BEX_STAT_TERMINATED]++;
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
      if (lay.any(corrLane)) {
        const double evStep = lay.min_keep(evLane);
        if (evStep < a.MinimumTimeStep) {
          if (leader) st[BEX_STAT_EVENTS_OMITTED]++;
        } else {
          dtNew = evStep;
          updateStep = 0;
        }
      }
    }

    if (updateStep == 1) {  // Solver.cuh:433-525
      // Interpolated dense output: the step's start point and derivative are kept only when the
      // next output instant falls inside the step (a few dozen of the thousands of steps of a
      // period), since only then is the step flushed.
      bool covers = false;
      if (interp) {
        covers = (S.DOIDX < a.NDO) &&
                 (S.gridT0 + (double)S.kq * a.DenseOutputMinimumTimeStep <= t + dt);
        if (covers && primal) {
#pragma unroll
          for (int i = 0; i < UD; i++) {
            a.DenseAux[gtid + i * TLT] = x[i];
            a.DenseAux[gtid + (UD + i) * TLT] = k1[i];
          }
        }
        if (uid == 0) S.tA = t;
      }
      t += dt;
#pragma unroll
      for (int i = 0; i < UD; i++) x[i] = xn[i];
      // Systems that keep stepping long after the typical one become the launch's tail: raise
      // the issue priority of their wave step by step so they run at close to single-wave
      // latency while the remaining waves fill the idle issue slots.
      ++accepted;
      if (accepted == a.PriorityStepThreshold) __builtin_amdgcn_s_setprio(1);
      if (accepted == 2 * a.PriorityStepThreshold) __builtin_amdgcn_s_setprio(2);
      if (accepted == 4 * a.PriorityStepThreshold) __builtin_amdgcn_s_setprio(3);
      if (uid == 0) {
        if (leader) count(st, BEX_STAT_ACCEPTED, 1);
        S.T = t;
        S.dT = dt;
      }
      lay.sync();
      if (run_actions)
        CoupledSystems_PerBlock_ActionAfterSuccessfulTimeStep<double>(
            sid, uid, S.UDT, S.T, S.dT, S.TD, x, upar, S.sPAR, a.GlobalParameters,
            a.IntegerGlobalParameters, uacc, iuacc, S.sACC, S.isACC);
      lay.sync();
      if (NE > 0) {
#pragma unroll
        for (int i = 0; i < NE; i++) {
          if (((ea[i] > etol[i]) && (bex::fabs_d(en[i]) < etol[i]) && (edir[i] <= 0)) ||
              ((ea[i] < -etol[i]) && (bex::fabs_d(en[i]) < etol[i]) && (edir[i] >= 0))) {
            if (run_actions)
              CoupledSystems_PerBlock_ActionAfterEventDetection<double>(
                  sid, uid, i, S.UDT, S.T, S.dT, S.TD, x, upar, S.sPAR, a.GlobalParameters,
                  a.IntegerGlobalParameters, uacc, iuacc, S.sACC, S.isACC);
            if (primal) atomicAdd(&st[BEX_STAT_EVENTS], 1);
          }
        }
        lay.sync();
        CoupledSystems_PerBlock_EventFunction<double>(sid, uid, en, S.T, dtNew, S.TD, x, upar,
                                                      S.sPAR, a.GlobalParameters,
                                                      a.IntegerGlobalParameters, uacc, iuacc,
                                                      S.sACC, S.isACC);
#pragma unroll
        for (int i = 0; i < NE; i++) ea[i] = en[i];
      }
      lay.sync();
      t = S.T;
      if (interp) {
        if (uid == 0) {
          S.tB = t;
          S.pending = covers ? 1 : 0;
        }
        lay.sync();
      }
    } else if (leader) {
      count(st, BEX_STAT_REJECTED, 1);
    }
    const int udt = S.UDT;

    if (DENSE == kDenseSteps && a.NDO > 0) {  // Solver.cuh:528-557
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

    if (updateStep == 1 && (endReached == 1 || udt == 1)) {  // :560-569
      if (interp && S.pending != 0 && S.DOIDX < a.NDO) {
        if (uid == 0) S.flushOnly = 1;  // one more first-stage evaluation, then stop
        lay.sync();
      } else {
        terminate = 1;
      }
    }
  }

  // Finalization hook and write back (Solver.cuh:577-647).
  if (uid == 0) S.T = t;
  lay.sync();
  CoupledSystems_PerBlock_Finalization<double>(sid, uid, S.DOIDX, S.T, S.dT, S.TD, x, upar,
                                               S.sPAR, a.GlobalParameters,
                                               a.IntegerGlobalParameters, uacc, iuacc, S.sACC,
                                               S.isACC);
  lay.sync();
  if (primal) {
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
  }
  if (leader) {
    a.ActualTime[sid] = S.T;
    a.TimeDomain[sid] = S.TD[0];
    a.TimeDomain[sid + NS] = S.TD[1];
    for (int i = 0; i < D::NSA; i++) a.SystemAccessories[sid + i * NS] = S.sACC[i];
    for (int i = 0; i < D::NiSA; i++) a.IntegerSystemAccessories[sid + i * NS] = S.isACC[i];
    if (DENSE != kDenseNone && a.NDO > 0) a.DenseOutputIndex[sid] = S.DOIDX;
  }
  lay.sync();
  if (primal)
    for (int k = uid; k < BEX_NSTAT; k += N) a.Stats[(long long)sid * BEX_NSTAT + k] += st[k];
  if (leader && a.Timeline) a.Timeline[2 * (long long)sid + 1] = (long long)wall_clock64();
}

template <class L, int ALG, int DENSE>
__global__ __launch_bounds__(L::kBlock) __attribute__((amdgpu_waves_per_eu(L::kWavesPerEU)))
void solve_kernel(const BexKernelArgs a, long long count) {
  extern __shared__ __align__(16) unsigned char smem[];
  L lay;
  lay.init(smem, a);
  lay.bind(a, nullptr, (long long)blockIdx.x * L::kItemsPerBlock + threadIdx.x / (L::kBlock / L::kItemsPerBlock),
           count);
  integrate<L, ALG, DENSE>(lay, a);
}

// Persistent variant: a fixed grid whose waves (or workgroups, for multi-wave systems) take
// work items from a device queue until it is empty. order lists the systems in the order they
// are handed out (longest expected run first; nullptr: first_system, first_system + 1, ...);
// items are kSystemsPerItem consecutive entries. *counter must be zero at launch; every wave
// exits once the counter passes nitems. A slot that frees up takes the next item at once,
// where the hardware's in-order workgroup dispatch of solve_kernel leaves slots idle once the
// systems' run times diverge. Pools run it without dense output, ProblemSolver::Solve() in
// every dense mode.
template <class L, int ALG, int DENSE = kDenseNone>
__global__ __launch_bounds__(L::kBlock) __attribute__((amdgpu_waves_per_eu(L::kWavesPerEU)))
void queue_kernel(const BexKernelArgs a, const int* __restrict__ order, long long count,
                 long long nitems, unsigned long long* counter) {
  extern __shared__ __align__(16) unsigned char smem[];
  L lay;
  lay.init(smem, a);
  __shared__ unsigned long long taken[L::kItemsPerBlock];
  const int slot = threadIdx.x / (L::kBlock / L::kItemsPerBlock);
  const bool leader = (threadIdx.x % (L::kBlock / L::kItemsPerBlock)) == 0;
  for (;;) {
    if (leader) taken[slot] = atomicAdd(counter, 1ull);
    lay.sync();
    const long long item = (long long)taken[slot];
    lay.sync();
    if (item >= nitems) break;
    lay.bind(a, order, item, count);
    integrate<L, ALG, DENSE>(lay, a);
    lay.sync();
  }
}

}  // namespace bexk
