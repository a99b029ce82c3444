// Deterministic FP64 elementary functions for the coupled Keller-Miksis solver.
//
// The reference right-hand side (BubbleEnsemble_SystemDefinition.cuh:15-23) and its
// step-size controller (CoupledSystems_PerBlock_ExplicitRungeKutta_ErrorControllers.cuh:503-506)
// call pow, sincospi, sin and cos from the device math library. Device and host math
// libraries differ in the last bits, and an adaptive integrator amplifies such differences
// into different accepted-step sequences. These functions are written once, in plain FP64
// arithmetic with explicit fma() and no contraction, so that the gfx950 kernel and the host
// build produce identical bits for identical inputs. Accuracy (measured against mpmath in the
// tests) is about 1-2 ulp over the argument ranges the solver uses.
//
// Every operation used here is exactly specified by IEEE-754 (add, mul, div, fma, rint,
// floor, ldexp, comparisons), so the results do not depend on the compiler as long as the
// translation unit is built with -ffp-contract=off and without fast-math.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define BEX_HD __host__ __device__ inline
#else
#define BEX_HD inline
#endif

namespace bex {

// Polynomial coefficients and splitting constants, one table (Taylor coefficients rounded to
// double, generated with mpmath). On the device the table sits in constant memory and is read
// with scalar loads at the point of use: Coef(dep) passes the table's address through an empty
// asm that takes the evaluation's varying argument as an input, so loop-invariant code motion
// cannot lift the loads out of the solver's time loop (held across the loop, ~80 coefficients
// would be spilled), while neighbouring coefficients still merge into wide scalar loads. A
// constant written as an immediate would instead cost two scalar moves per use, since 64-bit
// VOP3 operands take no literal on gfx950.
enum CoefIndex : int {
  kSinpiTail = 0,
  kCospiTail = 10,
  kPi = 20,
  kPio2 = 22,
  kSinTail = 26,
  kCosTail = 36,
  kLog = 47,
  kExp = 62,
  kMisc = 78,
  kNumCoef = 81
};

#define BEX_DETMATH_COEFFICIENTS \
    /* kSinpiTail: sin(pi r) - pi r: r^2-Horner coefficients, highest order first */ \
    0x1.2877020d52cf0p-31, -0x1.8a404211f9547p-26, 0x1.aaec32af93359p-21, \
    -0x1.6fadb9f155744p-16, 0x1.e8f434d018d63p-12, -0x1.e3074fde8871fp-8, \
    0x1.50783487ee782p-4, -0x1.32d2cce62bd86p-1, 0x1.466bc6775aae2p+1, \
    -0x1.4abbce625be53p+2, \
    /* kCospiTail: cos(pi r) - 1: r^2-Horner coefficients, highest order first */ \
    0x1.ef6e308d6d1c4p-29, -0x1.2a0c591af8314p-23, 0x1.20c62c2f2d7f5p-18, \
    -0x1.b6e24f44b128fp-14, 0x1.f9d38a3763cc3p-10, -0x1.a6d1f2a204a8cp-6, \
    0x1.e1f506891babbp-3, -0x1.55d3c7e3cbffap+0, 0x1.03c1f081b5ac4p+2, \
    -0x1.3bd3cc9be45dep+2, \
    /* kPi: pi = hi + lo */ \
    0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53, \
    /* kPio2: 2/pi; pi/2 in three parts (Cody-Waite) */ \
    0x1.45f306dc9c883p-1, 0x1.921fb54400000p+0, 0x1.0b4611a600000p-34, \
    0x1.3198a2e000000p-69, \
    /* kSinTail: sin(r) - r: r^2-Horner coefficients, highest order first */ \
    0x1.71b8ef6dcf572p-66, -0x1.2f49b46814157p-57, 0x1.952c77030ad4ap-49, \
    -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33, -0x1.ae64567f544e4p-26, \
    0x1.71de3a556c734p-19, -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7, \
    -0x1.5555555555555p-3, \
    /* kCosTail: cos(r) - 1: r^2-Horner coefficients, highest order first */ \
    -0x1.0ce396db7f853p-70, 0x1.e542ba4020225p-62, -0x1.6827863b97d97p-53, \
    0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37, 0x1.1eed8eff8d898p-29, \
    -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16, -0x1.6c16c16c16c17p-10, \
    0x1.5555555555555p-5, -0x1.0000000000000p-1, \
    /* kLog: log(m) = 2s + 2s^3 P(s^2), P: 1/27 ... 1/3; ln 2 = a + b, k*a exact */ \
    0x1.2f684bda12f68p-5, 0x1.47ae147ae147bp-5, 0x1.642c8590b2164p-5, \
    0x1.8618618618618p-5, 0x1.af286bca1af28p-5, 0x1.e1e1e1e1e1e1ep-5, \
    0x1.1111111111111p-4, 0x1.3b13b13b13b14p-4, 0x1.745d1745d1746p-4, \
    0x1.c71c71c71c71cp-4, 0x1.2492492492492p-3, 0x1.999999999999ap-3, \
    0x1.5555555555555p-2, 0x1.62e42ff000000p-1, -0x1.718432a1b0e26p-35, \
    /* kExp: 1/ln 2; ln 2 = a + b; exp(r) = 1 + r + r^2 P(r), P: 1/14! ... 1/2 */ \
    0x1.71547652b82fep+0, 0x1.62e42ff000000p-1, -0x1.718432a1b0e26p-35, \
    0x1.93974a8c07c9dp-37, 0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29, \
    0x1.ae64567f544e4p-26, 0x1.27e4fb7789f5cp-22, 0x1.71de3a556c734p-19, \
    0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-13, 0x1.6c16c16c16c17p-10, \
    0x1.1111111111111p-7, 0x1.5555555555555p-5, 0x1.5555555555555p-3, \
    0x1.0000000000000p-1, \
    /* kMisc: sqrt(2) (log reduction); exp overflow and underflow thresholds */ \
    0x1.6a09e667f3bcdp+0, 709.782712893384, -745.1332191019412

#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ const double kCoefDevice[kNumCoef] = {BEX_DETMATH_COEFFICIENTS};
#endif
static constexpr double kCoefHost[kNumCoef] = {BEX_DETMATH_COEFFICIENTS};

struct Coef {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(4))) double* p;  // constant memory: scalar loads
  __device__ explicit Coef(double dep)
      : p((const __attribute__((address_space(4))) double*)kCoefDevice) {
    asm("" : "+s"(p) : "v"(dep));
  }
#else
  const double* p;
  explicit Coef(double) : p(kCoefHost) {}
#endif
  BEX_HD double operator[](int i) const { return p[i]; }
};

// One Horner step p * x + c with a table coefficient c, as a single fma. On the device the
// coefficient is marked as living in scalar registers (an empty asm), which keeps it the addend
// of one v_fma_f64; left to itself the compiler copies it into vector registers for a
// v_fmac_f64 (three instructions).
BEX_HD double horner(double p, double x, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+s"(c));
#endif
  return fma(p, x, c);
}

BEX_HD uint64_t as_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
BEX_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

BEX_HD bool is_nan(double x) { return x != x; }
BEX_HD bool is_finite(double x) { return (as_bits(x) & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL; }
BEX_HD double quiet_nan() { return from_bits(0x7ff8000000000000ULL); }
BEX_HD double infinity() { return from_bits(0x7ff0000000000000ULL); }

// fmin/fmax with the C99 NaN rule (a NaN operand is ignored), written out so that the
// host and device agree on every input: b is taken when it is smaller (larger) or when a is
// NaN, otherwise a (which covers b NaN). One select, no branches.
BEX_HD double fmin_d(double a, double b) { return ((b < a) | is_nan(a)) ? b : a; }
BEX_HD double fmax_d(double a, double b) { return ((b > a) | is_nan(a)) ? b : a; }
BEX_HD double fabs_d(double x) { return from_bits(as_bits(x) & 0x7fffffffffffffffULL); }

// ---------------------------------------------------------------------------------------
// Polynomial kernels (Taylor coefficients rounded to double, generated with mpmath).

// sin(pi*r) - pi*r, for |r| <= 1/4, as r^3 * P(r^2).
BEX_HD double sinpi_tail(double r, double r2, const Coef& C) {
  double p = C[kSinpiTail];
  p = horner(p, r2, C[kSinpiTail + 1]);
  p = horner(p, r2, C[kSinpiTail + 2]);
  p = horner(p, r2, C[kSinpiTail + 3]);
  p = horner(p, r2, C[kSinpiTail + 4]);
  p = horner(p, r2, C[kSinpiTail + 5]);
  p = horner(p, r2, C[kSinpiTail + 6]);
  p = horner(p, r2, C[kSinpiTail + 7]);
  p = horner(p, r2, C[kSinpiTail + 8]);
  p = horner(p, r2, C[kSinpiTail + 9]);
  return (r * r2) * p;
}

// cos(pi*r) - 1, for |r| <= 1/4, as r^2 * Q(r^2).
BEX_HD double cospi_tail(double r2, const Coef& C) {
  double q = C[kCospiTail];
  q = horner(q, r2, C[kCospiTail + 1]);
  q = horner(q, r2, C[kCospiTail + 2]);
  q = horner(q, r2, C[kCospiTail + 3]);
  q = horner(q, r2, C[kCospiTail + 4]);
  q = horner(q, r2, C[kCospiTail + 5]);
  q = horner(q, r2, C[kCospiTail + 6]);
  q = horner(q, r2, C[kCospiTail + 7]);
  q = horner(q, r2, C[kCospiTail + 8]);
  q = horner(q, r2, C[kCospiTail + 9]);
  return r2 * q;
}

// sin(r) - r and cos(r) - 1 for |r| <= pi/4 (+ reduction slack).
BEX_HD double sin_tail(double r, double r2, const Coef& C) {
  double p = C[kSinTail];
  p = horner(p, r2, C[kSinTail + 1]);
  p = horner(p, r2, C[kSinTail + 2]);
  p = horner(p, r2, C[kSinTail + 3]);
  p = horner(p, r2, C[kSinTail + 4]);
  p = horner(p, r2, C[kSinTail + 5]);
  p = horner(p, r2, C[kSinTail + 6]);
  p = horner(p, r2, C[kSinTail + 7]);
  p = horner(p, r2, C[kSinTail + 8]);
  p = horner(p, r2, C[kSinTail + 9]);
  return (r * r2) * p;
}
BEX_HD double cos_tail(double r2, const Coef& C) {
  double q = C[kCosTail];
  q = horner(q, r2, C[kCosTail + 1]);
  q = horner(q, r2, C[kCosTail + 2]);
  q = horner(q, r2, C[kCosTail + 3]);
  q = horner(q, r2, C[kCosTail + 4]);
  q = horner(q, r2, C[kCosTail + 5]);
  q = horner(q, r2, C[kCosTail + 6]);
  q = horner(q, r2, C[kCosTail + 7]);
  q = horner(q, r2, C[kCosTail + 8]);
  q = horner(q, r2, C[kCosTail + 9]);
  q = horner(q, r2, C[kCosTail + 10]);
  return r2 * q;
}

// ---------------------------------------------------------------------------------------
// The functions below evaluate their main path on sanitised arguments and apply the special
// cases (non-finite input, integer-valued sincospi argument, pow's C99 cases, exp's overflow
// and underflow) at the end, instead of returning early: on the device an early return is a
// divergent branch (exec-mask bookkeeping on every call). sincospi and sincos select them (a
// select is one instruction per half); pow and exp_dd, whose chains are longer, put them
// behind one branch that ordinary arguments never take. The values are those of the same
// formulas with early returns, for every input.

// Quadrant selection: (s, c) = (sin, cos) of r + q pi/2 from (sin r, cos r) = (sr, cr).
BEX_HD void quadrant(int q, double sr, double cr, double* s, double* c) {
  const bool odd = (q & 1) != 0;
  const double sv = odd ? cr : sr;
  const double cv = odd ? sr : cr;
  *s = (q & 2) ? -sv : sv;
  *c = ((q + 1) & 2) ? -cv : cv;
}

// sincospi(x) = (sin(pi x), cos(pi x)). The reduction x = n/2 + r is exact, so the result is
// accurate for every finite x.
BEX_HD void sincospi(double x, double* s, double* c) {
  const Coef C(x);
  const double ax = fabs_d(x);
  const bool finite = is_finite(x);
  const bool integral = ax >= 0x1.0p52;  // x is an integer: sin = 0, cos = (-1)^x
  const double xm = (finite && !integral) ? x : 0.0;
  const double n = rint(2.0 * xm);              // exact: 2x is exact, rint is exact
  const double r = fma(-0.5, n, xm);            // exact: |r| <= 1/4
  const double q4 = n - 4.0 * floor(n * 0.25);  // n mod 4, exact
  const int q = (int)q4;
  const double r2 = r * r;
  // sin(pi r) = r*pi_hi + (r*pi_lo + tail)
  const double sp = fma(r, C[kPi], fma(r, C[kPi + 1], sinpi_tail(r, r2, C)));
  const double cp = 1.0 + cospi_tail(r2, C);
  double sv, cv;
  quadrant(q, sp, cp, &sv, &cv);
  const double half = x * 0.5;
  const bool odd = (floor(half) != half);
  sv = integral ? 0.0 : sv;
  cv = integral ? (odd ? -1.0 : 1.0) : cv;
  *s = finite ? sv : quiet_nan();
  *c = finite ? cv : quiet_nan();
}

// sin and cos with a three-part Cody-Waite reduction by pi/2. The reduction is exact for
// |x| < 2^20 * pi/2 (~1.6e6); beyond that the result stays finite but loses accuracy. The
// solver evaluates these at 2*pi*f2/f1*tau + theta with tau in [0, 1]. Non-finite x is
// replaced by 0 here; the callers return NaN for it.
BEX_HD void reduce_pio2(double x, double* r, int* q, const Coef& C) {
  const double xm = is_finite(x) ? x : 0.0;
  const double n = rint(xm * C[kPio2]);
  double t = fma(-n, C[kPio2 + 1], xm);
  t = fma(-n, C[kPio2 + 2], t);
  t = fma(-n, C[kPio2 + 3], t);
  *r = t;
  const double q4 = n - 4.0 * floor(n * 0.25);
  *q = (int)q4;
}

// sin and cos of the same argument.
BEX_HD void sincos(double x, double* s, double* c) {
  const Coef C(x);
  double r;
  int q;
  reduce_pio2(x, &r, &q, C);
  const double r2 = r * r;
  const double sr = r + sin_tail(r, r2, C);
  const double cr = 1.0 + cos_tail(r2, C);
  double sv, cv;
  quadrant(q, sr, cr, &sv, &cv);
  const bool finite = is_finite(x);
  *s = finite ? sv : quiet_nan();
  *c = finite ? cv : quiet_nan();
}

// sin(x) and cos(x): bit-identical to the matching output of sincos.
BEX_HD double sin(double x) {
  double s, c;
  sincos(x, &s, &c);
  return s;
}

BEX_HD double cos(double x) {
  double s, c;
  sincos(x, &s, &c);
  return c;
}

// ---------------------------------------------------------------------------------------
// log(x) for finite x > 0, returned as an unevaluated sum hi + lo carrying ~10 extra bits.
BEX_HD double log_dd(double x, double* lo) {
  const Coef C(x);
  const bool subnormal = x < 0x1.0p-1022;
  x = subnormal ? x * 0x1.0p54 : x;  // exact
  int k = subnormal ? -54 : 0;
  uint64_t b = as_bits(x);
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
  const bool above = m > C[kMisc];  // m > sqrt(2): move to [sqrt(1/2), sqrt(2))
  m = above ? m * 0.5 : m;
  k += above ? 1 : 0;
  // log(m) = 2 atanh(s), s = f / (2 + f), f = m - 1 (exact).
  double f = m - 1.0;
  double dh = 2.0 + f;
  double dl = f - (dh - 2.0);                 // 2 + f = dh + dl exactly
  double rinv = 1.0 / dh;
  double sh = f * rinv;
  double e = fma(-sh, dh, f) - sh * dl;       // f - sh*(dh+dl)
  double sl = e * rinv;
  double s2 = sh * sh;
  double p = C[kLog];                   // 1/27
  p = horner(p, s2, C[kLog + 1]);       // 1/25
  p = horner(p, s2, C[kLog + 2]);       // 1/23
  p = horner(p, s2, C[kLog + 3]);       // 1/21
  p = horner(p, s2, C[kLog + 4]);       // 1/19
  p = horner(p, s2, C[kLog + 5]);       // 1/17
  p = horner(p, s2, C[kLog + 6]);       // 1/15
  p = horner(p, s2, C[kLog + 7]);       // 1/13
  p = horner(p, s2, C[kLog + 8]);       // 1/11
  p = horner(p, s2, C[kLog + 9]);       // 1/9
  p = horner(p, s2, C[kLog + 10]);      // 1/7
  p = horner(p, s2, C[kLog + 11]);      // 1/5
  p = horner(p, s2, C[kLog + 12]);      // 1/3
  double tail = (sh * s2) * p;                // s^3 (1/3 + s^2/5 + ...)
  double lm_hi = 2.0 * sh;
  double lm_lo = 2.0 * (sl + tail);
  // + k*ln2, ln2 split so that k*LN2_A is exact.
  double kd = (double)k;
  double a = kd * C[kLog + 13];
  double bsum = a + lm_hi;
  // two-sum of a and lm_hi
  double bb = bsum - a;
  double err = (a - (bsum - bb)) + (lm_hi - bb);
  double low = err + (lm_lo + kd * C[kLog + 14]);
  double hi = bsum + low;
  *lo = low - (hi - bsum);
  return hi;
}

// exp(th + tl) for |tl| << |th|: +inf above the overflow threshold, 0 below the underflow
// threshold, NaN for NaN th.
BEX_HD double exp_dd(double th, double tl) {
  const Coef C(th);
  const bool over = th > C[kMisc + 1];
  const bool under = th < C[kMisc + 2];
  const bool nan = is_nan(th);
  const bool regular = !(over || under || nan);
  const double thm = regular ? th : 0.0;
  const double tlm = regular ? tl : 0.0;
  double n = rint(thm * C[kExp]);
  double r = fma(-n, C[kExp + 1], thm);   // exact (Sterbenz region)
  r = r + (tlm - n * C[kExp + 2]);
  double p = C[kExp + 3];               // 1/14!
  p = horner(p, r, C[kExp + 4]);
  p = horner(p, r, C[kExp + 5]);
  p = horner(p, r, C[kExp + 6]);
  p = horner(p, r, C[kExp + 7]);
  p = horner(p, r, C[kExp + 8]);
  p = horner(p, r, C[kExp + 9]);
  p = horner(p, r, C[kExp + 10]);
  p = horner(p, r, C[kExp + 11]);
  p = horner(p, r, C[kExp + 12]);
  p = horner(p, r, C[kExp + 13]);
  p = horner(p, r, C[kExp + 14]);
  p = horner(p, r, C[kExp + 15]);
  double em1 = fma(r * r, p, r);
  double res = 1.0 + em1;
  const int ni = (int)n;
  // Two-step scaling keeps 2^ni representable for results near the overflow and underflow
  // thresholds; for |ni| <= 1000 the second step is a scaling by 2^0 and is skipped, as are the
  // special values, behind one branch that ordinary arguments never take (same values).
  const int n1 = ni > 1000 ? 1000 : (ni < -1000 ? -1000 : ni);
  double out = ldexp(res, n1);
  if (!regular || n1 != ni) {
    asm volatile("");  // a branch, not selects evaluated for every argument
    out = ldexp(out, ni - n1);
    out = over ? infinity() : out;
    out = under ? 0.0 : out;
    out = nan ? quiet_nan() : out;
  }
  return out;
}

BEX_HD bool is_integer(double y) { return rint(y) == y; }
BEX_HD bool is_odd_integer(double y) {
  const double h = y * 0.5;
  return is_integer(y) & (fabs_d(y) < 0x1.0p53) & (rint(h) != h);
}

// pow(x, y) with the C99 special-case rules (F.9.4.4), in this order of precedence:
// y = 0 -> 1; x = 1 -> 1; NaN -> NaN; y = +-inf; x < 0 with non-integer y -> NaN; x = +-0;
// x = +-inf; otherwise exp(y log|x|), negated for x < 0 (or x = -0) and odd integer y.
BEX_HD double pow(double x, double y) {
  const double ax = fabs_d(x);
  const bool finite_x = is_finite(ax);
  const bool finite_y = is_finite(y);
  const bool negative = (as_bits(x) >> 63) != 0;  // x < 0 or x = -0
  const double sign = (negative && is_odd_integer(y)) ? -1.0 : 1.0;
  // Main path on a finite, nonzero |x| and a finite y (1 and 0 otherwise).
  const bool regular = finite_x && (ax != 0.0) && finite_y;
  const double axm = regular ? ax : 1.0;
  const double ym = regular ? y : 0.0;
  double ll;
  const double lh = log_dd(axm, &ll);
  const double th = ym * lh;
  const double tl = fma(ym, lh, -th) + ym * ll;
  double r = sign * exp_dd(th, tl);
  // Special cases, lowest precedence first. On a finite nonzero x and a finite y, other than a
  // negative x with a non-integer y, none of them changes r (x = 1 and y = 0 give exactly 1 on
  // the main path), so they sit behind one branch that ordinary arguments never take.
  if (!regular || (negative && !is_integer(y))) {
    asm volatile("");  // a branch, not selects evaluated for every argument
    r = !finite_x ? ((y > 0.0) ? sign * infinity() : sign * 0.0) : r;
    r = (ax == 0.0) ? ((y > 0.0) ? sign * 0.0 : sign * infinity()) : r;
    r = (negative && (x != 0.0) && finite_x && !is_integer(y)) ? quiet_nan() : r;
    const double at_inf = (ax == 1.0) ? 1.0 : (((ax > 1.0) == (y > 0.0)) ? infinity() : 0.0);
    r = !finite_y ? at_inf : r;
    r = (is_nan(x) || is_nan(y)) ? quiet_nan() : r;
    r = (x == 1.0) ? 1.0 : r;
    r = (y == 0.0) ? 1.0 : r;
  }
  return r;
}

}  // namespace bex
