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

// A polynomial coefficient or splitting constant. On the device the value is pinned into a
// scalar register pair right where it is used: without the opaque move, loop-invariant code
// motion lifts every coefficient of the inlined functions out of the solver's time loop into
// vector registers, and the register allocator then spills them (and the coupling rows) to
// scratch memory on every stage. The empty asm takes the function's varying argument `dep` as
// an input, so it cannot be hoisted, yet it is not volatile, so the scheduler may interleave
// independent evaluations (pow, sincospi, sincos of one right-hand side). The value itself is
// unchanged.
BEX_HD double kc(double c, double dep) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+s"(c) : "v"(dep));
#else
  (void)dep;
#endif
  return c;
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
BEX_HD double sinpi_tail(double r, double r2) {
  double p = kc(0x1.2877020d52cf0p-31, r2);
  p = fma(p, r2, kc(-0x1.8a404211f9547p-26, r2));
  p = fma(p, r2, kc(0x1.aaec32af93359p-21, r2));
  p = fma(p, r2, kc(-0x1.6fadb9f155744p-16, r2));
  p = fma(p, r2, kc(0x1.e8f434d018d63p-12, r2));
  p = fma(p, r2, kc(-0x1.e3074fde8871fp-8, r2));
  p = fma(p, r2, kc(0x1.50783487ee782p-4, r2));
  p = fma(p, r2, kc(-0x1.32d2cce62bd86p-1, r2));
  p = fma(p, r2, kc(0x1.466bc6775aae2p+1, r2));
  p = fma(p, r2, kc(-0x1.4abbce625be53p+2, r2));
  return (r * r2) * p;
}

// cos(pi*r) - 1, for |r| <= 1/4, as r^2 * Q(r^2).
BEX_HD double cospi_tail(double r2) {
  double q = kc(0x1.ef6e308d6d1c4p-29, r2);
  q = fma(q, r2, kc(-0x1.2a0c591af8314p-23, r2));
  q = fma(q, r2, kc(0x1.20c62c2f2d7f5p-18, r2));
  q = fma(q, r2, kc(-0x1.b6e24f44b128fp-14, r2));
  q = fma(q, r2, kc(0x1.f9d38a3763cc3p-10, r2));
  q = fma(q, r2, kc(-0x1.a6d1f2a204a8cp-6, r2));
  q = fma(q, r2, kc(0x1.e1f506891babbp-3, r2));
  q = fma(q, r2, kc(-0x1.55d3c7e3cbffap+0, r2));
  q = fma(q, r2, kc(0x1.03c1f081b5ac4p+2, r2));
  q = fma(q, r2, kc(-0x1.3bd3cc9be45dep+2, r2));
  return r2 * q;
}

// sin(r) - r and cos(r) - 1 for |r| <= pi/4 (+ reduction slack).
BEX_HD double sin_tail(double r, double r2) {
  double p = kc(0x1.71b8ef6dcf572p-66, r2);
  p = fma(p, r2, kc(-0x1.2f49b46814157p-57, r2));
  p = fma(p, r2, kc(0x1.952c77030ad4ap-49, r2));
  p = fma(p, r2, kc(-0x1.ae7f3e733b81fp-41, r2));
  p = fma(p, r2, kc(0x1.6124613a86d09p-33, r2));
  p = fma(p, r2, kc(-0x1.ae64567f544e4p-26, r2));
  p = fma(p, r2, kc(0x1.71de3a556c734p-19, r2));
  p = fma(p, r2, kc(-0x1.a01a01a01a01ap-13, r2));
  p = fma(p, r2, kc(0x1.1111111111111p-7, r2));
  p = fma(p, r2, kc(-0x1.5555555555555p-3, r2));
  return (r * r2) * p;
}
BEX_HD double cos_tail(double r2) {
  double q = kc(-0x1.0ce396db7f853p-70, r2);
  q = fma(q, r2, kc(0x1.e542ba4020225p-62, r2));
  q = fma(q, r2, kc(-0x1.6827863b97d97p-53, r2));
  q = fma(q, r2, kc(0x1.ae7f3e733b81fp-45, r2));
  q = fma(q, r2, kc(-0x1.93974a8c07c9dp-37, r2));
  q = fma(q, r2, kc(0x1.1eed8eff8d898p-29, r2));
  q = fma(q, r2, kc(-0x1.27e4fb7789f5cp-22, r2));
  q = fma(q, r2, kc(0x1.a01a01a01a01ap-16, r2));
  q = fma(q, r2, kc(-0x1.6c16c16c16c17p-10, r2));
  q = fma(q, r2, kc(0x1.5555555555555p-5, r2));
  q = fma(q, r2, kc(-0x1.0000000000000p-1, r2));
  return r2 * q;
}

// ---------------------------------------------------------------------------------------
// sincospi(x) = (sin(pi x), cos(pi x)). The reduction x = n/2 + r is exact, so the result is
// accurate for every finite x.
BEX_HD void sincospi(double x, double* s, double* c) {
  if (!is_finite(x)) {
    *s = quiet_nan();
    *c = quiet_nan();
    return;
  }
  double ax = fabs_d(x);
  if (ax >= 0x1.0p52) {  // x is an integer: sin = 0, cos = (-1)^x
    double half = x * 0.5;
    bool odd = (floor(half) != half);
    *s = 0.0;
    *c = odd ? -1.0 : 1.0;
    return;
  }
  double n = rint(2.0 * x);      // exact: 2x is exact, rint is exact
  double r = fma(-0.5, n, x);    // exact: |r| <= 1/4
  double q4 = n - 4.0 * floor(n * 0.25);  // n mod 4, exact
  int q = (int)q4;
  double r2 = r * r;
  // sin(pi r) = r*pi_hi + (r*pi_lo + tail)
  double sp = fma(r, kc(0x1.921fb54442d18p+1, r), fma(r, kc(0x1.1a62633145c07p-53, r), sinpi_tail(r, r2)));
  double cp = 1.0 + cospi_tail(r2);
  switch (q) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
  }
}

// sin and cos with a three-part Cody-Waite reduction by pi/2. The reduction is exact for
// |x| < 2^20 * pi/2 (~1.6e6); beyond that the result stays finite but loses accuracy. The
// solver evaluates these at 2*pi*f2/f1*tau + theta with tau in [0, 1].
BEX_HD void reduce_pio2(double x, double* r, int* q) {
  double n = rint(x * kc(0x1.45f306dc9c883p-1, x));
  double t = fma(-n, kc(0x1.921fb54400000p+0, x), x);
  t = fma(-n, kc(0x1.0b4611a600000p-34, x), t);
  t = fma(-n, kc(0x1.3198a2e000000p-69, x), t);
  *r = t;
  double q4 = n - 4.0 * floor(n * 0.25);
  *q = (int)q4;
}

BEX_HD double sin(double x) {
  if (!is_finite(x)) return quiet_nan();
  double r;
  int q;
  reduce_pio2(x, &r, &q);
  double r2 = r * r;
  double sr = r + sin_tail(r, r2);
  double cr = 1.0 + cos_tail(r2);
  switch (q) {
    case 0: return sr;
    case 1: return cr;
    case 2: return -sr;
    default: return -cr;
  }
}

BEX_HD double cos(double x) {
  if (!is_finite(x)) return quiet_nan();
  double r;
  int q;
  reduce_pio2(x, &r, &q);
  double r2 = r * r;
  double sr = r + sin_tail(r, r2);
  double cr = 1.0 + cos_tail(r2);
  switch (q) {
    case 0: return cr;
    case 1: return -sr;
    case 2: return -cr;
    default: return sr;
  }
}

// sin and cos of the same argument; bit-identical to sin(x) and cos(x) above (same reduction
// and polynomials), at the cost of one of them.
BEX_HD void sincos(double x, double* s, double* c) {
  if (!is_finite(x)) {
    *s = quiet_nan();
    *c = quiet_nan();
    return;
  }
  double r;
  int q;
  reduce_pio2(x, &r, &q);
  double r2 = r * r;
  double sr = r + sin_tail(r, r2);
  double cr = 1.0 + cos_tail(r2);
  switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// ---------------------------------------------------------------------------------------
// log(x) for finite x > 0, returned as an unevaluated sum hi + lo carrying ~10 extra bits.
BEX_HD double log_dd(double x, double* lo) {
  int k = 0;
  if (x < 0x1.0p-1022) {  // subnormal input
    x *= 0x1.0p54;
    k = -54;
  }
  uint64_t b = as_bits(x);
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
  if (m > 0x1.6a09e667f3bcdp+0) {  // m > sqrt(2): move to [sqrt(1/2), sqrt(2))
    m *= 0.5;
    k += 1;
  }
  // log(m) = 2 atanh(s), s = f / (2 + f), f = m - 1 (exact).
  double f = m - 1.0;
  double dh = 2.0 + f;
  double dl = f - (dh - 2.0);                 // 2 + f = dh + dl exactly
  double rinv = 1.0 / dh;
  double sh = f * rinv;
  double e = fma(-sh, dh, f) - sh * dl;       // f - sh*(dh+dl)
  double sl = e * rinv;
  double s2 = sh * sh;
  double p = kc(0x1.2f684bda12f68p-5, s2);            // 1/27
  p = fma(p, s2, kc(0x1.47ae147ae147bp-5, s2));       // 1/25
  p = fma(p, s2, kc(0x1.642c8590b2164p-5, s2));       // 1/23
  p = fma(p, s2, kc(0x1.8618618618618p-5, s2));       // 1/21
  p = fma(p, s2, kc(0x1.af286bca1af28p-5, s2));       // 1/19
  p = fma(p, s2, kc(0x1.e1e1e1e1e1e1ep-5, s2));       // 1/17
  p = fma(p, s2, kc(0x1.1111111111111p-4, s2));       // 1/15
  p = fma(p, s2, kc(0x1.3b13b13b13b14p-4, s2));       // 1/13
  p = fma(p, s2, kc(0x1.745d1745d1746p-4, s2));       // 1/11
  p = fma(p, s2, kc(0x1.c71c71c71c71cp-4, s2));       // 1/9
  p = fma(p, s2, kc(0x1.2492492492492p-3, s2));       // 1/7
  p = fma(p, s2, kc(0x1.999999999999ap-3, s2));       // 1/5
  p = fma(p, s2, kc(0x1.5555555555555p-2, s2));       // 1/3
  double tail = (sh * s2) * p;                // s^3 (1/3 + s^2/5 + ...)
  double lm_hi = 2.0 * sh;
  double lm_lo = 2.0 * (sl + tail);
  // + k*ln2, ln2 split so that k*LN2_A is exact.
  double kd = (double)k;
  double a = kd * kc(0x1.62e42ff000000p-1, kd);
  double bsum = a + lm_hi;
  // two-sum of a and lm_hi
  double bb = bsum - a;
  double err = (a - (bsum - bb)) + (lm_hi - bb);
  double low = err + (lm_lo + kd * kc(-0x1.718432a1b0e26p-35, kd));
  double hi = bsum + low;
  *lo = low - (hi - bsum);
  return hi;
}

// exp(th + tl) for |tl| << |th|, th finite.
BEX_HD double exp_dd(double th, double tl) {
  if (th > 709.782712893384) return infinity();
  if (th < -745.1332191019412) return 0.0;
  double n = rint(th * kc(0x1.71547652b82fep+0, th));
  double r = fma(-n, kc(0x1.62e42ff000000p-1, th), th);   // exact (Sterbenz region)
  r = r + (tl - n * kc(-0x1.718432a1b0e26p-35, th));
  double p = kc(0x1.93974a8c07c9dp-37, r);               // 1/14!
  p = fma(p, r, kc(0x1.6124613a86d09p-33, r));
  p = fma(p, r, kc(0x1.1eed8eff8d898p-29, r));
  p = fma(p, r, kc(0x1.ae64567f544e4p-26, r));
  p = fma(p, r, kc(0x1.27e4fb7789f5cp-22, r));
  p = fma(p, r, kc(0x1.71de3a556c734p-19, r));
  p = fma(p, r, kc(0x1.a01a01a01a01ap-16, r));
  p = fma(p, r, kc(0x1.a01a01a01a01ap-13, r));
  p = fma(p, r, kc(0x1.6c16c16c16c17p-10, r));
  p = fma(p, r, kc(0x1.1111111111111p-7, r));
  p = fma(p, r, kc(0x1.5555555555555p-5, r));
  p = fma(p, r, kc(0x1.5555555555555p-3, r));
  p = fma(p, r, kc(0x1.0000000000000p-1, r));
  double em1 = fma(r * r, p, r);
  double res = 1.0 + em1;
  int ni = (int)n;
  // Two-step scaling keeps 2^ni representable for results near the overflow and
  // underflow thresholds.
  if (ni > 1000) return ldexp(ldexp(res, 1000), ni - 1000);
  if (ni < -1000) return ldexp(ldexp(res, -1000), ni + 1000);
  return ldexp(res, ni);
}

BEX_HD bool is_integer(double y) { return rint(y) == y; }
BEX_HD bool is_odd_integer(double y) {
  if (!is_integer(y)) return false;
  if (fabs_d(y) >= 0x1.0p53) return false;
  double h = y * 0.5;
  return rint(h) != h;
}

// pow(x, y) with the C99 special-case rules.
BEX_HD double pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (is_nan(x) || is_nan(y)) return quiet_nan();
  double ax = fabs_d(x);
  if (!is_finite(y)) {
    if (ax == 1.0) return 1.0;
    bool big = ax > 1.0;
    return (big == (y > 0.0)) ? infinity() : 0.0;
  }
  double sign = 1.0;
  if (x < 0.0 || (x == 0.0 && (as_bits(x) >> 63))) {
    if (x != 0.0 && is_finite(x) && !is_integer(y)) return quiet_nan();
    if (is_odd_integer(y)) sign = -1.0;
  }
  if (ax == 0.0) return (y > 0.0) ? sign * 0.0 : sign * infinity();
  if (!is_finite(ax)) return (y > 0.0) ? sign * infinity() : sign * 0.0;
  double ll;
  double lh = log_dd(ax, &ll);
  double th = y * lh;
  double tl = fma(y, lh, -th) + y * ll;
  return sign * exp_dd(th, tl);
}

}  // namespace bex
