// Correctly rounded FP64 elementary functions for host and device.
//
// The reference evaluates its profile and wind formulas with numpy and glibc
// (src/bldfm/pbl_model.py:58-250, src/bldfm/utils.py:23-27).  Those
// libraries return the correctly rounded value for all but a small fraction
// of arguments, while the GPU's libm is accurate to about one ulp; and the
// shooting method amplifies a one-ulp change of a profile value into a 1e-8
// change of a stable footprint.  These functions therefore compute each
// result to about 100 bits in double-double arithmetic and round once at the
// end, which gives the correctly rounded value except when the exact result
// lies within about 2^-100 (relative) of a rounding midpoint.
//
// Method: double-double arithmetic (T. J. Dekker, "A floating-point technique
// for extending the available precision", Numer. Math. 18, 1971; the
// error-free transformations as in Hida, Li and Bailey's QD library), with
//  * exp:   x = k ln2 + r (ln2 in three parts), r scaled by 2^-8, Taylor series
//           of expm1 to degree 11, eight squarings of 1 + expm1;
//  * log:   one Newton step y + (x - e^y)/e^y from the libm estimate, with the
//           double-double exp; near 1, 2 atanh(d/(2+d)) by its series;
//  * log1p: the same on the double-double sum 1 + d;
//  * atan:  reflection about 1 and pi/4, three half-angle reductions
//           u/(1 + sqrt(1 + u^2)), Taylor series;
//  * sin/cos: reduction by pi/2 in three parts (Cody-Waite), Taylor series.
// Only IEEE-754 basic operations, fma and sqrt (all correctly rounded on the
// CPU and on gfx950) are used after the libm starting estimate, and nothing
// is contracted, so host and device builds agree bit for bit.
#pragma once
#include <math.h>
#include "bldfm_types.h"

struct ddouble {
  double hi, lo;
};

BLDFM_HD ddouble dd_make(double hi, double lo) { ddouble r; r.hi = hi; r.lo = lo; return r; }

// Every function body below starts with CR_EXACT: no contraction into fma.
#define CR_EXACT _Pragma("clang fp contract(off)")

// a + b = s + e exactly (Knuth)
BLDFM_HD ddouble dd_two_sum(double a, double b) {
  CR_EXACT
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return dd_make(s, e);
}
// |a| >= |b| (Dekker)
BLDFM_HD ddouble dd_fast_two_sum(double a, double b) {
  CR_EXACT
  const double s = a + b;
  const double e = b - (s - a);
  return dd_make(s, e);
}
BLDFM_HD ddouble dd_two_prod(double a, double b) {
  CR_EXACT
  const double p = a * b;
  return dd_make(p, fma(a, b, -p));
}
BLDFM_HD ddouble dd_neg(ddouble a) { return dd_make(-a.hi, -a.lo); }
BLDFM_HD ddouble dd_add(ddouble a, ddouble b) {
  CR_EXACT
  ddouble s = dd_two_sum(a.hi, b.hi);
  const ddouble t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}
BLDFM_HD ddouble dd_add_d(ddouble a, double b) {
  CR_EXACT
  ddouble s = dd_two_sum(a.hi, b);
  s.lo += a.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}
BLDFM_HD ddouble dd_sub(ddouble a, ddouble b) { return dd_add(a, dd_neg(b)); }
BLDFM_HD ddouble dd_mul(ddouble a, ddouble b) {
  CR_EXACT
  ddouble p = dd_two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return dd_fast_two_sum(p.hi, p.lo);
}
BLDFM_HD ddouble dd_mul_d(ddouble a, double b) {
  CR_EXACT
  ddouble p = dd_two_prod(a.hi, b);
  p.lo += a.lo * b;
  return dd_fast_two_sum(p.hi, p.lo);
}
BLDFM_HD ddouble dd_div(ddouble a, ddouble b) {
  CR_EXACT
  const double q1 = a.hi / b.hi;
  ddouble r = dd_sub(a, dd_mul_d(b, q1));
  const double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  const double q3 = r.hi / b.hi;
  return dd_add_d(dd_fast_two_sum(q1, q2), q3);
}
BLDFM_HD ddouble dd_sqrt(ddouble a) {
  CR_EXACT
  const double x = sqrt(a.hi);
  const ddouble r = dd_sub(a, dd_two_prod(x, x));
  return dd_fast_two_sum(x, r.hi / (2.0 * x));
}
// hi + lo rounded to the nearest double
BLDFM_HD double dd_round(ddouble a) { return a.hi + a.lo; }

// ln 2 = LN2_HI + LN2_MID + LN2_LO; pi/2 in 33-bit parts (k * part exact for
// |k| < 2^20) and as a double-double; pi/4 as a double-double.
#define CR_LN2_HI 0x1.62e42fefa39efp-1
#define CR_LN2_MID 0x1.abc9e3b39803fp-56
#define CR_LN2_LO 0x1.7b57a079a1934p-111
#define CR_INV_LN2 0x1.71547652b82fep+0
#define CR_PIO2_1 0x1.921fb544p+0
#define CR_PIO2_2 0x1.0b4611a6p-34
#define CR_PIO2_3 0x1.3198a2e037073p-69
#define CR_PIO2_HI 0x1.921fb54442d18p+0
#define CR_PIO2_LO 0x1.1a62633145c07p-54
#define CR_PIO4_HI 0x1.921fb54442d18p-1
#define CR_PIO4_LO 0x1.1a62633145c07p-55
#define CR_TWO_OVER_PI 0x1.45f306dc9c883p-1

// 1/n as double-doubles (hi, lo), n = 1..32: the series coefficients, so that
// no series step needs a double-double division.  In the unrolled loops below
// every index is a compile-time constant and the values become immediates.
static constexpr double CR_INV[64] = {
    0x1.0000000000000p+0, 0x0.0p+0,
    0x1.0000000000000p-1, 0x0.0p+0,
    0x1.5555555555555p-2, 0x1.5555555555555p-56,
    0x1.0000000000000p-2, 0x0.0p+0,
    0x1.999999999999ap-3, -0x1.999999999999ap-57,
    0x1.5555555555555p-3, 0x1.5555555555555p-57,
    0x1.2492492492492p-3, 0x1.2492492492492p-57,
    0x1.0000000000000p-3, 0x0.0p+0,
    0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58,
    0x1.999999999999ap-4, -0x1.999999999999ap-58,
    0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59,
    0x1.5555555555555p-4, 0x1.5555555555555p-58,
    0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58,
    0x1.2492492492492p-4, 0x1.2492492492492p-58,
    0x1.1111111111111p-4, 0x1.1111111111111p-60,
    0x1.0000000000000p-4, 0x0.0p+0,
    0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61,
    0x1.c71c71c71c71cp-5, 0x1.c71c71c71c71cp-59,
    0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59,
    0x1.999999999999ap-5, -0x1.999999999999ap-59,
    0x1.8618618618618p-5, 0x1.8618618618618p-59,
    0x1.745d1745d1746p-5, -0x1.745d1745d1746p-60,
    0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60,
    0x1.5555555555555p-5, 0x1.5555555555555p-59,
    0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61,
    0x1.3b13b13b13b14p-5, -0x1.3b13b13b13b14p-59,
    0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59,
    0x1.2492492492492p-5, 0x1.2492492492492p-59,
    0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61,
    0x1.1111111111111p-5, 0x1.1111111111111p-61,
    0x1.0842108421084p-5, 0x1.0842108421084p-60,
    0x1.0000000000000p-5, 0x0.0p+0,
};
BLDFM_HD ddouble dd_inv(int n) { return dd_make(CR_INV[2 * (n - 1)], CR_INV[2 * (n - 1) + 1]); }

// e^x as a double-double (|x| <= 700).
BLDFM_HD ddouble dd_exp(double x) {
  CR_EXACT
  const double k = rint(x * CR_INV_LN2);
  // r = x - k ln2: x - k*LN2_HI is formed from the exact product
  const ddouble p1 = dd_two_prod(k, CR_LN2_HI);
  const ddouble p2 = dd_two_prod(k, CR_LN2_MID);
  ddouble r = dd_two_sum(x, -p1.hi);
  r = dd_add_d(r, -p1.lo);
  r = dd_sub(r, p2);
  r = dd_add_d(r, -(k * CR_LN2_LO));
  // expm1(s), s = r / 256, by its Taylor series (|s| < 1.4e-3)
  const ddouble s = dd_make(r.hi * 0x1p-8, r.lo * 0x1p-8);
  ddouble e = dd_make(0.0, 0.0);
  // Horner: e = s(1 + s/2(1 + s/3(...)))
#pragma unroll
  for (int n = 11; n >= 1; --n) e = dd_mul(dd_add_d(dd_mul(e, s), 1.0), dd_inv(n));
  e = dd_mul(e, s);  // expm1(s) = s (1 + s/2 (...)) with the 1/n folded in
  // undo the scaling: (1 + e)^2 = 1 + (2e + e^2), eight times
  for (int i = 0; i < 8; ++i) e = dd_add(dd_make(2.0 * e.hi, 2.0 * e.lo), dd_mul(e, e));
  ddouble y = dd_add_d(e, 1.0);
  const int ki = (int)k;
  return dd_make(ldexp(y.hi, ki), ldexp(y.lo, ki));
}

// 2 atanh(u) = u (2 + 2u^2/3 + 2u^4/5 + ...) for small |u| (double-double u)
// (|u| < 0.032: 14 terms reach 2^-106 relative)
BLDFM_HD ddouble dd_atanh2_series(ddouble u) {
  CR_EXACT
  const ddouble u2 = dd_mul(u, u);
  ddouble s = dd_make(0.0, 0.0);
#pragma unroll
  for (int n = 13; n >= 0; --n) {
    const ddouble c = dd_inv(2 * n + 1);
    s = dd_add(dd_mul(s, u2), dd_make(2.0 * c.hi, 2.0 * c.lo));
  }
  return dd_mul(s, u);
}

// log of a positive double-double w, given a starting estimate y0 close to it.
BLDFM_HD ddouble dd_log_newton(ddouble w, double y0) {
  CR_EXACT
  const ddouble e = dd_exp(y0);
  const ddouble d = dd_sub(w, e);
  return dd_add_d(dd_make(y0, 0.0), d.hi / e.hi + (d.lo - (d.hi / e.hi) * e.lo) / e.hi);
}

BLDFM_HD double cr_exp(double x) {
  CR_EXACT
  if (!(x == x)) return x;
  if (x > 709.78) return INFINITY;
  if (x < -708.0) return exp(x);  // subnormal range: not used by the profile formulas
  return dd_round(dd_exp(x));
}

BLDFM_HD double cr_log(double x) {
  CR_EXACT
  if (!(x > 0.0) || x == INFINITY) return log(x);
  if (fabs(x - 1.0) < 0.0625) {
    // d = x - 1 is exact here; log x = 2 atanh(d/(2 + d))
    const double d = x - 1.0;
    if (d == 0.0) return 0.0;
    const ddouble u = dd_div(dd_make(d, 0.0), dd_two_sum(2.0, d));
    return dd_round(dd_atanh2_series(u));
  }
  return dd_round(dd_log_newton(dd_make(x, 0.0), log(x)));
}

BLDFM_HD double cr_log1p(double d) {
  CR_EXACT
  if (!(d > -1.0) || d == INFINITY) return log1p(d);
  if (d == 0.0) return d;
  if (fabs(d) < 0.0625) {
    const ddouble u = dd_div(dd_make(d, 0.0), dd_two_sum(2.0, d));
    return dd_round(dd_atanh2_series(u));
  }
  return dd_round(dd_log_newton(dd_two_sum(1.0, d), log1p(d)));
}

BLDFM_HD double cr_atan(double x) {
  CR_EXACT
  if (!(x == x)) return x;
  const bool neg = x < 0.0;
  const double ax = fabs(x);
  if (ax == 0.0) return x;
  if (ax > 1e17) return neg ? -CR_PIO2_HI : CR_PIO2_HI;
  // t in [0, 1] with atan(ax) = base + sgn * atan(t)
  ddouble base = dd_make(0.0, 0.0);
  double sgn = 1.0;
  ddouble t = dd_make(ax, 0.0);
  if (ax > 1.0) {
    t = dd_div(dd_make(1.0, 0.0), t);
    base = dd_make(CR_PIO2_HI, CR_PIO2_LO);
    sgn = -1.0;
  }
  if (t.hi > 0.4142135623730950) {  // atan t = pi/4 + atan((t - 1)/(t + 1))
    const ddouble u = dd_div(dd_add_d(t, -1.0), dd_add_d(t, 1.0));
    base = dd_add(base, dd_make(sgn * CR_PIO4_HI, sgn * CR_PIO4_LO));
    t = u;
  }
  // three half-angle steps: atan u = 2 atan(u / (1 + sqrt(1 + u^2)))
  for (int i = 0; i < 3; ++i) t = dd_div(t, dd_add_d(dd_sqrt(dd_add_d(dd_mul(t, t), 1.0)), 1.0));
  // atan u = u - u^3/3 + u^5/5 - ...  (|u| < 0.052)
  const ddouble u2 = dd_mul(t, t);
  ddouble s = dd_make(0.0, 0.0);
#pragma unroll
  for (int n = 13; n >= 0; --n) {
    const ddouble c = dd_inv(2 * n + 1);
    s = dd_add(dd_mul(s, u2), n % 2 ? dd_neg(c) : c);
  }
  s = dd_mul(s, t);
  const ddouble r = dd_add(base, dd_make(8.0 * sgn * s.hi, 8.0 * sgn * s.lo));
  const double v = dd_round(r);
  return neg ? -v : v;
}

// sin and cos of |x| < 2^19 * pi/2 (wind directions in radians).
BLDFM_HD void cr_sincos(double x, double* sn, double* cs) {
  CR_EXACT
  if (!(fabs(x) < 823549.0)) {
    *sn = sin(x);
    *cs = cos(x);
    return;
  }
  const double k = rint(x * CR_TWO_OVER_PI);
  // r = x - k pi/2: the first two products are exact
  ddouble r = dd_two_sum(x - k * CR_PIO2_1, -(k * CR_PIO2_2));
  r = dd_add(r, dd_neg(dd_two_prod(k, CR_PIO2_3)));
  // sin r = r S, S = 1 - r^2/(2*3) (1 - r^2/(4*5) (1 - ...));
  // cos r = C, C = 1 - r^2/(1*2) (1 - r^2/(3*4) (1 - ...))  (|r| <= pi/4, 15 terms)
  const ddouble r2 = dd_mul(r, r);
  ddouble s = dd_make(1.0, 0.0), c = dd_make(1.0, 0.0);
#pragma unroll
  for (int n = 15; n >= 1; --n) {
    s = dd_add_d(dd_neg(dd_mul(dd_mul(dd_mul(r2, s), dd_inv(2 * n)), dd_inv(2 * n + 1))), 1.0);
    c = dd_add_d(dd_neg(dd_mul(dd_mul(dd_mul(r2, c), dd_inv(2 * n - 1)), dd_inv(2 * n))), 1.0);
  }
  s = dd_mul(s, r);
  const int q = ((int)fmod(k, 4.0) + 4) % 4;
  const double sv = dd_round(s), cv = dd_round(c);
  if (q == 0) { *sn = sv; *cs = cv; }
  else if (q == 1) { *sn = cv; *cs = -sv; }
  else if (q == 2) { *sn = -sv; *cs = -cv; }
  else { *sn = -cv; *cs = sv; }
}


// ------------------------------------------------ glibc-compatible functions
// numpy's complex power, which the reference's psi and phi call
// (src/bldfm/pbl_model.py:222,248), is glibc's cpow = cexp(p clog(v)).  glibc's
// exp and log1p are not correctly rounded (log1p is off by one ulp for about
// 9% of arguments in (0, 3)), so these two follow glibc's own algorithms.

BLDFM_HD uint64_t bits_of(double x) { uint64_t b; __builtin_memcpy(&b, &x, 8); return b; }
BLDFM_HD double double_of(uint64_t b) { double x; __builtin_memcpy(&x, &b, 8); return x; }

// 2^(k/128) = H (1 + T) for 0 <= k < 128: H the nearest double, T the rounded
// relative remainder (the table of the exp below, computed instead of stored).
BLDFM_HD void gl_exp_table(int k, double* H, double* T) {
  CR_EXACT
  // v = e^{(k/128) ln 2} to ~100 bits: the argument is exact in double-double
  const ddouble a = dd_mul_d(dd_make(CR_LN2_HI, CR_LN2_MID), (double)k * 0x1p-7);
  const ddouble s = dd_make(a.hi * 0x1p-8, a.lo * 0x1p-8);
  ddouble e = dd_make(0.0, 0.0);
#pragma unroll
  for (int n = 11; n >= 1; --n) e = dd_mul(dd_add_d(dd_mul(e, s), 1.0), dd_inv(n));
  e = dd_mul(e, s);
  for (int i = 0; i < 8; ++i) e = dd_add(dd_make(2.0 * e.hi, 2.0 * e.lo), dd_mul(e, e));
  const ddouble v = dd_add_d(e, 1.0);
  *H = dd_round(v);
  *T = dd_round(dd_div(dd_add_d(v, -*H), dd_make(*H, 0.0)));
}

// glibc's exp (sysdeps/ieee754/dbl-64/e_exp.c, the ARM optimized-routines
// algorithm by Szabolcs Nagy, as built with fused multiply-adds on x86-64):
// x = k ln2/128 + r, exp x = 2^(k/128) (1 + T_k + r + r^2 (C2 + r C3) + r^4 (C4 + r C5)).
// The steps and the constants below are those of optimized-routines'
// math/exp.c, Copyright (c) 2018 Arm Limited,
// SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception; they are
// reproduced because the result must round exactly as the host's libm does.
BLDFM_HD double gl_exp(double x) {
  CR_EXACT
  const double ax = fabs(x);
  if (!(ax < 512.0)) return cr_exp(x);  // glibc's scaled special case; not reached by the profiles
  if (ax < 0x1p-54) return 1.0 + x;
  const double InvLn2N = 0x1.71547652b82fep0 * 128.0, NegLn2hiN = -0x1.62e42fefa0000p-8,
               NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  const double z = InvLn2N * x;
  double kd = z + Shift;
  kd -= Shift;
  const double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  const int ki = (int)kd;
  const int k = ki & 127, m = (ki - k) / 128;
  double H, T;
  gl_exp_table(k, &H, &T);
  const double r2 = r * r;
  const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), T + r));
  const double scale = ldexp(H, m);
  return fma(scale, tmp, scale);
}

// glibc's log1p (sysdeps/ieee754/dbl-64/s_log1p.c, from Sun's fdlibm):
// 1 + x = 2^k (1 + f) with sqrt(2)/2 < 1 + f < sqrt(2), s = f/(2 + f),
// log1p(f) = f - (f^2/2 - s (f^2/2 + R(s^2))), R a degree-7 polynomial in
// s^2 evaluated in glibc's split order, plus k ln2 and the rounding
// correction c/u of 1 + x.  The steps and the constants are fdlibm's
// (reproduced for the same reason as the exp above), under its notice:
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
BLDFM_HD double gl_log1p(double x) {
  CR_EXACT
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double Lp1 = 6.666666666666735130e-01, Lp2 = 3.999999999940941908e-01,
               Lp3 = 2.857142874366239149e-01, Lp4 = 2.222219843214978396e-01,
               Lp5 = 1.818357216161805012e-01, Lp6 = 1.531383769920937332e-01,
               Lp7 = 1.479819860511658591e-01;
  double f = 0.0, c = 0.0, u;
  const int32_t hx = (int32_t)(bits_of(x) >> 32);
  const int32_t ax = hx & 0x7fffffff;
  int32_t k = 1, hu = 0;
  if (hx < 0x3FDA827A) {                 // x < 0.41422
    if (ax >= 0x3ff00000) return log1p(x);  // x <= -1
    if (ax < 0x3e200000) {               // |x| < 2^-29
      if (ax < 0x3c900000) return x;
      return x - x * x * 0.5;
    }
    if (hx > 0 || hx <= (int32_t)0xbfd2bec4) {  // -0.2929 < x < 0.41422
      k = 0;
      f = x;
      hu = 1;
    }
  } else if (hx >= 0x7ff00000) {
    return x + x;
  }
  if (k != 0) {
    if (hx < 0x43400000) {
      u = 1.0 + x;
      hu = (int32_t)(bits_of(u) >> 32);
      k = (hu >> 20) - 1023;
      c = (k > 0) ? 1.0 - (u - x) : x - (u - 1.0);
      c /= u;
    } else {
      u = x;
      hu = (int32_t)(bits_of(u) >> 32);
      k = (hu >> 20) - 1023;
      c = 0.0;
    }
    hu &= 0x000fffff;
    const uint64_t lo = bits_of(u) & 0xffffffffull;
    if (hu < 0x6a09e) {
      u = double_of(((uint64_t)(uint32_t)(hu | 0x3ff00000) << 32) | lo);
    } else {
      k += 1;
      u = double_of(((uint64_t)(uint32_t)(hu | 0x3fe00000) << 32) | lo);
      hu = (0x00100000 - hu) >> 2;
    }
    f = u - 1.0;
  }
  const double hfsq = 0.5 * f * f;
  if (hu == 0) {  // |f| < 2^-20
    if (f == 0.0) {
      if (k == 0) return 0.0;
      c += k * ln2_lo;
      return k * ln2_hi + c;
    }
    const double R = hfsq * (1.0 - 0.66666666666666666 * f);
    if (k == 0) return f - R;
    return k * ln2_hi - ((R - (k * ln2_lo + c)) - f);
  }
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double R1 = z * Lp1, z2 = z * z;
  const double R2 = Lp2 + z * Lp3, z4 = z2 * z2;
  const double R3 = Lp4 + z * Lp5, z6 = z4 * z2;
  const double R4 = Lp6 + z * Lp7;
  const double R = R1 + z2 * R2 + z4 * R3 + z6 * R4;
  if (k == 0) return f - (hfsq - s * (hfsq + R));
  return k * ln2_hi - ((hfsq - (s * (hfsq + R) + (k * ln2_lo + c))) - f);
}

// Real part of glibc's cpow(v + 0i, p + 0i) for v >= 1 (numpy's complex
// power of a real base, src/bldfm/pbl_model.py:222,248): cexp(p clog(v)),
// where clog's real part is log1p((v - 1)(v + 1))/2 for 1 <= v < 2 and
// log(hypot(v, 0)) = log v above; the imaginary parts are exact zeros.
BLDFM_HD double gl_cpow_real(double v, double p) {
  CR_EXACT
  double lr;
  if (v >= 1.0 && v < 2.0) lr = gl_log1p((v - 1.0) * (v + 1.0)) / 2.0;
  else lr = cr_log(v);
  return gl_exp(p * lr);
}
