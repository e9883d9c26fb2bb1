/* eh_math.h -- FP64 sin, cos, tan, asin and log used identically on the host and
 * on gfx950.
 *
 * Why this exists: the tracking maps call transcendental functions (CSBEND's
 * exact curvilinear drift uses asin, sin and cos per sub-step; RFCA uses sin and
 * log).  The C library on the host and the device math library on the GPU do not
 * return bit-identical results, so a particle near an aperture edge could be lost
 * on one side and kept on the other.  Both the CPU checker and the HIP kernels
 * therefore call these routines, which use only IEEE-754 double +, -, *, /, sqrt
 * and fma (all correctly rounded on both targets) in a fixed order.  Every file that
 * includes this header must be compiled without floating-point contraction
 * (-ffp-contract=off) and without fast-math.
 *
 * Methods (all published, none copied from a library):
 *  - sin/cos: Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software
 *    Manual for the Elementary Functions", 1980) with pi/2 split into four
 *    24-bit pieces plus a tail, so k*piece is exact for |k| < 2^29; the reduced
 *    argument |r| <= ~pi/4 is kept as r + rl (the rounding error of r in rl)
 *    and fed to the Taylor series of sin (to r^21) and cos (to r^20), evaluated
 *    by Horner's rule, plus the first-order terms in rl.
 *  - asin: Taylor series sum_n (2n)!/(4^n (n!)^2 (2n+1)) x^(2n+1) for |x| <= 1/2
 *    (10 terms for |x| <= 1/8, 40 terms above), and asin(x) = pi/2 - 2 asin(
 *    sqrt((1-|x|)/2)) for |x| > 1/2.
 *  - log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *    s = (m-1)/(m+1), atanh series to s^25, plus e*ln2 split in hi/lo parts;
 *    for x = 1 + f with |f| < 1/16 the Taylor series of log(1+f) to f^16.
 *  - cos and log(1+f) form their leading terms with exact rounding errors
 *    (Dekker's Fast2Sum; the product error through fma), which makes them
 *    correctly rounded almost always in the small-argument regime of the
 *    tracking maps, i.e. equal to a correctly rounded C library there.
 * Constants come from eh_math_consts.h, generated from exact rational arithmetic
 * by gen_math_consts.py.
 *
 * Accuracy: a few ulp against correctly rounded results for |x| <= 2^29 (sin,
 * cos); beyond 2^50 (or for non-finite x) sin and cos return NaN.
 */
#ifndef EH_MATH_H
#define EH_MATH_H

#include "eh_math_consts.h"

#if defined(__HIPCC__)
#define EH_FN static __host__ __device__ inline
#else
#define EH_FN static inline
#endif

#ifdef __cplusplus
#include <cstdint>
#else
#include <stdint.h>
#endif

EH_FN uint64_t eh_bits(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}

EH_FN double eh_from_bits(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof x);
  return x;
}

EH_FN double eh_fabs(double x) { return eh_from_bits(eh_bits(x) & 0x7fffffffffffffffULL); }

EH_FN double eh_nan(void) { return eh_from_bits(0x7ff8000000000000ULL); }

EH_FN double eh_sqrt(double x) { return __builtin_sqrt(x); }

/* fused multiply-add: exactly specified by IEEE-754 (one rounding), hence
 * identical on every conforming host and on the GPU; used for the exact rounding
 * errors of products.  (Horner steps stay a multiply and an add: with fused
 * steps the compiler keeps the coefficients in vector registers across the
 * tracking kernels' loops, which then spill.)  The host build of this header
 * must compile it to the fused instruction (gcc/clang -mfma) or to the C
 * library's fma, which is correctly rounded too. */
EH_FN double eh_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

/* Taylor polynomial of sin(r) for 2^-4 <= |r| < pi/4 (to r^21) */
EH_FN double eh_sin_poly10(double r) {
  double z = r * r;
  double p = EH_SIN_10;
  p = EH_SIN_9 + z * p;
  p = EH_SIN_8 + z * p;
  p = EH_SIN_7 + z * p;
  p = EH_SIN_6 + z * p;
  p = EH_SIN_5 + z * p;
  p = EH_SIN_4 + z * p;
  p = EH_SIN_3 + z * p;
  p = EH_SIN_2 + z * p;
  p = EH_SIN_1 + z * p;
  return eh_fma(r * z, p, r);
}

/* Taylor polynomial of sin(r) for |r| < 2^-4, cut after the r^11 term (the
 * next one is below 2^-80 of the result) */
EH_FN double eh_sin_poly(double r) {
  double z = r * r;
  double q = EH_SIN_5;
  q = EH_SIN_4 + z * q;
  q = EH_SIN_3 + z * q;
  q = EH_SIN_2 + z * q;
  q = EH_SIN_1 + z * q;
  return eh_fma(r * z, q, r);
}

/* sin(r + rl) for a reduced argument carried as r + rl (|r| <= ~pi/4,
 * |rl| <= ulp(r)/2), for the arguments that needed a reduction.  The series is
 * r + r^3 (S1 + S2 z + z^2 Q(z)) with z = r^2: r^2, r^3, S1, S2 and the products
 * that carry the leading terms are kept with their exact rounding errors
 * (double-double by fma), Q in plain doubles, so that only the final addition
 * rounds at the level of the result; rl enters as rl cos(r) = rl (1 - z/2). */
EH_FN double eh_sin_poly_l(double r, double rl) {
  const double z = r * r, zl = eh_fma(r, r, -z);
  double q = EH_SIN_10;
  q = EH_SIN_9 + z * q;
  q = EH_SIN_8 + z * q;
  q = EH_SIN_7 + z * q;
  q = EH_SIN_6 + z * q;
  q = EH_SIN_5 + z * q;
  q = EH_SIN_4 + z * q;
  q = EH_SIN_3 + z * q;
  /* S = S1 + z S2 + z^2 q as sh + sl */
  const double t2 = z * EH_SIN_2;
  const double t2e = eh_fma(z, EH_SIN_2, -t2) + (zl * EH_SIN_2 + z * EH_SIN_2_LO);
  const double sh = EH_SIN_1 + t2;
  const double sl = ((EH_SIN_1 - sh) + t2) + (EH_SIN_1_LO + t2e + (z * z) * q);
  /* r^3 = c3 + c3l */
  const double c3 = r * z, c3l = eh_fma(r, z, -c3) + r * zl;
  /* r^3 S = T + te */
  const double T = c3 * sh;
  const double te = eh_fma(c3, sh, -T) + (c3 * sl + c3l * sh);
  const double h = r + T;
  const double hl = (r - h) + T;
  return h + (hl + (te + rl * (1.0 - 0.5 * z)));
}

/* Taylor polynomial of cos(r) for 2^-4 <= |r| < pi/4 (to r^20), leading terms
 * as in eh_cos_poly */
EH_FN double eh_cos_poly10(double r) {
  double z = r * r;
  double p = EH_COS_10;
  p = EH_COS_9 + z * p;
  p = EH_COS_8 + z * p;
  p = EH_COS_7 + z * p;
  p = EH_COS_6 + z * p;
  p = EH_COS_5 + z * p;
  p = EH_COS_4 + z * p;
  p = EH_COS_3 + z * p;
  p = EH_COS_2 + z * p;
  double zl = eh_fma(r, r, -z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  double t = (1.0 - w) - hz;
  return w + eh_fma(z * z, p, t - 0.5 * zl);
}

/* Taylor polynomial of cos(r) for |r| < 2^-4, cut after the r^12 term (the
 * next one is below 2^-76 of the result) */
EH_FN double eh_cos_poly(double r) {
  double z = r * r;
  double p = EH_COS_6;
  p = EH_COS_5 + z * p;
  p = EH_COS_4 + z * p;
  p = EH_COS_3 + z * p;
  p = EH_COS_2 + z * p;
  /* 1 - z/2 + z^2 p.  1 - z/2 is formed with its exact rounding error (Fast2Sum,
   * valid since z/2 <= 1) and the exact error of z = r*r, so that the final
   * addition is the only rounding of the leading terms: the result is then
   * correctly rounded except within a tiny fraction of an ulp of a midpoint. */
  double zl = eh_fma(r, r, -z);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  double t = (1.0 - w) - hz;
  return w + eh_fma(z * z, p, t - 0.5 * zl);
}

/* cos(r + rl) for a reduced argument carried as r + rl: 1 - z/2 + z^2 C(z),
 * C = C2 + C3 z + z^2 Q(z), with 1 - z/2 formed with its exact error (as in
 * eh_cos_poly), z^2 and the leading terms of C as double-double, the sum of
 * the two leading parts formed exactly, so that only the final addition rounds
 * at the level of the result; rl enters as -rl sin(r). */
EH_FN double eh_cos_poly_l(double r, double rl) {
  const double z = r * r, zl = eh_fma(r, r, -z);
  double q = EH_COS_10;
  q = EH_COS_9 + z * q;
  q = EH_COS_8 + z * q;
  q = EH_COS_7 + z * q;
  q = EH_COS_6 + z * q;
  q = EH_COS_5 + z * q;
  q = EH_COS_4 + z * q;
  /* C = ch + cl */
  const double t3 = z * EH_COS_3;
  const double t3e = eh_fma(z, EH_COS_3, -t3) + (zl * EH_COS_3 + z * EH_COS_3_LO);
  const double ch = EH_COS_2 + t3;
  const double cl = ((EH_COS_2 - ch) + t3) + (EH_COS_2_LO + t3e + (z * z) * q);
  /* z^2 = z2 + z2l, z^2 C = P + pe */
  const double z2 = z * z, z2l = eh_fma(z, z, -z2) + 2.0 * z * zl;
  const double P = z2 * ch;
  const double pe = eh_fma(z2, ch, -P) + (z2 * cl + z2l * ch);
  /* 1 - z/2 = w + t exactly (up to the -zl/2 term) */
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double t = (1.0 - w) - hz;
  /* w + P = h + hl exactly (|w| > |P|) */
  const double h = w + P;
  const double hl = (w - h) + P;
  /* rl sin(r), with sin(r) = r (1 - z/6 + z^2/120) to the precision that matters */
  const double rs = rl * (r * (1.0 + z * (EH_SIN_1 + z * EH_SIN_2)));
  return h + (hl + ((t - 0.5 * zl) + pe - rs));
}

/* Reduce x (|x| >= pi/4) to r + rl = x - k pi/2 with the rounding error of r kept
 * in rl: the first two pieces of pi/2 are subtracted exactly (Cody-Waite), the
 * rest of k pi/2 beyond r's precision is formed separately and subtracted with
 * an exact two-sum.  Returns k mod 4, or -1 outside the domain (non-finite or
 * |x| >= 2^50). */
EH_FN int eh_reduce_pio2_l(double x, double *r, double *rl) {
  if (!(eh_fabs(x) < 0x1p50)) {
    *r = *rl = eh_nan();
    return -1;
  }
  double kf = eh_fma(x, EH_TWO_OVER_PI, 0x1.8p52) - 0x1.8p52;
  double t = eh_fma(-kf, EH_PIO2_1, x);
  t = eh_fma(-kf, EH_PIO2_2, t);
  double c = eh_fma(kf, EH_PIO2_3, eh_fma(kf, EH_PIO2_4, kf * EH_PIO2_5));
  /* h + rl = t - c exactly (Knuth's TwoSum: no ordering of |t|, |c| needed) */
  double h = t - c;
  double v = h - t;
  *rl = (t - (h - v)) + (-c - v);
  *r = h;
  int64_t k = (int64_t)kf;
  return (int)(k & 3);
}

/* Reduce x to r = x - k pi/2 with k the nearest integer to x 2/pi (the rounding
 * of r is not kept); returns k mod 4, or -1 outside the domain */
EH_FN int eh_reduce_pio2(double x, double *r) {
  if (!(eh_fabs(x) < 0x1p50)) {
    *r = eh_nan();
    return -1;
  }
  double kf = eh_fma(x, EH_TWO_OVER_PI, 0x1.8p52) - 0x1.8p52;
  double t = eh_fma(-kf, EH_PIO2_1, x);
  t = eh_fma(-kf, EH_PIO2_2, t);
  t = eh_fma(-kf, EH_PIO2_3, t);
  t = eh_fma(-kf, EH_PIO2_4, t);
  t = eh_fma(-kf, EH_PIO2_5, t);
  *r = t;
  int64_t k = (int64_t)kf;
  return (int)(k & 3);
}

/* sin and cos of x for the tracking maps (CSBEND's sub-steps), whose arguments
 * lie below pi/4: the short series below 2^-4, the full series below pi/4, and
 * a plain reduction beyond (kept small: this path sits inside the maps' inner
 * loops, where registers are scarce) */
EH_FN void eh_sincos(double x, double *s, double *c) {
  const double ax = eh_fabs(x);
  if (ax < 0x1p-4) {
    *s = eh_sin_poly(x);
    *c = eh_cos_poly(x);
    return;
  }
  double r = x;
  int q = 0;
  if (!(ax < 0x1.921fb54442d18p-1)) {
    q = eh_reduce_pio2(x, &r);
    if (q < 0) {
      *s = *c = eh_nan();
      return;
    }
  }
  double ps = eh_sin_poly10(r), pc = eh_cos_poly10(r);
  switch (q) {
  case 0:
    *s = ps;
    *c = pc;
    break;
  case 1:
    *s = pc;
    *c = -ps;
    break;
  case 2:
    *s = -ps;
    *c = -pc;
    break;
  default:
    *s = -pc;
    *c = ps;
    break;
  }
}

EH_FN double eh_cos(double x) {
  const double ax = eh_fabs(x);
  if (ax < 0x1p-4)
    return eh_cos_poly(x);
  if (ax < 0x1.921fb54442d18p-1)
    return eh_cos_poly10(x);
  double r;
  switch (eh_reduce_pio2(x, &r)) {
  case 0:
    return eh_cos_poly10(r);
  case 1:
    return -eh_sin_poly10(r);
  case 2:
    return -eh_cos_poly10(r);
  case 3:
    return eh_sin_poly10(r);
  default:
    return eh_nan();
  }
}

/* sin of any argument: below pi/4 as eh_sincos; beyond, the reduction keeps
 * its rounding error and the series is evaluated in double-double, so that the
 * RF cavity's phase (about 3 rad) is correctly rounded as often as the C
 * library's */
EH_FN double eh_sin(double x) {
  const double ax = eh_fabs(x);
  if (ax < 0x1p-4)
    return eh_sin_poly(x);
  if (ax < 0x1.921fb54442d18p-1)
    return eh_sin_poly10(x);
  double r, rl;
  int q = eh_reduce_pio2_l(x, &r, &rl);
  switch (q) {
  case 0:
    return eh_sin_poly_l(r, rl);
  case 1:
    return eh_cos_poly_l(r, rl);
  case 2:
    return -eh_sin_poly_l(r, rl);
  case 3:
    return -eh_cos_poly_l(r, rl);
  default:
    return eh_nan();
  }
}

EH_FN double eh_tan(double x) {
  double s, c;
  eh_sincos(x, &s, &c);
  return s / c;
}

/* asin Taylor series; z = x*x, valid for |x| <= 1/8.  Shorter below 2^-8
 * (4 terms) and 2^-4 (7 terms): the first omitted term is below 2^-80 resp.
 * 2^-64 of the result. */
EH_FN double eh_asin_small(double x) {
  double z = x * x;
  double ax = eh_fabs(x);
  if (ax < 0x1p-8) {
    double q = EH_ASIN_4;
    q = EH_ASIN_3 + z * q;
    q = EH_ASIN_2 + z * q;
    q = EH_ASIN_1 + z * q;
    return eh_fma(x * z, q, x);
  }
  if (ax < 0x1p-4) {
    double q = EH_ASIN_7;
    q = EH_ASIN_6 + z * q;
    q = EH_ASIN_5 + z * q;
    q = EH_ASIN_4 + z * q;
    q = EH_ASIN_3 + z * q;
    q = EH_ASIN_2 + z * q;
    q = EH_ASIN_1 + z * q;
    return eh_fma(x * z, q, x);
  }
  double p = EH_ASIN_10;
  p = EH_ASIN_9 + z * p;
  p = EH_ASIN_8 + z * p;
  p = EH_ASIN_7 + z * p;
  p = EH_ASIN_6 + z * p;
  p = EH_ASIN_5 + z * p;
  p = EH_ASIN_4 + z * p;
  p = EH_ASIN_3 + z * p;
  p = EH_ASIN_2 + z * p;
  p = EH_ASIN_1 + z * p;
  return eh_fma(x * z, p, x);
}

/* asin Taylor series valid for |x| <= 1/2 (40 terms) */
EH_FN double eh_asin_mid(double x) {
  const double a[40] = {EH_ASIN_1,  EH_ASIN_2,  EH_ASIN_3,  EH_ASIN_4,  EH_ASIN_5,  EH_ASIN_6,  EH_ASIN_7,
                        EH_ASIN_8,  EH_ASIN_9,  EH_ASIN_10, EH_ASIN_11, EH_ASIN_12, EH_ASIN_13, EH_ASIN_14,
                        EH_ASIN_15, EH_ASIN_16, EH_ASIN_17, EH_ASIN_18, EH_ASIN_19, EH_ASIN_20, EH_ASIN_21,
                        EH_ASIN_22, EH_ASIN_23, EH_ASIN_24, EH_ASIN_25, EH_ASIN_26, EH_ASIN_27, EH_ASIN_28,
                        EH_ASIN_29, EH_ASIN_30, EH_ASIN_31, EH_ASIN_32, EH_ASIN_33, EH_ASIN_34, EH_ASIN_35,
                        EH_ASIN_36, EH_ASIN_37, EH_ASIN_38, EH_ASIN_39, EH_ASIN_40};
  double z = x * x;
  double p = a[39];
#pragma unroll
  for (int n = 38; n >= 0; n--)
    p = a[n] + z * p;
  return eh_fma(x * z, p, x);
}

EH_FN double eh_asin_half(double x) {
  return eh_fabs(x) <= 0.125 ? eh_asin_small(x) : eh_asin_mid(x);
}

EH_FN double eh_asin(double x) {
  double ax = eh_fabs(x);
  if (ax <= 0.5)
    return eh_asin_half(x);
  if (!(ax <= 1.0))
    return eh_nan();
  double w = eh_sqrt((1.0 - ax) * 0.5);
  double r = (EH_PIO2_HI - 2.0 * eh_asin_half(w)) + EH_PIO2_LO;
  return x < 0 ? -r : r;
}

/* exp: x = k ln2 + r with |r| <= ln2/2 (Cody-Waite, ln2 split so that k ln2_hi
 * is exact), exp(r) = (1 + r) + r^2 (1/2! + r/3! + ... + r^12/14!) with the
 * rounding error of 1 + r carried (Fast2Sum), so that only the last addition
 * rounds at the scale of the result; then the exact scaling by 2^k. */
EH_FN double eh_exp(double x) {
  if (x != x)
    return x;
  if (x > 0x1.62e42fefa39efp+9) /* ln(DBL_MAX) */
    return eh_from_bits(0x7ff0000000000000ULL);
  if (x < -0x1.74910d52d3052p+9) /* below ln of half the least subnormal */
    return 0.0;
  const double kd = __builtin_rint(x * EH_INV_LN2);
  const int k = (int)kd;
  const double r = (x - kd * EH_LN2_HI) - kd * EH_LN2_LO;
  double p = EH_EXP_14;
  p = EH_EXP_13 + r * p;
  p = EH_EXP_12 + r * p;
  p = EH_EXP_11 + r * p;
  p = EH_EXP_10 + r * p;
  p = EH_EXP_9 + r * p;
  p = EH_EXP_8 + r * p;
  p = EH_EXP_7 + r * p;
  p = EH_EXP_6 + r * p;
  p = EH_EXP_5 + r * p;
  p = EH_EXP_4 + r * p;
  p = EH_EXP_3 + r * p;
  p = EH_EXP_2 + r * p;
  const double s = (r * r) * p;
  const double hi = 1.0 + r;
  const double lo = (1.0 - hi) + r;
  const double e = hi + (lo + s);
  if (k >= -1021 && k <= 1023)
    return e * eh_from_bits((uint64_t)(k + 1023) << 52);
  if (k > 1023)
    return (e * 2.0) * eh_from_bits((uint64_t)(k - 1 + 1023) << 52);
  return (e * eh_from_bits((uint64_t)(k + 1000 + 1023) << 52)) * 0x1p-1000;
}

EH_FN double eh_log(double x) {
  if (x != x || x < 0)
    return eh_nan();
  if (x == 0)
    return -eh_from_bits(0x7ff0000000000000ULL);
  uint64_t u = eh_bits(x);
  int e = (int)((u >> 52) & 0x7ff);
  if (e == 0x7ff)
    return x; /* +inf */
  if (e == 0) { /* subnormal: scale into the normal range */
    x = x * 0x1p54;
    u = eh_bits(x);
    e = (int)((u >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  /* mantissa in [1, 2) */
  double m = eh_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 0x1.6a09e667f3bcdp+0) { /* > sqrt(2): use m/2 in [sqrt(1/2), 1) */
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  if (e == 0 && eh_fabs(f) < 0.0625) {
    /* log(1+f) = f - f^2/2 + f^3/3 - ...: f - f^2/2 is formed with its exact
     * rounding error so that the leading terms are rounded once */
    double h = f * f;
    double hl = eh_fma(f, f, -h);
    double hh = 0.5 * h;
    double w = f - hh;
    double t = (f - w) - hh;
    double q = EH_LOG1P_16;
    q = EH_LOG1P_15 + f * q;
    q = EH_LOG1P_14 + f * q;
    q = EH_LOG1P_13 + f * q;
    q = EH_LOG1P_12 + f * q;
    q = EH_LOG1P_11 + f * q;
    q = EH_LOG1P_10 + f * q;
    q = EH_LOG1P_9 + f * q;
    q = EH_LOG1P_8 + f * q;
    q = EH_LOG1P_7 + f * q;
    q = EH_LOG1P_6 + f * q;
    q = EH_LOG1P_5 + f * q;
    q = EH_LOG1P_4 + f * q;
    q = EH_LOG1P_3 + f * q;
    return w + eh_fma(h * f, q, t - 0.5 * hl);
  }
  double s = f / (2.0 + f);
  double z = s * s;
  double p = EH_LOG_12;
  p = EH_LOG_11 + z * p;
  p = EH_LOG_10 + z * p;
  p = EH_LOG_9 + z * p;
  p = EH_LOG_8 + z * p;
  p = EH_LOG_7 + z * p;
  p = EH_LOG_6 + z * p;
  p = EH_LOG_5 + z * p;
  p = EH_LOG_4 + z * p;
  p = EH_LOG_3 + z * p;
  p = EH_LOG_2 + z * p;
  p = EH_LOG_1 + z * p;
  double lm = eh_fma(s * z, p, 2.0 * s);
  double de = (double)e;
  return de * EH_LN2_HI + (lm + de * EH_LN2_LO);
}

#endif
