/* Elementary functions for the spray (exp, log, cos, pow) and the bullet (sin, cos, asin),
 * shared by the kernels and the CPU oracle so that both give the same bits. WGSL leaves the
 * accuracy of these functions to the adapter (several ulp), so any implementation within a few
 * ulp is as faithful to the reference as the adapter's own.
 *
 * Each function is evaluated in double and rounded once to float:
 *  - exp: x = k ln2 + r with |r| <= ln2 / 2 (ln2 split in a high part exact in the product and
 *    a low part), exp(r) by its Taylor series to r^13 / 13!, scaled by 2^k;
 *  - log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1) / (m + 1),
 *    |s| <= 0.172, by the atanh series to s^21 / 21, plus e ln2;
 *  - cos and sin: x = q pi/2 + t with |t| <= pi/4 (pi/2 split high and low), then +-cos t or
 *    +-sin t by their Taylor series to t^18 / 18! and t^17 / 17!;
 *  - asin: for |x| <= 1/2 the series sum of (2n)! / (4^n n!^2 (2n+1)) x^(2n+1) to n = 29 (the
 *    terms fall by x^2 <= 1/4 each); above, asin x = pi/2 - 2 asin(sqrt((1 - x) / 2)), whose
 *    argument is at most 1/2 (1 - x is exact there);
 *  - pow(x, y) = exp(y log x) for x > 0, in double.
 * The double results are within ~1e-15 relative of the true values, so the float results are
 * the correctly rounded ones except when the true value lies within that of a rounding
 * boundary. Every expression is written out (no contraction: both builds use
 * -ffp-contract=off). */
#ifndef NATRAY_MATH_H
#define NATRAY_MATH_H

#include <math.h>

#ifdef __HIPCC__
#define NM_FN __host__ __device__ static inline
#else
#define NM_FN static inline
#endif

#define NM_LN2_HI 6.93147180369123816490e-01 /* 0x3fe62e42fee00000: k * NM_LN2_HI is exact */
#define NM_LN2_LO 1.90821492927058770002e-10
#define NM_LN2 6.93147180559945286227e-01
#define NM_INV_LN2 1.44269504088896338700e+00
#define NM_PIO2_HI 1.57079632673412561417e+00 /* 0x3ff921fb54400000: q * NM_PIO2_HI is exact */
#define NM_PIO2_LO 6.07710050650619224932e-11
#define NM_2_PI 6.36619772367581382433e-01
#define NM_PIO2 1.57079632679489655800e+00

NM_FN double nm_exp_d(double x) {
  if (x != x) return x;
  if (x > 709.0) return HUGE_VAL;
  if (x < -745.0) return 0.0;
  const double k = rint(x * NM_INV_LN2);
  const double r = (x - k * NM_LN2_HI) - k * NM_LN2_LO;
  double p = 1.0 / 6227020800.0; /* 1 / 13! */
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  return ldexp(p, (int)k);
}

NM_FN double nm_log_d(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e = 0;
  double m = frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
  if (m < 7.07106781186547524401e-01) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = 1.0 / 19.0 + s2 * p;
  p = 1.0 / 17.0 + s2 * p;
  p = 1.0 / 15.0 + s2 * p;
  p = 1.0 / 13.0 + s2 * p;
  p = 1.0 / 11.0 + s2 * p;
  p = 1.0 / 9.0 + s2 * p;
  p = 1.0 / 7.0 + s2 * p;
  p = 1.0 / 5.0 + s2 * p;
  p = 1.0 / 3.0 + s2 * p;
  p = 1.0 + s2 * p;
  return (double)e * NM_LN2 + 2.0 * s * p;
}

/* cos x for shift 0, sin x = cos(x - pi/2) for shift 1. */
NM_FN double nm_trig_d(double x, int shift) {
  if (x != x || x == HUGE_VAL || x == -HUGE_VAL) return NAN;
  const double q = rint(x * NM_2_PI);
  const double t = (x - q * NM_PIO2_HI) - q * NM_PIO2_LO;
  const double t2 = t * t;
  /* cos t = sum of (-1)^n t^2n / (2n)!, n = 0..9, and sin t = t sum of (-1)^n t^2n / (2n+1)!,
   * n = 0..8, both by Horner in t^2. */
  double c = -1.0 / 6402373705728000.0; /* -1 / 18! */
  c = 1.0 / 20922789888000.0 + t2 * c;
  c = -1.0 / 87178291200.0 + t2 * c;
  c = 1.0 / 479001600.0 + t2 * c;
  c = -1.0 / 3628800.0 + t2 * c;
  c = 1.0 / 40320.0 + t2 * c;
  c = -1.0 / 720.0 + t2 * c;
  c = 1.0 / 24.0 + t2 * c;
  c = -0.5 + t2 * c;
  c = 1.0 + t2 * c;
  double s = 1.0 / 355687428096000.0; /* 1 / 17! */
  s = -1.0 / 1307674368000.0 + t2 * s;
  s = 1.0 / 6227020800.0 + t2 * s;
  s = -1.0 / 39916800.0 + t2 * s;
  s = 1.0 / 362880.0 + t2 * s;
  s = -1.0 / 5040.0 + t2 * s;
  s = 1.0 / 120.0 + t2 * s;
  s = -1.0 / 6.0 + t2 * s;
  s = t + t * (t2 * s);
  const long long iq = (long long)q;
  const int quad = (int)((((iq - shift) % 4) + 4) % 4);
  return quad == 0 ? c : (quad == 1 ? -s : (quad == 2 ? -c : s));
}
NM_FN double nm_cos_d(double x) { return nm_trig_d(x, 0); }
NM_FN double nm_sin_d(double x) { return nm_trig_d(x, 1); }

/* The asin series for 0 <= x <= 1/2. */
NM_FN double nm_asin_series(double x) {
  const double x2 = x * x;
  double a = x, sum = x;
  for (int n = 0; n < 29; ++n) {
    a = a * x2 * (double)(2 * n + 1) / (double)(2 * n + 2);
    sum = sum + a / (double)(2 * n + 3);
  }
  return sum;
}
NM_FN double nm_asin_d(double x) {
  if (x != x || x > 1.0 || x < -1.0) return NAN;
  const double ax = x < 0.0 ? -x : x;
  const double r = ax <= 0.5 ? nm_asin_series(ax) : NM_PIO2 - 2.0 * nm_asin_series(sqrt((1.0 - ax) * 0.5));
  return x < 0.0 ? -r : r;
}

NM_FN float nm_expf(float x) { return (float)nm_exp_d((double)x); }
NM_FN float nm_logf(float x) { return (float)nm_log_d((double)x); }
NM_FN float nm_cosf(float x) { return (float)nm_cos_d((double)x); }
NM_FN float nm_sinf(float x) { return (float)nm_sin_d((double)x); }
NM_FN float nm_asinf(float x) { return (float)nm_asin_d((double)x); }
/* pow(x, y) for x > 0 (the spray's only use); 0 for x == 0 and y > 0. */
NM_FN float nm_powf(float x, float y) {
  if (x == 0.0f) return y > 0.0f ? 0.0f : (float)HUGE_VAL;
  return (float)nm_exp_d((double)y * nm_log_d((double)x));
}

#endif
