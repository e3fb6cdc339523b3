/* Deterministic elementary functions for f32 arguments.
 *
 * Each function widens its f32 arguments to f64, evaluates with IEEE
 * +, -, *, / and sqrt only (plus exact bit operations and floor), in a fixed
 * order, and rounds the f64 result to f32 once. Built with -ffp-contract=off,
 * host (x86-64 SSE2) and device (gfx950) builds therefore return bit-identical
 * results, which lets the path tracer and its CPU checker draw the same paths
 * sample for sample. Accuracy: the f64 intermediate is good to a few units of
 * 2^-52 relative, so the f32 result is the correctly rounded value except in
 * rare near-halfway cases (at most 1 ulp away).
 *
 * Methods (textbook):
 *   sin/cos  Cody-Waite reduction by pi/2 with a three-part constant, Taylor
 *            series on |r| <= pi/4 (exact to ~1e-16 there). Arguments up to
 *            |x| ~ 1e6 reduce exactly; larger ones lose accuracy (not used).
 *   exp      reduction by ln 2 (two-part constant), Taylor series to r^13.
 *   log      x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1))
 *            as an odd series in s = (m-1)/(m+1).
 *   atan     table of atan(j/8), j = 0..8, and the addition formula
 *            atan(a) = atan(c) + atan((a - c) / (1 + a c)), |reduced| <= 1/16.
 *   pow      exp(y log x) for x > 0 (C99 special cases otherwise).
 *   log2     log x times the f64 rounding of 1 / ln 2.
 *   acos     atan2(sqrt((1 - x)(1 + x)), x).
 * The split constants below are the leading 25 significant bits of the
 * constant, then the next 25, then the f64 rounding of the remainder.
 */
#ifndef MTX_FMATH_H
#define MTX_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define MTX_FM_HD __host__ __device__ inline
/* The public functions stay out of line: inlined at every call site of
 * the shading code their f64 temporaries multiply the kernels' register
 * pressure (and spills); as calls they cost a few instructions each. */
#define MTX_FM_CALL __host__ __device__ inline __attribute__((noinline))
#else
#define MTX_FM_HD static inline
#define MTX_FM_CALL static inline
#endif

namespace mtxfm {

MTX_FM_HD uint64_t dbits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
MTX_FM_HD double dfrom(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
MTX_FM_HD uint32_t fbits(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
MTX_FM_HD float ffrom(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }
MTX_FM_HD bool is_nan(double x) { return x != x; }
MTX_FM_HD bool is_inf(double x) { return (dbits(x) & 0x7fffffffffffffffull) == 0x7ff0000000000000ull; }

constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double kPiO2 = 0x1.921fb54442d18p+0;
constexpr double kPiO2Hi = 0x1.921fb50000000p+0;
constexpr double kPiO2Mid = 0x1.110b460000000p-26;
constexpr double kPiO2Lo = 0x1.1a62633145c07p-54;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kLn2Hi = 0x1.62e42f0000000p-1;
constexpr double kLn2Lo = 0x1.df473de6af279p-26;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;

/* 2^k for integer k in [-1074, 1023] (exact). */
MTX_FM_HD double exp2i(int k) {
  if (k >= -1022) return dfrom((uint64_t)(k + 1023) << 52);
  return dfrom((uint64_t)(k + 1023 + 54) << 52) * 0x1p-54;
}

/* sin and cos of r, |r| <= pi/4 (Taylor, Horner in r^2). */
MTX_FM_HD double sin_poly(double r) {
  double z = r * r;
  double p = 1.0 / 355687428096000.0;
  p = p * z - 1.0 / 1307674368000.0;
  p = p * z + 1.0 / 6227020800.0;
  p = p * z - 1.0 / 39916800.0;
  p = p * z + 1.0 / 362880.0;
  p = p * z - 1.0 / 5040.0;
  p = p * z + 1.0 / 120.0;
  p = p * z - 1.0 / 6.0;
  return r + (r * z) * p;
}
MTX_FM_HD double cos_poly(double r) {
  double z = r * r;
  double p = 1.0 / 20922789888000.0;
  p = p * z - 1.0 / 87178291200.0;
  p = p * z + 1.0 / 479001600.0;
  p = p * z - 1.0 / 3628800.0;
  p = p * z + 1.0 / 40320.0;
  p = p * z - 1.0 / 720.0;
  p = p * z + 1.0 / 24.0;
  p = p * z - 0.5;
  return 1.0 + z * p;
}
/* Reduces x to r in [-pi/4, pi/4] and the quadrant k mod 4. */
MTX_FM_HD double reduce_pio2(double x, int &quadrant) {
  double k = __builtin_floor(x * kTwoOverPi + 0.5);
  double r = x - k * kPiO2Hi;
  r = r - k * kPiO2Mid;
  r = r - k * kPiO2Lo;
  int64_t ki = (int64_t)k;
  quadrant = (int)(ki & 3);
  return r;
}
MTX_FM_CALL float sinf(float xf) {
  double x = xf;
  if (is_nan(x) || is_inf(x)) return ffrom(0x7fc00000u);
  if (x == 0.0) return xf;  /* keeps the sign of zero */
  int q;
  double r = reduce_pio2(x, q);
  double v = (q & 1) ? cos_poly(r) : sin_poly(r);
  return (float)((q & 2) ? -v : v);
}
MTX_FM_CALL float cosf(float xf) {
  double x = xf;
  if (is_nan(x) || is_inf(x)) return ffrom(0x7fc00000u);
  int q;
  double r = reduce_pio2(x, q);
  double v = (q & 1) ? sin_poly(r) : cos_poly(r);
  return (float)(((q + 1) & 2) ? -v : v);
}

/* sin and cos of one argument from one reduction: the same two values as
 * sinf(x) and cosf(x), for half the work of two calls. */
struct SinCos {
  float s, c;
};
MTX_FM_CALL SinCos sincosf(float xf) {
  double x = xf;
  if (is_nan(x) || is_inf(x)) return SinCos{ffrom(0x7fc00000u), ffrom(0x7fc00000u)};
  int q;
  double r = reduce_pio2(x, q);
  double s = sin_poly(r), c = cos_poly(r);
  double sv = (q & 1) ? c : s, cv = (q & 1) ? s : c;
  SinCos out;
  out.s = x == 0.0 ? xf : (float)((q & 2) ? -sv : sv);
  out.c = (float)(((q + 1) & 2) ? -cv : cv);
  return out;
}

/* e^t for finite t. */
MTX_FM_HD double exp_d(double t) {
  if (t > 709.0) return dfrom(0x7ff0000000000000ull);
  if (t < -745.0) return 0.0;
  double k = __builtin_floor(t * kInvLn2 + 0.5);
  double r = t - k * kLn2Hi;
  r = r - k * kLn2Lo;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  int ki = (int)k;
  if (ki > 1023) return (p * exp2i(1023)) * exp2i(ki - 1023);
  if (ki < -1022) return (p * exp2i(ki + 600)) * 0x1p-600;
  return p * exp2i(ki);
}

/* Natural log for finite x > 0. */
MTX_FM_HD double log_d(double x) {
  int e = 0;
  if (x < 0x1p-1022) { x = x * 0x1p54; e = -54; }
  uint64_t u = dbits(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  double m = dfrom((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  double p = 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  double lm = (2.0 * s) + (2.0 * s) * (z * p);
  double de = (double)e;
  return (de * kLn2Hi + lm) + de * kLn2Lo;
}

MTX_FM_CALL float expf(float xf) {
  double x = xf;
  if (is_nan(x)) return ffrom(0x7fc00000u);
  if (is_inf(x)) return x > 0.0 ? ffrom(0x7f800000u) : 0.0f;
  return (float)exp_d(x);
}

MTX_FM_CALL float powf(float xf, float yf) {
  double x = xf, y = yf;
  if (y == 0.0 || x == 1.0) return 1.0f;
  if (is_nan(x) || is_nan(y)) return ffrom(0x7fc00000u);
  bool yInt = false, yOdd = false;
  if (!is_inf(y) && __builtin_floor(y) == y) {
    yInt = true;
    double h = y * 0.5;
    yOdd = __builtin_floor(h) != h;
  }
  double sign = 1.0;
  if (x < 0.0 || (x == 0.0 && (fbits(xf) >> 31))) {
    if (x != 0.0 && !is_inf(x) && !yInt) return ffrom(0x7fc00000u);
    if (yOdd) sign = -1.0;
    x = -x;
  }
  if (x == 0.0) return (float)(sign * (y > 0.0 ? 0.0 : dfrom(0x7ff0000000000000ull)));
  if (is_inf(x)) return (float)(sign * (y > 0.0 ? dfrom(0x7ff0000000000000ull) : 0.0));
  if (is_inf(y)) {
    if (x == 1.0) return 1.0f;
    bool big = x > 1.0;
    return (big == (y > 0.0)) ? ffrom(0x7f800000u) : 0.0f;
  }
  return (float)(sign * exp_d(y * log_d(x)));
}

/* log2 = log(x) / ln 2 with the C99 special cases (log2(0) = -inf,
 * log2(x < 0) = NaN, log2(+inf) = +inf). */
MTX_FM_CALL float log2f(float xf) {
  double x = xf;
  if (is_nan(x) || x < 0.0) return ffrom(0x7fc00000u);
  if (x == 0.0) return ffrom(0xff800000u);
  if (is_inf(x)) return ffrom(0x7f800000u);
  return (float)(log_d(x) * kInvLn2);
}

/* atan for a >= 0 (finite or +inf). */
MTX_FM_HD double atan_pos(double a) {
  bool inv = a > 1.0;
  if (inv) a = 1.0 / a;
  int j = (int)(a * 8.0 + 0.5);
  double c = (double)j * 0.125;
  double t = (a - c) / (1.0 + a * c);
  double z = t * t;
  double p = -1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z - 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z - 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z - 1.0 / 3.0;
  double at = t + (t * z) * p;
  /* atan(j / 8), j = 0..8 (f64 roundings of the exact values). */
  double base = j == 0 ? 0.0
              : j == 1 ? 0x1.fd5ba9aac2f6ep-4
              : j == 2 ? 0x1.f5b75f92c80ddp-3
              : j == 3 ? 0x1.6f61941e4def1p-2
              : j == 4 ? 0x1.dac670561bb4fp-2
              : j == 5 ? 0x1.1e00babdefeb4p-1
              : j == 6 ? 0x1.4978fa3269ee1p-1
              : j == 7 ? 0x1.700a7c5784634p-1
                       : 0x1.921fb54442d18p-1;
  double r = base + at;
  return inv ? kPiO2 - r : r;
}

/* atan2 on f64 arguments with the C99 (Annex F) special cases. */
MTX_FM_HD double atan2_d(double y, double x, bool yNeg, bool xNeg) {
  if (is_nan(x) || is_nan(y)) return dfrom(0x7ff8000000000000ull);
  double r;
  if (y == 0.0) {
    r = xNeg ? kPi : 0.0;
  } else if (x == 0.0) {
    r = kPiO2;
  } else if (is_inf(y) && is_inf(x)) {
    r = xNeg ? 3.0 * kPi / 4.0 : kPi / 4.0;
  } else if (is_inf(x)) {
    r = xNeg ? kPi : 0.0;
  } else if (is_inf(y)) {
    r = kPiO2;
  } else {
    double ay = y < 0.0 ? -y : y, ax = x < 0.0 ? -x : x;
    r = atan_pos(ay / ax);
    if (xNeg) r = kPi - r;
  }
  return yNeg ? -r : r;
}
MTX_FM_CALL float atan2f(float y, float x) {
  return (float)atan2_d((double)y, (double)x, (fbits(y) >> 31) != 0, (fbits(x) >> 31) != 0);
}
MTX_FM_CALL float acosf(float xf) {
  double x = xf;
  if (is_nan(x) || x > 1.0 || x < -1.0) return ffrom(0x7fc00000u);
  double s = __builtin_sqrt((1.0 - x) * (1.0 + x));
  return (float)atan2_d(s, x, false, x < 0.0);
}

}  // namespace mtxfm

#endif /* MTX_FMATH_H */
