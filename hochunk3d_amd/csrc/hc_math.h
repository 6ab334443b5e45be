/*
 * hc_math.h -- deterministic FP64 elementary functions for the plain FP64
 * transport path.
 *
 * The same source is compiled for the GPU and for the host: every function
 * is a fixed sequence of IEEE additions, multiplications, divisions and
 * square roots (all correctly rounded on both sides; build with
 * -ffp-contract=off), so a packet's arithmetic is bit-identical wherever it
 * runs.  This is what makes the deterministic-trace comparison exact.
 *
 * Methods (each from the textbook mathematics, coefficients are exact
 * rationals rounded once to double):
 *   exp   x = k ln2 + r, |r| <= ln2/2 (ln2 split in a 32-bit head and a tail),
 *         Taylor series of e^r to r^13, scaled by 2^k.
 *   log   x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f),
 *         log(1+f) = f - s (f - R(s^2)) with R the series of 2 atanh(s) - 2s.
 *   sin, cos  x = q pi/2 + r (pi/2 in three parts), Taylor series of sin r
 *         and cos r on |r| <= pi/4 (valid for |x| < 2^20).
 *   acos  asin series on |y| <= 1/2 with acos(x) = pi/2 - asin(x) there and
 *         acos(x) = 2 asin(sqrt((1-x)/2)) (x > 1/2), pi - 2 asin(sqrt((1+x)/2))
 *         (x < -1/2).
 *   atan2 angle from the x axis as asin(|y|/r) (|y|/r <= 1/2) or acos(|x|/r),
 *         r = sqrt(x^2 + y^2), placed in its quadrant.
 *   pow   x^0.25 = sqrt(sqrt(x)), x^2 = x x, else exp(y log x) (x > 0).
 *   log10 log(x) times 1/ln 10.
 *   exp10 10^y = e^p (1 + e) with p = y ln 10 rounded, e its rounding error
 *         (a fused multiply-add, correctly rounded on every target) plus
 *         y times the tail of ln 10.
 * Accuracy is within a few ulp of the correctly rounded result over the
 * arguments the transport path uses; tests/test_hc_math.py measures it.
 *
 * The includer may define HCM_FN (e.g. with __device__ qualifiers); by
 * default the functions are static inline C.
 */
#ifndef HC_MATH_H
#define HC_MATH_H

#include <stdint.h>

#ifndef HCM_FN
#define HCM_FN static inline
#endif

typedef union {
  double d;
  uint64_t u;
} hcm_bits;

HCM_FN double hcm_pow2i(int k) { /* 2^k for -1022 <= k <= 1023 */
  hcm_bits b;
  b.u = (uint64_t)(k + 1023) << 52;
  return b.d;
}

HCM_FN double hcm_exp(double x) {
  if (x != x) return x;
  if (x > 709.78) return 1.0 / 0.0;
  if (x < -745.2) return 0.0;
  const double inv_ln2 = 1.4426950408889634;
  const double ln2_hi = 0.6931471806019545; /* 32 significant bits */
  const double ln2_lo = -4.2009150726810846e-11;
  double kd = x * inv_ln2;
  int k = (int)(kd + (kd >= 0.0 ? 0.5 : -0.5));
  double r = (x - (double)k * ln2_hi) - (double)k * ln2_lo;
  double p = 1.6059043836821613e-10;
  p = 2.08767569878681e-09 + r * p;
  p = 2.505210838544172e-08 + r * p;
  p = 2.755731922398589e-07 + r * p;
  p = 2.7557319223985893e-06 + r * p;
  p = 2.48015873015873e-05 + r * p;
  p = 0.0001984126984126984 + r * p;
  p = 0.001388888888888889 + r * p;
  p = 0.008333333333333333 + r * p;
  p = 0.041666666666666664 + r * p;
  p = 0.16666666666666666 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  /* scale by 2^k in two steps so each factor stays normal */
  int k1 = k / 2;
  return (p * hcm_pow2i(k1)) * hcm_pow2i(k - k1);
}

HCM_FN double hcm_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return 0.0 / 0.0;
  if (x == 0.0) return -1.0 / 0.0;
  if (x == 1.0 / 0.0) return x;
  hcm_bits b;
  int e = 0;
  b.d = x;
  if ((b.u >> 52) == 0) { /* subnormal */
    b.d = x * 18014398509481984.0; /* 2^54 */
    e = -54;
  }
  e += (int)((b.u >> 52) & 0x7ff) - 1023;
  b.u = (b.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull; /* m in [1,2) */
  double m = b.d;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 0.10526315789473684;
  R = 0.11764705882352941 + z * R;
  R = 0.13333333333333333 + z * R;
  R = 0.15384615384615385 + z * R;
  R = 0.18181818181818182 + z * R;
  R = 0.2222222222222222 + z * R;
  R = 0.2857142857142857 + z * R;
  R = 0.4 + z * R;
  R = 0.6666666666666666 + z * R;
  R = z * R;
  const double ln2_hi = 0.6931471806019545;
  const double ln2_lo = -4.2009150726810846e-11;
  const double ed = (double)e;
  return ed * ln2_hi + ((ed * ln2_lo - s * (f - R)) + f);
}

/* x = q pi/2 + r; returns r, q in *q */
HCM_FN double hcm_reduce(double x, int *q) {
  const double two_over_pi = 0.6366197723675814;
  const double p1 = 1.5707963267341256; /* 33 significant bits */
  const double p2 = 6.077100506303966e-11;
  const double p3 = 2.0222662487959506e-21;
  double kd = x * two_over_pi;
  int k = (int)(kd + (kd >= 0.0 ? 0.5 : -0.5));
  *q = k;
  return ((x - (double)k * p1) - (double)k * p2) - (double)k * p3;
}

HCM_FN double hcm_sin_poly(double r) {
  const double z = r * r;
  double p = 2.8114572543455206e-15;
  p = -7.647163731819816e-13 + z * p;
  p = 1.6059043836821613e-10 + z * p;
  p = -2.505210838544172e-08 + z * p;
  p = 2.7557319223985893e-06 + z * p;
  p = -0.0001984126984126984 + z * p;
  p = 0.008333333333333333 + z * p;
  p = -0.16666666666666666 + z * p;
  return r + r * (z * p);
}

HCM_FN double hcm_cos_poly(double r) {
  const double z = r * r;
  double p = -1.5619206968586225e-16;
  p = 4.779477332387385e-14 + z * p;
  p = -1.1470745597729725e-11 + z * p;
  p = 2.08767569878681e-09 + z * p;
  p = -2.755731922398589e-07 + z * p;
  p = 2.48015873015873e-05 + z * p;
  p = -0.001388888888888889 + z * p;
  p = 0.041666666666666664 + z * p;
  return 1.0 - (0.5 * z - (z * z) * p);
}

HCM_FN double hcm_sin(double x) {
  int q;
  const double r = hcm_reduce(x, &q);
  switch (q & 3) {
    case 0: return hcm_sin_poly(r);
    case 1: return hcm_cos_poly(r);
    case 2: return -hcm_sin_poly(r);
    default: return -hcm_cos_poly(r);
  }
}

HCM_FN double hcm_cos(double x) {
  int q;
  const double r = hcm_reduce(x, &q);
  switch (q & 3) {
    case 0: return hcm_cos_poly(r);
    case 1: return -hcm_sin_poly(r);
    case 2: return -hcm_cos_poly(r);
    default: return hcm_sin_poly(r);
  }
}

/* asin(y) for |y| <= 1/2 */
HCM_FN double hcm_asin_small(double y) {
  const double z = y * y;
  double p = 0.0022014739737101384;
  p = 0.002338091892111975 + z * p;
  p = 0.0024894486782468836 + z * p;
  p = 0.00265787063820729 + z * p;
  p = 0.002846178401108942 + z * p;
  p = 0.0030578216492580306 + z * p;
  p = 0.003297059503473485 + z * p;
  p = 0.0035692053938259347 + z * p;
  p = 0.003880964558837669 + z * p;
  p = 0.004240907093679363 + z * p;
  p = 0.004660143486915096 + z * p;
  p = 0.005153309682319905 + z * p;
  p = 0.005740037670841924 + z * p;
  p = 0.006447210311889649 + z * p;
  p = 0.0073125258735988454 + z * p;
  p = 0.008390335809616815 + z * p;
  p = 0.009761609529194078 + z * p;
  p = 0.011551800896139705 + z * p;
  p = 0.01396484375 + z * p;
  p = 0.017352764423076924 + z * p;
  p = 0.022372159090909092 + z * p;
  p = 0.030381944444444444 + z * p;
  p = 0.044642857142857144 + z * p;
  p = 0.075 + z * p;
  p = 0.16666666666666666 + z * p;
  return y + y * (z * p);
}

HCM_FN double hcm_acos(double x) {
  const double pio2_hi = 1.5707963267948966, pio2_lo = 6.123233995736766e-17;
  const double pi_hi = 3.141592653589793, pi_lo = 1.2246467991473532e-16;
  if (x != x || x > 1.0 || x < -1.0) return 0.0 / 0.0;
  if (x <= 0.5 && x >= -0.5) return pio2_hi - (hcm_asin_small(x) - pio2_lo);
  if (x > 0.5) return 2.0 * hcm_asin_small(__builtin_sqrt((1.0 - x) * 0.5));
  return (pi_hi - 2.0 * hcm_asin_small(__builtin_sqrt((1.0 + x) * 0.5))) + pi_lo;
}

HCM_FN double hcm_atan2(double y, double x) {
  const double pi_hi = 3.141592653589793, pi_lo = 1.2246467991473532e-16;
  const double r = __builtin_sqrt(x * x + y * y);
  if (r == 0.0) return 0.0;
  const double ay = y < 0.0 ? -y : y, ax = x < 0.0 ? -x : x;
  /* angle from the x axis: asin of the small ratio keeps relative accuracy */
  double t = ay / r, u = ax / r, a0;
  if (t > 1.0) t = 1.0;
  if (u > 1.0) u = 1.0;
  a0 = (t <= 0.5) ? hcm_asin_small(t) : hcm_acos(u);
  const double a = (x >= 0.0) ? a0 : (pi_hi - a0) + pi_lo;
  return y < 0.0 ? -a : a;
}

HCM_FN double hcm_pow(double x, double y) {
  if (y == 2.0) return x * x;
  if (y == 0.25) return __builtin_sqrt(__builtin_sqrt(x));
  if (y == 0.0) return 1.0;
  if (x == 0.0) return y > 0.0 ? 0.0 : 1.0 / 0.0;
  if (x == 1.0) return 1.0;
  return hcm_exp(y * hcm_log(x));
}

HCM_FN double hcm_log10(double x) { return hcm_log(x) * 0.43429448190325182765; }

HCM_FN double hcm_exp10(double y) {
  const double l_hi = 2.302585092994046;       /* ln 10 rounded to double */
  const double l_lo = -2.1707562233822494e-16; /* ln 10 - l_hi */
  const double p = y * l_hi;
  /* the rounding error of p (a correctly rounded fused multiply-add on
   * every target) plus the tail of ln 10: 10^y = e^p (1 + e) */
  const double e = __builtin_fma(y, l_hi, -p) + y * l_lo;
  const double x = hcm_exp(p);
  return x + x * e;
}

#endif
