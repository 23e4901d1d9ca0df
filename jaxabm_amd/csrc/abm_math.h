/* abm_math.h -- exp, log, log1p and pow written with IEEE basic operations
 * only (+, -, *, /, exact conversions and bit moves), shared by the device
 * kernels, the kernels generated at run time (hiprtc) and any host code that
 * must compute the same bits.
 *
 * Each operation is rounded to nearest by IEEE 754, so any compiler that
 * neither contracts a multiply and an add into a fused multiply-add nor keeps
 * excess precision (hipcc and gcc with -ffp-contract=off on x86-64 SSE)
 * computes the same result for every input.  float32 exp and log are
 * accurate to one ulp; the float64 ones to about one float64 ulp, so log1p
 * and pow, computed through them, are nearly correctly rounded in float32.
 *
 * Methods:
 *  * exp: Cody & Waite range reduction (W. J. Cody, W. Waite, "Software
 *    Manual for the Elementary Functions", 1980): x = n ln2 + r with ln2
 *    split into a high part whose products with n are exact and a low part,
 *    |r| <= ln2/2; exp(r) by its Taylor series (degree 7 in float32, 13 in
 *    float64, truncation below a quarter ulp); scaling by 2^n in at most two
 *    exact steps so that a subnormal result is rounded once.
 *  * log: x = 2^e m with m in [sqrt(2)/2, sqrt(2)), f = m - 1 (exact),
 *    s = f / (2 + f), log(1 + f) = 2 atanh(s) = 2s + s R(s^2) with
 *    R(z) = sum_k 2 z^k / (2k + 1), evaluated as
 *    f - (f^2/2 - s (f^2/2 + R)) (the arrangement of Sun's fdlibm e_log.c,
 *    which keeps the large term f exact), plus e ln2 split as for exp.
 *  * log1p: D. Goldberg, "What every computer scientist should know about
 *    floating-point arithmetic" (1991), Theorem 4: u = 1 + x,
 *    log1p(x) = x when u == 1, else log(u) x / (u - 1); in float64, rounded
 *    to float32 once.
 *  * pow(x, y) for x >= 0: exp(y log x) in float64, rounded once to float32.
 */
#ifndef ABM_MATH_H
#define ABM_MATH_H

#if defined(__HIPCC_RTC__) || defined(__HIP__)
#define ABM_MATH_FN __host__ __device__ static inline
#else
#include <stdint.h>
#define ABM_MATH_FN static inline
#endif

ABM_MATH_FN uint32_t abm_f32_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, sizeof u);
  return u;
}
ABM_MATH_FN float abm_f32_from_bits(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, sizeof f);
  return f;
}
ABM_MATH_FN uint64_t abm_f64_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, sizeof u);
  return u;
}
ABM_MATH_FN double abm_f64_from_bits(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, sizeof d);
  return d;
}

/* ---- float32 ---- */

/* 2^k for -126 <= k <= 127. */
ABM_MATH_FN float abm_pow2f(int k) { return abm_f32_from_bits((uint32_t)(k + 127) << 23); }

/* p 2^n for p in [0.5, 2] and -152 <= n <= 129, rounded once. */
ABM_MATH_FN float abm_scale2f(float p, int n) {
  if (n < -126) {
    const float t = p * abm_pow2f(n + 64); /* normal: exact */
    return t * 0x1p-64f;                   /* the one rounding */
  }
  if (n > 127) {
    const float t = p * 0x1p127f;
    return t * abm_pow2f(n - 127);
  }
  return p * abm_pow2f(n);
}

ABM_MATH_FN float abm_expf(float x) {
  if (!(x == x)) return x + x;
  if (x > 0x1.62e42ep6f) return abm_f32_from_bits(0x7f800000u); /* exp overflows */
  if (x < -104.0f) return 0.0f;                                   /* below 2^-150 */
  const float n = __builtin_rintf(x * 0x1.715476p0f);             /* x log2(e) */
  /* ln2 = 0x1.62e4p-1 (16 significant bits, n * it exact) + 0x1.7f7d1cp-20 */
  const float r = (x - n * 0x1.62e4p-1f) - n * 0x1.7f7d1cp-20f;
  float q = 0x1.a01a02p-13f; /* 1/7! */
  q = 0x1.6c16c2p-10f + q * r; /* 1/6! */
  q = 0x1.111112p-7f + q * r;  /* 1/5! */
  q = 0x1.555556p-5f + q * r;  /* 1/4! */
  q = 0x1.555556p-3f + q * r;  /* 1/3! */
  q = 0.5f + q * r;
  const float p = 1.0f + (r + (r * r) * q);
  return abm_scale2f(p, (int)n);
}

ABM_MATH_FN float abm_logf(float x) {
  if (!(x == x)) return x + x;
  if (x < 0.0f) return abm_f32_from_bits(0x7fc00000u);
  if (x == 0.0f) return abm_f32_from_bits(0xff800000u);
  uint32_t b = abm_f32_bits(x);
  if (b == 0x7f800000u) return x;
  int e = 0;
  if (b < 0x00800000u) { /* subnormal: scale into the normal range (exact) */
    x = x * 0x1p23f;
    b = abm_f32_bits(x);
    e = -23;
  }
  e += (int)(b >> 23) - 127;
  float m = abm_f32_from_bits((b & 0x007fffffu) | 0x3f800000u);
  if (m > 0x1.6a09e6p0f) { /* sqrt(2) */
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float R = 0x1.745d18p-3f;  /* 2/11 */
  R = 0x1.c71c72p-3f + R * z; /* 2/9 */
  R = 0x1.24924ap-2f + R * z; /* 2/7 */
  R = 0x1.99999ap-2f + R * z; /* 2/5 */
  R = 0x1.555556p-1f + R * z; /* 2/3 */
  R = R * z;
  const float hfsq = 0.5f * (f * f);
  const float ef = (float)e;
  return ef * 0x1.62e4p-1f - ((hfsq - (s * (hfsq + R) + ef * 0x1.7f7d1cp-20f)) - f);
}

/* ---- float64 ---- */

/* 2^k for -1022 <= k <= 1023. */
ABM_MATH_FN double abm_pow2(int k) { return abm_f64_from_bits((uint64_t)(k + 1023) << 52); }

/* p 2^n for p in [0.5, 2] and -1077 <= n <= 1025, rounded once. */
ABM_MATH_FN double abm_scale2(double p, int n) {
  if (n < -1022) {
    const double t = p * abm_pow2(n + 200);
    return t * 0x1p-200;
  }
  if (n > 1023) {
    const double t = p * 0x1p1023;
    return t * abm_pow2(n - 1023);
  }
  return p * abm_pow2(n);
}

ABM_MATH_FN double abm_exp(double x) {
  if (!(x == x)) return x + x;
  if (x > 0x1.62e42fefa39efp9) return abm_f64_from_bits(0x7ff0000000000000ull);
  if (x < -0x1.74910d52d3051p9) return 0.0; /* below 2^-1075 */
  const double n = __builtin_rint(x * 0x1.71547652b82fep0);
  /* ln2 = 0x1.62e42feep-1 (32 significant bits) + 0x1.a39ef35793c76p-33 */
  const double r = (x - n * 0x1.62e42feep-1) - n * 0x1.a39ef35793c76p-33;
  double q = 0x1.6124613a86d09p-33; /* 1/13! */
  q = 0x1.1eed8eff8d898p-29 + q * r; /* 1/12! */
  q = 0x1.ae64567f544e4p-26 + q * r; /* 1/11! */
  q = 0x1.27e4fb7789f5cp-22 + q * r; /* 1/10! */
  q = 0x1.71de3a556c734p-19 + q * r; /* 1/9! */
  q = 0x1.a01a01a01a01ap-16 + q * r; /* 1/8! */
  q = 0x1.a01a01a01a01ap-13 + q * r; /* 1/7! */
  q = 0x1.6c16c16c16c17p-10 + q * r; /* 1/6! */
  q = 0x1.1111111111111p-7 + q * r;  /* 1/5! */
  q = 0x1.5555555555555p-5 + q * r;  /* 1/4! */
  q = 0x1.5555555555555p-3 + q * r;  /* 1/3! */
  q = 0.5 + q * r;
  const double p = 1.0 + (r + (r * r) * q);
  return abm_scale2(p, (int)n);
}

ABM_MATH_FN double abm_log(double x) {
  if (!(x == x)) return x + x;
  if (x < 0.0) return abm_f64_from_bits(0x7ff8000000000000ull);
  if (x == 0.0) return abm_f64_from_bits(0xfff0000000000000ull);
  uint64_t b = abm_f64_bits(x);
  if (b == 0x7ff0000000000000ull) return x;
  int e = 0;
  if (b < 0x0010000000000000ull) {
    x = x * 0x1p54;
    b = abm_f64_bits(x);
    e = -54;
  }
  e += (int)(b >> 52) - 1023;
  double m = abm_f64_from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp0) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 0x1.642c8590b2164p-4; /* 2/23 */
  R = 0x1.8618618618618p-4 + R * z; /* 2/21 */
  R = 0x1.af286bca1af28p-4 + R * z; /* 2/19 */
  R = 0x1.e1e1e1e1e1e1ep-4 + R * z; /* 2/17 */
  R = 0x1.1111111111111p-3 + R * z; /* 2/15 */
  R = 0x1.3b13b13b13b14p-3 + R * z; /* 2/13 */
  R = 0x1.745d1745d1746p-3 + R * z; /* 2/11 */
  R = 0x1.c71c71c71c71cp-3 + R * z; /* 2/9 */
  R = 0x1.2492492492492p-2 + R * z; /* 2/7 */
  R = 0x1.999999999999ap-2 + R * z; /* 2/5 */
  R = 0x1.5555555555555p-1 + R * z; /* 2/3 */
  R = R * z;
  const double hfsq = 0.5 * (f * f);
  const double ef = (double)e;
  return ef * 0x1.62e42feep-1 - ((hfsq - (s * (hfsq + R) + ef * 0x1.a39ef35793c76p-33)) - f);
}

/* log1p(x) by Goldberg's method in float64, rounded to float32 once. */
ABM_MATH_FN float abm_log1pf(float x) {
  if (!(x == x)) return x + x;
  const double xd = (double)x;
  const double u = 1.0 + xd;
  if (u == 1.0) return x;
  if (u == abm_f64_from_bits(0x7ff0000000000000ull)) return (float)u;
  return (float)(abm_log(u) * (xd / (u - 1.0)));
}

/* x^y for x >= 0 (NaN for x < 0): exp(y log x) in float64, rounded to
 * float32 once. */
ABM_MATH_FN float abm_powf(float x, float y) {
  if (x == 1.0f) return 1.0f;
  return (float)abm_exp((double)y * abm_log((double)x));
}

#endif /* ABM_MATH_H */
