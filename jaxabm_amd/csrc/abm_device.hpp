// abm_device.hpp -- device arithmetic shared by the library's kernels and by
// kernels generated at run time (compiled with hiprtc from this same text).
//
// threefry2x32-20 (Random123, Salmon et al. SC'11) and JAX's counter layout
// for random bits, so that every kernel can draw element e of
// jax.random.bits(key, (n,)) directly from (key, n, e) without materialising
// the array (JAX call sites: jaxabm/model.py:46,129,156,164,183,
// jaxabm/agent.py:115,156); the uniform, normal and randint constructions of
// jax.random on top of them.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "abm_math.h"
#endif

#define ABM_HD __host__ __device__ __forceinline__

namespace abm {

enum : int { kModeOriginal = 0, kModePartitionable = 1 };
enum : int { kOk = 0, kErrInvalid = 1, kErrHip = 2, kErrNoMem = 3, kErrState = 4 };

struct Key {
  uint32_t k0, k1;
};

ABM_HD uint32_t rotl(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

// One threefry2x32-20 block.
ABM_HD void threefry(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1,
                     uint32_t &y0, uint32_t &y1) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  x0 += k0;
  x1 += k1;
#define ABM_R(r) \
  x0 += x1;      \
  x1 = rotl(x1, r) ^ x0;
  ABM_R(13) ABM_R(15) ABM_R(26) ABM_R(6) x0 += k1; x1 += k2 + 1u;
  ABM_R(17) ABM_R(29) ABM_R(16) ABM_R(24) x0 += k2; x1 += k0 + 2u;
  ABM_R(13) ABM_R(15) ABM_R(26) ABM_R(6) x0 += k0; x1 += k1 + 3u;
  ABM_R(17) ABM_R(29) ABM_R(16) ABM_R(24) x0 += k1; x1 += k2 + 4u;
  ABM_R(13) ABM_R(15) ABM_R(26) ABM_R(6) x0 += k2; x1 += k0 + 5u;
#undef ABM_R
  y0 = x0;
  y1 = x1;
}

// Element e of jax.random.bits(key, (n,)) (32-bit).
//  original:      iota(n) padded to even length, lane i hashes
//                 (i, h+i), output concat(y0s, y1s)[:n], h = ceil(n/2)
//  partitionable: y0 ^ y1 of threefry(key, (hi32(e), lo32(e)))
ABM_HD uint32_t bits_at(Key k, int64_t n, int64_t e, int mode) {
  uint32_t y0, y1;
  if (mode == kModePartitionable) {
    threefry(k.k0, k.k1, (uint32_t)((uint64_t)e >> 32), (uint32_t)e, y0, y1);
    return y0 ^ y1;
  }
  const int64_t h = (n + 1) >> 1;
  if (e < h) {
    const int64_t j = h + e;
    threefry(k.k0, k.k1, (uint32_t)e, j < n ? (uint32_t)j : 0u, y0, y1);
    return y0;
  }
  threefry(k.k0, k.k1, (uint32_t)(e - h), (uint32_t)e, y0, y1);
  return y1;
}

// Key i of jax.random.split(key, num).
ABM_HD Key split_at(Key k, int64_t num, int64_t i, int mode) {
  uint32_t y0, y1;
  if (mode == kModePartitionable) {
    threefry(k.k0, k.k1, (uint32_t)((uint64_t)i >> 32), (uint32_t)i, y0, y1);
    return Key{y0, y1};
  }
  // flat element 2i and 2i+1 of threefry_2x32(key, iota(2*num))
  const int64_t n = 2 * num;
  return Key{bits_at(k, n, 2 * i, kModeOriginal),
             bits_at(k, n, 2 * i + 1, kModeOriginal)};
}

// Both halves of split(key) (num = 2): original mode hashes lanes (0,2) and
// (1,3); partitionable hashes (0,0) and (0,1).
ABM_HD void split2(Key k, Key &a, Key &b, int mode) {
  if (mode == kModePartitionable) {
    threefry(k.k0, k.k1, 0u, 0u, a.k0, a.k1);
    threefry(k.k0, k.k1, 0u, 1u, b.k0, b.k1);
    return;
  }
  uint32_t p0, p1, q0, q1;
  threefry(k.k0, k.k1, 0u, 2u, p0, p1);
  threefry(k.k0, k.k1, 1u, 3u, q0, q1);
  a = Key{p0, q0};
  b = Key{p1, q1};
}

// Scalar draw: jax.random.bits(key, ()) == element 0 of a length-1 array.
ABM_HD uint32_t bits_scalar(Key k, int mode) { return bits_at(k, 1, 0, mode); }

// float32 in [0, 1) as jax.random.uniform builds it: mantissa fill of [1, 2)
// minus one.
ABM_HD float unit_float(uint32_t b) {
  return __uint_as_float((b >> 9) | 0x3F800000u) - 1.0f;
}

// uniform(key, lo, hi) element: max(lo, f * (hi - lo) + lo), no contraction.
ABM_HD float scale_uniform(float f, float lo, float span) {
  float t = f * span;
  float v = t + lo;
  return v > lo ? v : lo;
}

// jax.random.randint offset construction for one element, uint32 arithmetic.
ABM_HD uint32_t randint_offset(uint32_t hb, uint32_t lb, uint32_t span) {
  uint32_t mult = 65536u % span;
  mult = mult * mult;
  mult = mult % span;
  uint32_t off = (hb % span) * mult + (lb % span);
  return off % span;
}

// jnp.mod(x, 1.0): truncated remainder, plus one when negative and non-zero.
// For finite x the truncated remainder fmod(x, 1) is x - trunc(x) exactly
// (the fraction bits of x, subtracted without rounding) with the sign of x
// (-0 for a negative integer): two instructions instead of fmodf's loop.
__device__ __forceinline__ float mod1(float x) {
  float r = copysignf(x - truncf(x), x);
  if (r != 0.0f && r < 0.0f) r = r + 1.0f;
  return r;
}

// XLA's float32 erf_inv: Giles (2010) single-precision polynomial pair.
__device__ __forceinline__ float erfinv_f32(float x) {
  float w = -abm_log1pf(-x * x);
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f;
    p = 3.43273939e-07f + p * w;
    p = -3.5233877e-06f + p * w;
    p = -4.39150654e-06f + p * w;
    p = 0.00021858087f + p * w;
    p = -0.00125372503f + p * w;
    p = -0.00417768164f + p * w;
    p = 0.246640727f + p * w;
    p = 1.50140941f + p * w;
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = 0.000100950558f + p * w;
    p = 0.00134934322f + p * w;
    p = -0.00367342844f + p * w;
    p = 0.00573950773f + p * w;
    p = -0.0076224613f + p * w;
    p = 0.00943887047f + p * w;
    p = 1.00167406f + p * w;
    p = 2.83297682f + p * w;
  }
  if (fabsf(x) == 1.0f) return x * __builtin_inff();
  return p * x;
}

// jax.random.normal scalar from 32 random bits.
__device__ __forceinline__ float normal_from_bits(uint32_t b) {
  const float lo = -0.99999994f;  // nextafter(-1, 0)
  const float span = 2.0f;        // f32(1 - lo) rounds to 2
  float u = scale_uniform(unit_float(b), lo, span);
  return 1.41421354f * erfinv_f32(u);
}

// ---- gamma, log-gamma and beta draws (jax.random.gamma / loggamma / beta)
//
// Marsaglia & Tsang, "A simple method for generating gamma variables",
// ACM TOMS 26(3), 2000, with jax.random's key layout (jax/_src/random.py):
//  * element j of a draw of `count` elements uses key split(key, count)[j];
//  * that key splits in two: the first half drives the rejection loop, the
//    second draws the boost uniform;
//  * each rejection round splits its key in three (next round, normal,
//    uniform); the redraw of the normal while v <= 0 splits its key in two
//    (next, draw);
//  * alpha < 1 samples Gamma(alpha + 1) and scales by u^(1/alpha) with
//    u = 1 - uniform, or in log space adds log1p(-uniform) / alpha unless
//    that is zero.
// float32 throughout, as jax_enable_x64 off.  Rejection rounds are capped far
// beyond any reachable count (acceptance is at least 0.95 per round) so that
// every lane ends; a capped lane returns NaN.
constexpr int kGammaMaxRounds = 4096;

__device__ __forceinline__ bool gamma_reject(float X, float V, float U, float d) {
  const bool squeeze = U >= 1.0f - 0.0331f * (X * X);
  const bool full = abm_logf(U) >= X * 0.5f + d * ((1.0f - V) + abm_logf(V));
  return squeeze && full;
}

__device__ inline float gamma_one(Key key, float alpha, bool log_space, int mode) {
  const bool boost = alpha >= 1.0f;
  const float a = boost ? alpha : alpha + 1.0f;
  const float third = 0.333333343f;   // float32(1/3)
  const float d = a - third;
  const float c = third / sqrtf(d);
  Key k, sub;
  split2(key, k, sub, mode);
  float X = 0.0f, V = 1.0f, U = 2.0f;   // a start the loop condition rejects
  for (int round = 0; gamma_reject(X, V, U, d); ++round) {
    if (round == kGammaMaxRounds) return __builtin_nanf("");
    const Key next = split_at(k, 3, 0, mode);
    Key xk = split_at(k, 3, 1, mode);
    const Key uk = split_at(k, 3, 2, mode);
    k = next;
    float x = 0.0f, v = -1.0f;
    for (int redraw = 0; v <= 0.0f; ++redraw) {
      if (redraw == kGammaMaxRounds) return __builtin_nanf("");
      Key nk, dk;
      split2(xk, nk, dk, mode);
      xk = nk;
      x = normal_from_bits(bits_scalar(dk, mode));
      v = 1.0f + x * c;
    }
    X = x * x;
    V = (v * v) * v;
    U = unit_float(bits_scalar(uk, mode));
  }
  const float u = unit_float(bits_scalar(sub, mode));
  if (log_space) {
    const float ls = abm_log1pf(-u);
    const float lb = (boost || ls == 0.0f) ? 0.0f : ls * (1.0f / alpha);
    return (abm_logf(d) + abm_logf(V)) + lb;
  }
  const float s = 1.0f - u;
  const float b = boost ? 1.0f : abm_powf(s, 1.0f / alpha);
  return (d * V) * b;
}

// Element j of gamma(key, alpha, (count,)) (log_space: loggamma).
__device__ inline float gamma_draw(Key key, int64_t count, int64_t j, float alpha,
                                   bool log_space, int mode) {
  return gamma_one(split_at(key, count, j, mode), alpha, log_space, mode);
}

// Element j of beta(key, a, b, (count,)): split(key) into the keys of
// loggamma(a) and loggamma(b), then exp(la - m) / (exp(la - m) + exp(lb - m))
// with m = max(la, lb) (NaN if either is NaN).
__device__ inline float beta_draw(Key key, int64_t count, int64_t j, float a, float b,
                                  int mode) {
  Key ka, kb;
  split2(key, ka, kb, mode);
  const float la = gamma_draw(ka, count, j, a, true, mode);
  const float lb = gamma_draw(kb, count, j, b, true, mode);
  const float m = (la != la || lb != lb) ? la + lb : (la > lb ? la : lb);
  const float ea = abm_expf(la - m), eb = abm_expf(lb - m);
  return ea / (ea + eb);
}

}  // namespace abm
