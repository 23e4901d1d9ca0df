/*
 * oracle.c -- CPU restatement of the reference's agent-step semantics.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.
 * The shipped engine never links or calls this file.
 *
 * Parity pins (each function's header names what pins it):
 *   threefry block      -- Random123 known-answer vectors (tests/test_oracle_rng.py)
 *   split / bits / uniform -- JAX-documented values of PRNGKey(0) in both modes
 *   randint             -- closed form + distribution + span==1 special case
 *   normal / erfinv     -- Giles (2010) polynomial vs scipy.special.erfinv
 *   permutation         -- brute-force stable sort on tiny n, invariants
 *   BA / CSR            -- edge count m(n-m), symmetry, degree sum, no dup
 *   Gini                -- brute-force pairwise |xi-xj| identity
 *   programs            -- invariants (conservation, monotone R, one agent
 *                          per cell, sum of opinions) + brute-force step
 *                          re-evaluation in tests/test_oracle_programs.py
 *   Deffuant adjoint    -- central finite differences of the float64
 *                          forward of the same program: agreement to 1e-5
 *                          at h = 1e-6 and monotone convergence from
 *                          h = 1e-3 (the gate's curvature separates the
 *                          secant from the tangent at h = 1e-3)
 *
 *   exp / log / log1p / pow -- the shared IEEE-basic-operation functions of
 *                          jaxabm_amd/csrc/abm_math.h (also compiled into the
 *                          device kernels, so both compute the same bits),
 *                          pinned against float64 numpy to 1 ulp and by
 *                          identities in tests/test_oracle_math.py
 *
 * Parity unpinned (no JAX in this environment to produce vectors): XLA:CPU's
 * own rounding of log1p inside erf_inv, of exp in the sigmoid gate and of
 * log / pow / exp in the gamma sampler (this restatement and the device use
 * abm_math.h, accurate to about one ulp, where XLA uses its own
 * approximations), and end-to-end trajectories under an installed JAX
 * version beyond the documented key and draw values above.
 */
#include "oracle.h"

#include "abm_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* threefry2x32-20: Salmon et al., "Parallel random numbers: as easy as 1, 2,
 * 3" (SC'11), the block function JAX wraps (call sites jaxabm/model.py:46,
 * 129,156,164,183; jaxabm/agent.py:115,156).  Five groups of four rounds with
 * rotation sets (13,15,26,6) / (17,29,16,24), key injection after each group.
 */
static inline uint32_t rotl32(uint32_t v, int r) {
  return (v << r) | (v >> (32 - r));
}

void orc_threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1,
                      uint32_t out[2]) {
  static const int rot[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  x0 += ks[0];
  x1 += ks[1];
  for (int g = 0; g < 5; ++g) {
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl32(x1, rot[g & 1][r]);
      x1 ^= x0;
    }
    int i = g + 1;
    x0 += ks[i % 3];
    x1 += ks[(i + 1) % 3] + (uint32_t)i;
  }
  out[0] = x0;
  out[1] = x1;
}

/* PRNGKey(seed) with 32-bit seeds (jax_enable_x64 off): [seed >> 32, seed &
 * 0xFFFFFFFF] where the shift of a 32-bit value by 32 is 0.  Reference call
 * site: jaxabm/model.py:46. */
void orc_prng_key(int64_t seed, uint32_t key[2]) {
  key[0] = 0u;
  key[1] = (uint32_t)(seed & 0xFFFFFFFFll);
}

/* Element e of JAX's threefry_2x32(key, iota(n)) in the original stream: the
 * count vector is padded to even length m and split in halves; lane i hashes
 * (count[i], count[h+i]) and the output is concat(lane_out0, lane_out1)[:n]. */
static uint32_t orig_element(const uint32_t key[2], int64_t n, int64_t e) {
  int64_t h = (n + 1) / 2;
  uint32_t o[2];
  if (e < h) {
    int64_t j = h + e;
    uint32_t x1 = (j < n) ? (uint32_t)j : 0u;
    orc_threefry2x32(key[0], key[1], (uint32_t)e, x1, o);
    return o[0];
  }
  orc_threefry2x32(key[0], key[1], (uint32_t)(e - h), (uint32_t)e, o);
  return o[1];
}

/* jax.random.split(key, num).  Original: threefry_2x32(key, iota(2*num))
 * reshaped (num, 2).  Partitionable: key_i = threefry(key, (hi32(i), lo32(i))).
 * Pinned by split(PRNGKey(0)) documented values in both modes. */
void orc_split(const uint32_t key[2], int64_t num, uint32_t *out, int mode) {
  if (mode == ORC_MODE_PARTITIONABLE) {
    for (int64_t i = 0; i < num; ++i) {
      uint32_t o[2];
      orc_threefry2x32(key[0], key[1], (uint32_t)((uint64_t)i >> 32),
                       (uint32_t)i, o);
      out[2 * i] = o[0];
      out[2 * i + 1] = o[1];
    }
    return;
  }
  int64_t n = 2 * num;
  for (int64_t i = 0; i < num; ++i) {
    uint32_t o[2];
    orc_threefry2x32(key[0], key[1], (uint32_t)i, (uint32_t)(num + i), o);
    /* o[0] is flat element i, o[1] is flat element num+i */
    out[i] = o[0];
    out[num + i] = o[1];
  }
  (void)n;
}

/* jax.random.bits / _random_bits(key, 32, (n,)).  Original: threefry_2x32
 * over iota(n) (n < 2^32-1 so a single block).  Partitionable: y0 ^ y1 of
 * threefry(key, (hi32(j), lo32(j))). */
void orc_random_bits32(const uint32_t key[2], int64_t n, uint32_t *out,
                       int mode) {
  if (mode == ORC_MODE_PARTITIONABLE) {
    for (int64_t j = 0; j < n; ++j) {
      uint32_t o[2];
      orc_threefry2x32(key[0], key[1], (uint32_t)((uint64_t)j >> 32),
                       (uint32_t)j, o);
      out[j] = o[0] ^ o[1];
    }
    return;
  }
  for (int64_t e = 0; e < n; ++e) out[e] = orig_element(key, n, e);
}

static inline float bits_to_unit(uint32_t b) {
  uint32_t fb = (b >> 9) | 0x3F800000u;
  float f;
  memcpy(&f, &fb, 4);
  return f - 1.0f;
}

/* jax.random.uniform(key, (n,), float32, lo, hi): mantissa fill of [1,2),
 * minus 1, scale by (hi-lo) in f32, add lo, max with lo. */
void orc_uniform_f32(const uint32_t key[2], int64_t n, float lo, float hi,
                     float *out, int mode) {
  uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  orc_random_bits32(key, n, b, mode);
  volatile float span = hi - lo;
  for (int64_t i = 0; i < n; ++i) {
    volatile float t = bits_to_unit(b[i]) * span;
    float v = t + lo;
    out[i] = v > lo ? v : lo;
  }
  free(b);
}

/* jax.random.randint(key, (n,), lo, hi, int32): k1, k2 = split(key);
 * higher/lower bits from each; offset = ((hb % span) * mult + lb % span) % span
 * in uint32 arithmetic, mult = ((2^16 % span)^2) % span; span = 1 if hi<=lo. */
void orc_randint_i32(const uint32_t key[2], int64_t n, int32_t lo, int32_t hi,
                     int32_t *out, int mode) {
  uint32_t ks[4];
  orc_split(key, 2, ks, mode);
  uint32_t *hb = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t *lb = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  orc_random_bits32(ks, n, hb, mode);
  orc_random_bits32(ks + 2, n, lb, mode);
  uint32_t span = (hi <= lo) ? 1u : (uint32_t)(hi - lo);
  uint32_t mult = 65536u % span;
  mult = mult * mult;
  mult = mult % span;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t off = (hb[i] % span) * mult + (lb[i] % span);
    off = off % span;
    out[i] = (int32_t)((uint32_t)lo + off);
  }
  free(hb);
  free(lb);
}

/* The transcendental functions the float32 programs use, from abm_math.h,
 * elementwise (pinned against float64 numpy in tests/test_oracle_math.py). */
void orc_math_f32(int fn, const float *x, const float *y, float *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = abm_expf(x[i]); break;
      case 1: out[i] = abm_logf(x[i]); break;
      case 2: out[i] = abm_log1pf(x[i]); break;
      default: out[i] = abm_powf(x[i], y[i]); break;
    }
  }
}

void orc_math_f64(int fn, const double *x, double *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = fn == 0 ? abm_exp(x[i]) : abm_log(x[i]);
}

/* erf_inv for float32: M. Giles, "Approximating the erfinv function", GPU
 * Computing Gems Jade Edition (2010), single-precision branch pair; the form
 * XLA lowers lax.erf_inv to for f32. */
float orc_erfinv_f32(float x) {
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
  if (fabsf(x) == 1.0f) return x * INFINITY;
  return p * x;
}

/* jax.random.normal(key, (n,), float32) = sqrt(2) * erf_inv(u),
 * u = uniform(key, lo=nextafter(-1, 0), hi=1). */
void orc_normal_f32(const uint32_t key[2], int64_t n, float *out, int mode) {
  float lo = nextafterf(-1.0f, 0.0f);
  orc_uniform_f32(key, n, lo, 1.0f, out, mode);
  const float sqrt2 = (float)1.4142135623730951;
  for (int64_t i = 0; i < n; ++i) out[i] = sqrt2 * orc_erfinv_f32(out[i]);
}

/* ---- gamma, log-gamma and beta: jax.random.gamma / loggamma / beta (the
 * reference's examples/models/advanced_economic_model.py:106,351,635,958,962
 * call beta).  Marsaglia & Tsang, "A simple method for generating gamma
 * variables", ACM TOMS 26(3), 2000, in JAX's key layout (JAX is an
 * un-vendored dependency, jax>=0.4.1 unpinned, requirements.txt:9-10):
 *   element j of a draw of `count` values uses k = split(key, count)[j];
 *   (k_loop, k_boost) = split(k);
 *   a rejection round: (k_loop, k_x, k_u) = split(k_loop, 3); then, while
 *   v <= 0: (k_x, k_d) = split(k_x), x = normal(k_d), v = 1 + x c;
 *   X = x^2, V = v^3, U = uniform(k_u); the round repeats while
 *   U >= 1 - 0.0331 X^2 and log U >= X/2 + d (1 - V + log V), starting from
 *   (X, V, U) = (0, 1, 2);
 *   alpha < 1 runs with alpha + 1 and multiplies by (1 - uniform(k_boost))^
 *   (1/alpha), or in log space adds log1p(-uniform(k_boost)) * (1/alpha)
 *   unless that logarithm is 0;
 *   beta(a, b): (ka, kb) = split(key), la = loggamma(ka, a), lb =
 *   loggamma(kb, b), m = max(la, lb), exp(la - m) / (exp(la - m) +
 *   exp(lb - m)).
 * Bit parity with JAX is unpinned (JAX cannot run here); the tests pin the
 * sampled distributions against scipy.stats. */
static float orc_gamma_one(const uint32_t key[2], float alpha, int log_space, int mode) {
  const int boost = alpha >= 1.0f;
  const float a = boost ? alpha : alpha + 1.0f;
  const float third = (float)(1.0 / 3.0);
  const float d = a - third;
  const float c = third / sqrtf(d);
  uint32_t ks[4], k[2], sub[2];
  orc_split(key, 2, ks, mode);
  k[0] = ks[0]; k[1] = ks[1];
  sub[0] = ks[2]; sub[1] = ks[3];
  float X = 0.0f, V = 1.0f, U = 2.0f;
  for (;;) {
    const float squeeze = 1.0f - 0.0331f * (X * X);
    const float rhs = X * 0.5f + d * ((1.0f - V) + abm_logf(V));
    if (!(U >= squeeze && abm_logf(U) >= rhs)) break;
    uint32_t k3[6], xk[2], uk[2];
    orc_split(k, 3, k3, mode);
    k[0] = k3[0]; k[1] = k3[1];
    xk[0] = k3[2]; xk[1] = k3[3];
    uk[0] = k3[4]; uk[1] = k3[5];
    float x = 0.0f, v = -1.0f;
    while (v <= 0.0f) {
      uint32_t k2[4];
      orc_split(xk, 2, k2, mode);
      xk[0] = k2[0]; xk[1] = k2[1];
      orc_normal_f32(k2 + 2, 1, &x, mode);
      v = 1.0f + x * c;
    }
    X = x * x;
    V = (v * v) * v;
    orc_uniform_f32(uk, 1, 0.0f, 1.0f, &U, mode);
  }
  float u;
  orc_uniform_f32(sub, 1, 0.0f, 1.0f, &u, mode);
  if (log_space) {
    const float ls = abm_log1pf(-u);
    const float lb = (boost || ls == 0.0f) ? 0.0f : ls * (1.0f / alpha);
    return (abm_logf(d) + abm_logf(V)) + lb;
  }
  const float s = 1.0f - u;
  const float b = boost ? 1.0f : abm_powf(s, 1.0f / alpha);
  return (d * V) * b;
}

void orc_gamma_f32(const uint32_t key[2], int64_t count, const float *alpha, int log_space,
                   float *out, int mode) {
  uint32_t *ks = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(count > 0 ? count : 1));
  orc_split(key, count, ks, mode);
  for (int64_t j = 0; j < count; ++j) out[j] = orc_gamma_one(ks + 2 * j, alpha[j], log_space, mode);
  free(ks);
}

void orc_beta_f32(const uint32_t key[2], int64_t count, const float *a, const float *b,
                  float *out, int mode) {
  uint32_t kab[4];
  orc_split(key, 2, kab, mode);
  float *la = (float *)malloc(sizeof(float) * (size_t)(count > 0 ? count : 1));
  float *lb = (float *)malloc(sizeof(float) * (size_t)(count > 0 ? count : 1));
  orc_gamma_f32(kab, count, a, 1, la, mode);
  orc_gamma_f32(kab + 2, count, b, 1, lb, mode);
  for (int64_t j = 0; j < count; ++j) {
    const float m = (isnan(la[j]) || isnan(lb[j])) ? la[j] + lb[j] : (la[j] > lb[j] ? la[j] : lb[j]);
    const float ea = abm_expf(la[j] - m), eb = abm_expf(lb[j] - m);
    out[j] = ea / (ea + eb);
  }
  free(la);
  free(lb);
}

/* jax.random.permutation(key, n) via _shuffle: rounds = ceil(3 ln(max(1,n)) /
 * ln(2^32-1)); each round key, sub = split(key); sort_key_val(bits(sub), x)
 * stable by key. */
int32_t orc_permutation_rounds(int64_t n) {
  double nn = (double)(n > 1 ? n : 1);
  return (int32_t)ceil(3.0 * log(nn) / log(4294967295.0));
}

typedef struct {
  uint32_t k;
  int32_t v;
  int64_t pos;
} kv_t;

static int kv_cmp(const void *a, const void *b) {
  const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos ? 1 : 0);
}

void orc_permutation(const uint32_t key[2], int64_t n, int32_t *out,
                     int mode) {
  for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)i;
  int32_t rounds = orc_permutation_rounds(n);
  uint32_t k[2] = {key[0], key[1]};
  uint32_t *bits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t r = 0; r < rounds; ++r) {
    uint32_t ks[4];
    orc_split(k, 2, ks, mode);
    k[0] = ks[0];
    k[1] = ks[1];
    orc_random_bits32(ks + 2, n, bits, mode);
    for (int64_t i = 0; i < n; ++i) {
      kv[i].k = bits[i];
      kv[i].v = out[i];
      kv[i].pos = i;
    }
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);
    for (int64_t i = 0; i < n; ++i) out[i] = kv[i].v;
  }
  free(bits);
  free(kv);
}

/* ------------------------------------------------------------------------ */
/* Barabasi-Albert graph, the program's own deterministic generator (the
 * reference has only Network.add_edge, jaxabm/agentpy.py:559-582).  Start from
 * a star on nodes 0..m (centre 0); node s = m+1..n-1 draws targets from the
 * degree-repeated node list with index floor(r * len / 2^32),
 * r = threefry(key, (s, attempt)).y0, rejecting repeats until m distinct. */
int64_t orc_ba_num_edges(int64_t n, int32_t m) {
  if (n <= m) return n > 0 ? n - 1 : 0;
  return (int64_t)m * (n - m);
}

int64_t orc_ba_edges(int64_t n, int32_t m, const uint32_t key[2],
                     int32_t *src, int32_t *dst) {
  int64_t E = 0;
  if (n <= 1) return 0;
  if (n <= m) {
    for (int64_t i = 1; i < n; ++i) {
      src[E] = 0;
      dst[E] = (int32_t)i;
      ++E;
    }
    return E;
  }
  int64_t cap = 2 * orc_ba_num_edges(n, m) + 2;
  int32_t *rep = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  int64_t len = 0;
  for (int32_t i = 1; i <= m; ++i) {
    src[E] = 0;
    dst[E] = i;
    ++E;
  }
  for (int32_t i = 0; i < m; ++i) rep[len++] = 0;
  for (int32_t i = 1; i <= m; ++i) rep[len++] = i;
  int32_t *chosen = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
  for (int64_t s = m + 1; s < n; ++s) {
    int32_t c = 0;
    uint32_t attempt = 0;
    while (c < m) {
      uint32_t o[2];
      orc_threefry2x32(key[0], key[1], (uint32_t)s, attempt++, o);
      int64_t idx = (int64_t)(((uint64_t)o[0] * (uint64_t)len) >> 32);
      int32_t t = rep[idx];
      int dup = 0;
      for (int32_t q = 0; q < c; ++q)
        if (chosen[q] == t) dup = 1;
      if (!dup) chosen[c++] = t;
    }
    for (int32_t q = 0; q < m; ++q) {
      src[E] = (int32_t)s;
      dst[E] = chosen[q];
      ++E;
      rep[len++] = chosen[q];
    }
    for (int32_t q = 0; q < m; ++q) rep[len++] = (int32_t)s;
  }
  free(chosen);
  free(rep);
  return E;
}

static int i32_cmp(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* Undirected CSR: both directions of every edge; row v holds its neighbours
 * in ascending id order (the observable order of Network.get_neighbors,
 * jaxabm/agentpy.py:605-613, which returns jnp.unique'd ids). */
void orc_csr_undirected(int64_t n, int64_t n_edges, const int32_t *src,
                        const int32_t *dst, int32_t *row_ptr, int32_t *col) {
  for (int64_t v = 0; v <= n; ++v) row_ptr[v] = 0;
  for (int64_t e = 0; e < n_edges; ++e) {
    row_ptr[src[e] + 1]++;
    row_ptr[dst[e] + 1]++;
  }
  for (int64_t v = 0; v < n; ++v) row_ptr[v + 1] += row_ptr[v];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t v = 0; v < n; ++v) fill[v] = row_ptr[v];
  for (int64_t e = 0; e < n_edges; ++e) {
    col[fill[src[e]]++] = dst[e];
    col[fill[dst[e]]++] = src[e];
  }
  for (int64_t v = 0; v < n; ++v)
    qsort(col + row_ptr[v], (size_t)(row_ptr[v + 1] - row_ptr[v]),
          sizeof(int32_t), i32_cmp);
  free(fill);
}

/* ------------------------------------------------------------------------ */
/* Gini, reference definition examples/models/advanced_economic_model.py:
 * 1798-1811 under jax_enable_x64=False: x = sort(v) (int32); S = sum(i*x_i)
 * and T = sum(x) in int32 (wrapping); G = f32(2*S) / f32(n*T) - f32((n+1)/n)
 * if T > 1e-6 else 0. */
float orc_gini_i32(const int32_t *v, int64_t n) {
  if (n <= 0) return 0.0f;
  int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  memcpy(x, v, sizeof(int32_t) * (size_t)n);
  qsort(x, (size_t)n, sizeof(int32_t), i32_cmp);
  uint32_t S = 0, T = 0;
  for (int64_t i = 0; i < n; ++i) {
    S += (uint32_t)(i + 1) * (uint32_t)x[i];
    T += (uint32_t)x[i];
  }
  free(x);
  int32_t Ti = (int32_t)T;
  if (!((float)Ti > 1e-6f)) return 0.0f;
  int32_t num = (int32_t)(2u * S);
  int32_t den = (int32_t)((uint32_t)n * T);
  volatile float q = (float)num / (float)den;
  float c = (float)((double)(n + 1) / (double)n);
  return q - c;
}

/* jnp.mean of an int32 vector: f32 sum / f32 n (exact for |sum| < 2^24). */
float orc_mean_i32(const int32_t *v, int64_t n) {
  int64_t s = 0;
  for (int64_t i = 0; i < n; ++i) s += v[i];
  volatile float q = (float)s / (float)n;
  return q;
}

/* ------------------------------------------------------------------------ */
/* Reference step-key schedule, jaxabm/model.py:129-130 (initialize) and
 * 156, 164, 183 (step).  Every program below has one agent collection and an
 * env-phase callback, so a step draws: rng, sk = split(rng);
 * sk, ck = split(sk) (collection key); sk, uk = split(sk) (env-phase key). */
static void split2(const uint32_t k[2], uint32_t a[2], uint32_t b[2],
                   int mode) {
  uint32_t o[4];
  orc_split(k, 2, o, mode);
  a[0] = o[0];
  a[1] = o[1];
  b[0] = o[2];
  b[1] = o[3];
}

static void init_keys(int64_t seed, int mode, uint32_t rng[2],
                      uint32_t init_key[2]) {
  uint32_t k[2];
  orc_prng_key(seed, k);
  split2(k, rng, init_key, mode); /* split(rng, C+1) with C = 1 */
}

static void step_keys(uint32_t rng[2], uint32_t ck[2], uint32_t uk[2],
                      int mode) {
  uint32_t sk[2], nr[2], s2[2];
  split2(rng, nr, sk, mode);
  rng[0] = nr[0];
  rng[1] = nr[1];
  split2(sk, s2, ck, mode);
  split2(s2, sk, uk, mode);
}

/* ------------------------------------------------------------------------ */
/* Network SIR program (reference API: one 'people' collection with identity
 * update, env-phase callback with update_key; jaxabm/model.py:146-216).
 * Init: rng, init_key = split(PRNGKey(seed)); infected = permutation(init_key,
 * n)[:n0].  Step: k_inf, k_rec = split(uk); u = uniform(k_inf, (nnz,));
 * r = uniform(k_rec, (n,)); S->I if some CSR entry e of row v has
 * old[col[e]] == I and u[e] < beta; I->R if r[v] < gamma.  Records (S,I,R). */
void orc_sir_init(int64_t n, int32_t n_seed_infected, int64_t seed, int mode,
                  uint8_t *state, uint32_t rng[2]) {
  uint32_t ik[2];
  init_keys(seed, mode, rng, ik);
  memset(state, 0, (size_t)n);
  int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  orc_permutation(ik, n, perm, mode);
  int64_t n0 = n_seed_infected < n ? n_seed_infected : n;
  for (int64_t i = 0; i < n0; ++i) state[perm[i]] = 1;
  free(perm);
}

static uint8_t sir_next(int64_t v, int64_t n, int64_t nnz, const int32_t *row_ptr,
                        const int32_t *col, const uint8_t *old, const uint32_t kinf[2],
                        const uint32_t krec[2], float beta, float gamma, int mode) {
  uint8_t s = old[v];
  if (s == 0) {
    for (int64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
      if (old[col[e]] != 1) continue;
      uint32_t b;
      if (mode == ORC_MODE_PARTITIONABLE) {
        uint32_t o[2];
        orc_threefry2x32(kinf[0], kinf[1], (uint32_t)((uint64_t)e >> 32),
                         (uint32_t)e, o);
        b = o[0] ^ o[1];
      } else {
        b = orig_element(kinf, nnz, e);
      }
      if (bits_to_unit(b) < beta) return 1;
    }
    return 0;
  }
  if (s == 1) {
    uint32_t b;
    if (mode == ORC_MODE_PARTITIONABLE) {
      uint32_t o[2];
      orc_threefry2x32(krec[0], krec[1], (uint32_t)((uint64_t)v >> 32),
                       (uint32_t)v, o);
      b = o[0] ^ o[1];
    } else {
      b = orig_element(krec, n, v);
    }
    if (bits_to_unit(b) < gamma) return 2;
  }
  return s;
}

/* The same run with the agents of each step split over `threads` OpenMP
 * threads: every agent's next state depends only on the pre-step copy, so
 * the result does not depend on the thread count. */
void orc_sir_run_threads(int64_t n, const int32_t *row_ptr, const int32_t *col,
                         uint8_t *state, uint32_t rng[2], float beta, float gamma,
                         int32_t steps, int mode, int32_t *counts, int threads) {
  int64_t nnz = row_ptr[n];
  uint8_t *old = (uint8_t *)malloc((size_t)n);
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2], kinf[2], krec[2];
    step_keys(rng, ck, uk, mode);
    split2(uk, kinf, krec, mode);
    memcpy(old, state, (size_t)n);
    int32_t c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : c0, c1, c2)
    for (int64_t v = 0; v < n; ++v) {
      const uint8_t s = sir_next(v, n, nnz, row_ptr, col, old, kinf, krec, beta, gamma, mode);
      state[v] = s;
      c0 += s == 0;
      c1 += s == 1;
      c2 += s == 2;
    }
    counts[3 * t + 0] = c0;
    counts[3 * t + 1] = c1;
    counts[3 * t + 2] = c2;
  }
  free(old);
}

void orc_sir_run(int64_t n, const int32_t *row_ptr, const int32_t *col,
                 uint8_t *state, uint32_t rng[2], float beta, float gamma,
                 int32_t steps, int mode, int32_t *counts) {
  orc_sir_run_threads(n, row_ptr, col, state, rng, beta, gamma, steps, mode, counts, 1);
}

/* ------------------------------------------------------------------------ */
/* jnp.mod(x, 1.0) for float32: r = fmod(x, 1); r + 1 if r != 0 and r < 0. */
static inline float mod1(float x) {
  float r = fmodf(x, 1.0f);
  if (r != 0.0f && r < 0.0f) {
    volatile float t = r + 1.0f;
    r = t;
  }
  return r;
}

/* Wealth-exchange random walk (reference API: 'walkers' collection with a
 * per-agent update that uses its agent key, env-phase callback that performs
 * the exchange; jaxabm/model.py:146-216, jaxabm/agent.py:132-177).
 * Init (agent keys = split(init_key, n)): kx, ky = split(agent_key);
 * x = uniform(kx), y = uniform(ky), wealth = 1.
 * Agent phase (agent keys = split(ck, n)): k1, k2 = split(agent_key);
 * x = mod(x + 0.01*normal(k1), 1), y = mod(y + 0.01*normal(k2), 1).
 * Env phase (uk): t = randint(uk, (n,), 0, n-1); t += (t >= i);
 * give = wealth > 0; wealth = wealth - give + bincount(t, give).
 * Record: mean(wealth), Gini(wealth). */
static void agent_xy_draw(const uint32_t akey[2], float *a, float *b,
                          int normal, int mode) {
  uint32_t k1[2], k2[2];
  split2(akey, k1, k2, mode);
  if (normal) {
    orc_normal_f32(k1, 1, a, mode);
    orc_normal_f32(k2, 1, b, mode);
  } else {
    orc_uniform_f32(k1, 1, 0.0f, 1.0f, a, mode);
    orc_uniform_f32(k2, 1, 0.0f, 1.0f, b, mode);
  }
}

/* ------------------------------------------------------------------------ */
/* Generic trader model (the per-agent code of bench.py's generic workload,
 * written against the reference API: one collection with an update and an
 * env-phase callback, jaxabm/model.py:118-216).  Init with agent key a:
 * k1, k2 = split(a); x = uniform(k1); w = randint(k2, (), 0, 6).  Step with
 * agent key a: k1, k2 = split(a); x = mod(x + normal(k1) * 0.01, 1);
 * w = (uniform(k2) < 0.5 and w > 0) ? w - 1 : w + 1. */
void orc_trader_init(int64_t n, int64_t seed, int mode, float *x, int32_t *w,
                     uint32_t rng[2]) {
  uint32_t ik[2];
  init_keys(seed, mode, rng, ik);
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(n > 0 ? n : 1));
  orc_split(ik, n, ak, mode);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t k1[2], k2[2];
    split2(ak + 2 * i, k1, k2, mode);
    orc_uniform_f32(k1, 1, 0.0f, 1.0f, &x[i], mode);
    orc_randint_i32(k2, 1, 0, 6, &w[i], mode);
  }
  free(ak);
}

void orc_trader_run(int64_t n, float *x, int32_t *w, uint32_t rng[2], int32_t steps,
                    int mode) {
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(n > 0 ? n : 1));
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2];
    step_keys(rng, ck, uk, mode);
    orc_split(ck, n, ak, mode);
    for (int64_t i = 0; i < n; ++i) {
      uint32_t k1[2], k2[2];
      split2(ak + 2 * i, k1, k2, mode);
      float z, u;
      orc_normal_f32(k1, 1, &z, mode);
      orc_uniform_f32(k2, 1, 0.0f, 1.0f, &u, mode);
      volatile float d = z * 0.01f;
      x[i] = mod1(x[i] + d);
      w[i] = (u < 0.5f && w[i] > 0) ? w[i] - 1 : w[i] + 1;
    }
  }
  free(ak);
}

void orc_wealth_init(int64_t n, int64_t seed, int mode, float *x, float *y,
                     int32_t *wealth, uint32_t rng[2]) {
  uint32_t ik[2];
  init_keys(seed, mode, rng, ik);
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)n);
  orc_split(ik, n, ak, mode);
  for (int64_t i = 0; i < n; ++i) {
    agent_xy_draw(ak + 2 * i, &x[i], &y[i], 0, mode);
    wealth[i] = 1;
  }
  free(ak);
}

void orc_wealth_run(int64_t n, float *x, float *y, int32_t *wealth,
                    uint32_t rng[2], int32_t steps, int mode, float *mean_out,
                    float *gini_out) {
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)n);
  int32_t *tgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int32_t *nw = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  const float step_sd = 0.01f;
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2];
    step_keys(rng, ck, uk, mode);
    orc_split(ck, n, ak, mode);
    for (int64_t i = 0; i < n; ++i) {
      float dx, dy;
      agent_xy_draw(ak + 2 * i, &dx, &dy, 1, mode);
      volatile float mx = step_sd * dx;
      volatile float my = step_sd * dy;
      volatile float sx = x[i] + mx;
      volatile float sy = y[i] + my;
      x[i] = mod1(sx);
      y[i] = mod1(sy);
    }
    orc_randint_i32(uk, n, 0, (int32_t)(n - 1), tgt, mode);
    for (int64_t i = 0; i < n; ++i) nw[i] = wealth[i];
    for (int64_t i = 0; i < n; ++i) {
      int32_t ti = tgt[i] + (tgt[i] >= i ? 1 : 0);
      if (wealth[i] > 0) {
        nw[i] -= 1;
        nw[ti] += 1;
      }
    }
    memcpy(wealth, nw, sizeof(int32_t) * (size_t)n);
    mean_out[t] = orc_mean_i32(wealth, n);
    gini_out[t] = orc_gini_i32(wealth, n);
  }
  free(ak);
  free(tgt);
  free(nw);
}

/* Element f of JAX's threefry_2x32(key, iota(2 num)) as orc_split lays it
 * out (key i of split(key, num) is elements 2i and 2i+1), and element j of
 * random_bits(key, (n,)): the per-element forms of orc_split and
 * orc_random_bits32, for loops over agents on several threads. */
static uint32_t split_element(const uint32_t key[2], int64_t num, int64_t f, int mode) {
  uint32_t o[2];
  if (mode == ORC_MODE_PARTITIONABLE) {
    const int64_t i = f >> 1;
    orc_threefry2x32(key[0], key[1], (uint32_t)((uint64_t)i >> 32), (uint32_t)i, o);
    return o[f & 1];
  }
  if (f < num) {
    orc_threefry2x32(key[0], key[1], (uint32_t)f, (uint32_t)(num + f), o);
    return o[0];
  }
  orc_threefry2x32(key[0], key[1], (uint32_t)(f - num), (uint32_t)f, o);
  return o[1];
}

static uint32_t bits_element(const uint32_t key[2], int64_t n, int64_t j, int mode) {
  if (mode == ORC_MODE_PARTITIONABLE) {
    uint32_t o[2];
    orc_threefry2x32(key[0], key[1], (uint32_t)((uint64_t)j >> 32), (uint32_t)j, o);
    return o[0] ^ o[1];
  }
  return orig_element(key, n, j);
}

/* orc_wealth_run with the per-agent work of each step (agent keys, position
 * draws, exchange targets) on `threads` OpenMP threads: the same results,
 * since every draw is a function of (key, index) alone; the exchange and the
 * records stay sequential. */
void orc_wealth_run_threads(int64_t n, float *x, float *y, int32_t *wealth, uint32_t rng[2],
                            int32_t steps, int mode, float *mean_out, float *gini_out,
                            int threads) {
  int32_t *tgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int32_t *nw = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  const float step_sd = 0.01f;
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2], ks[4];
    step_keys(rng, ck, uk, mode);
    orc_split(uk, 2, ks, mode);
    const uint32_t span = n - 1 <= 0 ? 1u : (uint32_t)(n - 1);
    uint32_t mult = 65536u % span;
    mult = mult * mult;
    mult = mult % span;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t ak[2] = {split_element(ck, n, 2 * i, mode),
                              split_element(ck, n, 2 * i + 1, mode)};
      float dx, dy;
      agent_xy_draw(ak, &dx, &dy, 1, mode);
      volatile float mx = step_sd * dx;
      volatile float my = step_sd * dy;
      volatile float sx = x[i] + mx;
      volatile float sy = y[i] + my;
      x[i] = mod1(sx);
      y[i] = mod1(sy);
      const uint32_t hb = bits_element(ks, n, i, mode), lb = bits_element(ks + 2, n, i, mode);
      uint32_t off = (hb % span) * mult + (lb % span);
      tgt[i] = (int32_t)(off % span);
    }
    for (int64_t i = 0; i < n; ++i) nw[i] = wealth[i];
    for (int64_t i = 0; i < n; ++i) {
      int32_t ti = tgt[i] + (tgt[i] >= i ? 1 : 0);
      if (wealth[i] > 0) {
        nw[i] -= 1;
        nw[ti] += 1;
      }
    }
    memcpy(wealth, nw, sizeof(int32_t) * (size_t)n);
    mean_out[t] = orc_mean_i32(wealth, n);
    gini_out[t] = orc_gini_i32(wealth, n);
  }
  free(tgt);
  free(nw);
}

/* ------------------------------------------------------------------------ */
/* Schelling segregation on a periodic Grid (reference Grid stores shape and
 * the periodic flag only, jaxabm/agentpy.py:465-527; the neighbourhood step is
 * the program's own).  Cells are row-major, index r*W + c.
 * Init: u = uniform(PRNGKey(seed), (H*W,)); type 1 if u < 0.45, 2 if u < 0.9,
 * else empty; agent k = k-th occupied cell in row-major order.
 * Agent phase (agent keys = split(ck, n_agents)): Moore-8 on the torus,
 * happy iff 8*same >= 3*occupied (same/occupied >= 0.375; isolated agents are
 * happy); unhappy agents draw j = randint(agent_key, (), 0, n_empty) and
 * target the j-th empty cell of the pre-step grid (row-major order).
 * Env phase: each target cell goes to its lowest-id claimant; winners move.
 * Record: fraction happy (agent phase), number of moves. */
int64_t orc_schelling_init(int32_t height, int32_t width, int64_t seed,
                           int mode, int8_t *cells, int32_t *agent_pos,
                           int8_t *agent_type, uint32_t rng[2]) {
  int64_t nc = (int64_t)height * width;
  uint32_t k[2], ik[2];
  orc_prng_key(seed, k);
  float *u = (float *)malloc(sizeof(float) * (size_t)nc);
  orc_uniform_f32(k, nc, 0.0f, 1.0f, u, mode);
  int64_t na = 0;
  for (int64_t c = 0; c < nc; ++c) {
    int8_t ty = u[c] < 0.45f ? 1 : (u[c] < 0.9f ? 2 : 0);
    cells[c] = ty;
    if (ty) {
      agent_pos[na] = (int32_t)c;
      agent_type[na] = ty;
      ++na;
    }
  }
  free(u);
  init_keys(seed, mode, rng, ik);
  return na;
}

/* Agent k's phase of a Schelling step: Moore-8 happiness on the torus and,
 * when unhappy, the claimed empty cell (-1 otherwise). */
static int schelling_agent(int32_t height, int32_t width, const int8_t *cells,
                           const int32_t *agent_pos, const int8_t *agent_type, int64_t k,
                           const uint32_t ak[2], const int32_t *empty, int64_t ne, int mode,
                           int32_t *target) {
  int32_t p = agent_pos[k];
  int32_t r = p / width, c = p % width;
  int32_t occ = 0, same = 0;
  for (int dr = -1; dr <= 1; ++dr)
    for (int dc = -1; dc <= 1; ++dc) {
      if (!dr && !dc) continue;
      int32_t rr = (r + dr + height) % height;
      int32_t cc = (c + dc + width) % width;
      int8_t v = cells[(int64_t)rr * width + cc];
      if (v) {
        occ++;
        if (v == agent_type[k]) same++;
      }
    }
  int h = 8 * same >= 3 * occ;
  target[k] = -1;
  if (!h && ne > 0) {
    int32_t j;
    orc_randint_i32(ak, 1, 0, (int32_t)ne, &j, mode);
    target[k] = empty[j];
  }
  return h;
}

/* The env phase: each claimed cell goes to its lowest-id claimant, winners
 * move; returns the moves. */
static int32_t schelling_claims(int64_t nc, int8_t *cells, int32_t *agent_pos,
                                const int8_t *agent_type, int64_t n_agents,
                                const int32_t *target, int32_t *claim) {
  for (int64_t c = 0; c < nc; ++c) claim[c] = INT32_MAX;
  for (int64_t k = 0; k < n_agents; ++k)
    if (target[k] >= 0 && (int32_t)k < claim[target[k]])
      claim[target[k]] = (int32_t)k;
  int32_t moves = 0;
  for (int64_t k = 0; k < n_agents; ++k) {
    if (target[k] >= 0 && claim[target[k]] == (int32_t)k) {
      cells[agent_pos[k]] = 0;
      cells[target[k]] = agent_type[k];
      agent_pos[k] = target[k];
      moves++;
    }
  }
  return moves;
}

void orc_schelling_run(int32_t height, int32_t width, int8_t *cells,
                       int32_t *agent_pos, const int8_t *agent_type,
                       int64_t n_agents, uint32_t rng[2], int32_t steps,
                       int mode, float *happy_frac_out, int32_t *moves_out) {
  int64_t nc = (int64_t)height * width;
  int32_t *empty = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
  int32_t *claim = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
  int32_t *target = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_agents + 1));
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(n_agents + 1));
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2];
    step_keys(rng, ck, uk, mode);
    orc_split(ck, n_agents, ak, mode);
    int64_t ne = 0;
    for (int64_t c = 0; c < nc; ++c)
      if (cells[c] == 0) empty[ne++] = (int32_t)c;
    int64_t happy = 0;
    for (int64_t k = 0; k < n_agents; ++k)
      happy += schelling_agent(height, width, cells, agent_pos, agent_type, k, ak + 2 * k,
                               empty, ne, mode, target);
    int32_t moves = schelling_claims(nc, cells, agent_pos, agent_type, n_agents, target, claim);
    volatile float hf = (float)happy / (float)n_agents;
    happy_frac_out[t] = hf;
    moves_out[t] = moves;
  }
  free(empty);
  free(claim);
  free(target);
  free(ak);
}

/* orc_schelling_run with the agent phase (neighbourhoods, happiness, agent
 * keys and draws) on `threads` OpenMP threads: the same results, since each
 * agent's phase reads only the pre-step grid and its own key; the empty list,
 * the claims and the moves stay sequential. */
void orc_schelling_run_threads(int32_t height, int32_t width, int8_t *cells,
                               int32_t *agent_pos, const int8_t *agent_type,
                               int64_t n_agents, uint32_t rng[2], int32_t steps,
                               int mode, float *happy_frac_out, int32_t *moves_out,
                               int threads) {
  int64_t nc = (int64_t)height * width;
  int32_t *empty = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
  int32_t *claim = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
  int32_t *target = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_agents + 1));
  for (int32_t t = 0; t < steps; ++t) {
    uint32_t ck[2], uk[2];
    step_keys(rng, ck, uk, mode);
    int64_t ne = 0;
    for (int64_t c = 0; c < nc; ++c)
      if (cells[c] == 0) empty[ne++] = (int32_t)c;
    int64_t happy = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : happy)
    for (int64_t k = 0; k < n_agents; ++k) {
      const uint32_t ak[2] = {split_element(ck, n_agents, 2 * k, mode),
                              split_element(ck, n_agents, 2 * k + 1, mode)};
      happy += schelling_agent(height, width, cells, agent_pos, agent_type, k, ak, empty, ne,
                               mode, target);
    }
    int32_t moves = schelling_claims(nc, cells, agent_pos, agent_type, n_agents, target, claim);
    volatile float hf = (float)happy / (float)n_agents;
    happy_frac_out[t] = hf;
    moves_out[t] = moves;
  }
  free(empty);
  free(claim);
  free(target);
}

/* ------------------------------------------------------------------------ */
/* Deffuant bounded-confidence model, differentiable (sigmoid gate).
 * Layout x[i*batch + b]: agent-major, instance-minor; all instances share the
 * seed and therefore the pairing (common random numbers).
 * Init (agent keys = split(init_key, n)): x_i = uniform(agent_key_i).
 * Step: env-phase key uk; perm = permutation(uk, n); pairs (perm[2i],
 * perm[2i+1]); d = xb - xa; g = sigmoid((eps - |d|) / 0.02);
 * delta = (mu * g) * d; xa += delta; xb -= delta. */
#define DEFF_SCALE 0.02f
#define HIST_TAU_INV 128.0

void orc_deffuant_init(int64_t n, int32_t batch, int64_t seed, int mode,
                       float *x, uint32_t rng[2]) {
  uint32_t ik[2];
  init_keys(seed, mode, rng, ik);
  uint32_t *ak = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)n);
  orc_split(ik, n, ak, mode);
  for (int64_t i = 0; i < n; ++i) {
    float v;
    orc_uniform_f32(ak + 2 * i, 1, 0.0f, 1.0f, &v, mode);
    for (int32_t b = 0; b < batch; ++b) x[i * batch + b] = v;
  }
  free(ak);
}

static inline float sigmoidf_ref(float z) {
  volatile float e = abm_expf(-z);
  volatile float d = 1.0f + e;
  return 1.0f / d;
}

static void deffuant_step_pairs(uint32_t rng[2], int64_t n, int mode,
                                int32_t *perm) {
  uint32_t ck[2], uk[2];
  step_keys(rng, ck, uk, mode);
  orc_permutation(uk, n, perm, mode);
}

void orc_deffuant_forward(int64_t n, int32_t batch, const float *eps,
                          const float *mu, float *x, uint32_t rng[2],
                          int32_t steps, int mode) {
  int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  for (int32_t t = 0; t < steps; ++t) {
    deffuant_step_pairs(rng, n, mode, perm);
    for (int64_t i = 0; i + 1 < n; i += 2) {
      int64_t a = perm[i], bb = perm[i + 1];
      for (int32_t b = 0; b < batch; ++b) {
        float xa = x[a * batch + b], xb = x[bb * batch + b];
        volatile float d = xb - xa;
        volatile float z = (eps[b] - fabsf(d)) / DEFF_SCALE;
        float g = sigmoidf_ref(z);
        volatile float mg = mu[b] * g;
        volatile float delta = mg * d;
        volatile float na = xa + delta;
        volatile float nb = xb - delta;
        x[a * batch + b] = na;
        x[bb * batch + b] = nb;
      }
    }
  }
  free(perm);
}

static inline double sigmoid_d(double z) { return 1.0 / (1.0 + exp(-z)); }

/* Per-instance statistics: [mean, variance, 32-bin soft histogram].
 * hist_k = mean_i sigmoid((x_i - k/32) * 128) - sigmoid((x_i - (k+1)/32) * 128).
 * Accumulated in float64. */
void orc_deffuant_stats(int64_t n, int32_t batch, const float *x,
                        double *stats) {
  for (int32_t b = 0; b < batch; ++b) {
    double *s = stats + (size_t)b * ORC_STATS_LEN;
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) m += x[i * batch + b];
    m /= (double)n;
    double v = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double d = x[i * batch + b] - m;
      v += d * d;
    }
    v /= (double)n;
    s[0] = m;
    s[1] = v;
    for (int k = 0; k < ORC_HIST_BINS; ++k) s[2 + k] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double xi = x[i * batch + b];
      for (int k = 0; k < ORC_HIST_BINS; ++k) {
        double lo = (double)k / ORC_HIST_BINS, hi = (double)(k + 1) / ORC_HIST_BINS;
        s[2 + k] += sigmoid_d((xi - lo) * HIST_TAU_INV) -
                    sigmoid_d((xi - hi) * HIST_TAU_INV);
      }
    }
    for (int k = 0; k < ORC_HIST_BINS; ++k) s[2 + k] /= (double)n;
  }
}

/* loss = (m - m*)^2 + (v - v*)^2 + sum_k (h_k - h*_k)^2 */
double orc_deffuant_loss(const double *stats, const double *target) {
  double l = 0.0;
  for (int k = 0; k < ORC_STATS_LEN; ++k) {
    double d = stats[k] - target[k];
    l += d * d;
  }
  return l;
}

/* Restated jax.grad of the loss through the whole run (reverse mode with the
 * full trajectory kept, as jit(grad(loss_fn)) does, jaxabm/analysis.py:
 * 494-522), evaluated in float64 from a float64 forward of the same program
 * (same pairings).  Also returns the float64 loss.  eps/mu/loss/grads are
 * per instance [batch]. */
void orc_deffuant_grad(int64_t n, int32_t batch, const double *eps,
                       const double *mu, const float *x0,
                       const uint32_t rng0[2], int32_t steps, int mode,
                       const double *target, double *loss_out,
                       double *grad_eps, double *grad_mu) {
  size_t nb = (size_t)n * (size_t)batch;
  double *traj = (double *)malloc(sizeof(double) * nb * (size_t)(steps + 1));
  int32_t *perms = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * (size_t)steps);
  uint32_t rng[2] = {rng0[0], rng0[1]};
  for (size_t i = 0; i < nb; ++i) traj[i] = x0[i];
  const double s = (double)DEFF_SCALE;
  for (int32_t t = 0; t < steps; ++t) {
    int32_t *perm = perms + (size_t)t * n;
    deffuant_step_pairs(rng, n, mode, perm);
    const double *xo = traj + (size_t)t * nb;
    double *xn = traj + (size_t)(t + 1) * nb;
    memcpy(xn, xo, sizeof(double) * nb);
    for (int64_t i = 0; i + 1 < n; i += 2) {
      int64_t a = perm[i], c = perm[i + 1];
      for (int32_t b = 0; b < batch; ++b) {
        double xa = xo[a * batch + b], xc = xo[c * batch + b];
        double d = xc - xa;
        double g = sigmoid_d((eps[b] - fabs(d)) / s);
        double delta = mu[b] * g * d;
        xn[a * batch + b] = xa + delta;
        xn[c * batch + b] = xc - delta;
      }
    }
  }
  /* final stats and dL/dx */
  const double *xf = traj + (size_t)steps * nb;
  double *lam = (double *)malloc(sizeof(double) * nb);
  for (int32_t b = 0; b < batch; ++b) {
    double m = 0.0, v = 0.0, h[ORC_HIST_BINS];
    for (int64_t i = 0; i < n; ++i) m += xf[i * batch + b];
    m /= (double)n;
    for (int64_t i = 0; i < n; ++i) {
      double d = xf[i * batch + b] - m;
      v += d * d;
    }
    v /= (double)n;
    for (int k = 0; k < ORC_HIST_BINS; ++k) h[k] = 0.0;
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < ORC_HIST_BINS; ++k) {
        double xi = xf[i * batch + b];
        h[k] += sigmoid_d((xi - (double)k / ORC_HIST_BINS) * HIST_TAU_INV) -
                sigmoid_d((xi - (double)(k + 1) / ORC_HIST_BINS) * HIST_TAU_INV);
      }
    double loss = (m - target[0]) * (m - target[0]) +
                  (v - target[1]) * (v - target[1]);
    double gh[ORC_HIST_BINS];
    for (int k = 0; k < ORC_HIST_BINS; ++k) {
      h[k] /= (double)n;
      loss += (h[k] - target[2 + k]) * (h[k] - target[2 + k]);
      gh[k] = 2.0 * (h[k] - target[2 + k]) / (double)n;
    }
    loss_out[b] = loss;
    double gm = 2.0 * (m - target[0]) / (double)n;
    double gv = 2.0 * (v - target[1]) * 2.0 / (double)n;
    for (int64_t i = 0; i < n; ++i) {
      double xi = xf[i * batch + b];
      double g = gm + gv * (xi - m);
      for (int k = 0; k < ORC_HIST_BINS; ++k) {
        double sl = sigmoid_d((xi - (double)k / ORC_HIST_BINS) * HIST_TAU_INV);
        double sh = sigmoid_d((xi - (double)(k + 1) / ORC_HIST_BINS) * HIST_TAU_INV);
        g += gh[k] * HIST_TAU_INV * (sl * (1.0 - sl) - sh * (1.0 - sh));
      }
      lam[i * batch + b] = g;
    }
    grad_eps[b] = 0.0;
    grad_mu[b] = 0.0;
  }
  /* reverse sweep */
  for (int32_t t = steps - 1; t >= 0; --t) {
    const int32_t *perm = perms + (size_t)t * n;
    const double *xo = traj + (size_t)t * nb;
    for (int64_t i = 0; i + 1 < n; i += 2) {
      int64_t a = perm[i], c = perm[i + 1];
      for (int32_t b = 0; b < batch; ++b) {
        double xa = xo[a * batch + b], xc = xo[c * batch + b];
        double d = xc - xa;
        double ad = fabs(d);
        double g = sigmoid_d((eps[b] - ad) / s);
        double gp = g * (1.0 - g) / s;
        double sgn = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
        double fd = mu[b] * (g - d * sgn * gp);
        double fe = mu[b] * d * gp;
        double fm = g * d;
        double la = lam[a * batch + b], lc = lam[c * batch + b];
        double w = la - lc;
        lam[a * batch + b] = la - w * fd;
        lam[c * batch + b] = lc + w * fd;
        grad_eps[b] += w * fe;
        grad_mu[b] += w * fm;
      }
    }
  }
  free(lam);
  free(traj);
  free(perms);
}

/* orc_deffuant_grad with the instances spread over `threads` OpenMP threads
 * (a CPU baseline on all cores).  The pairings are drawn first, in step
 * order; then each instance runs alone on its own column arrays, with the
 * same operations in the same order as its column of orc_deffuant_grad, so
 * every output has the same bits. */
static void deffuant_grad_column(int64_t n, int32_t batch, int32_t b, double eps,
                                 double mu, const float *x0, const int32_t *perms,
                                 int32_t steps, const double *target, double *loss_out,
                                 double *grad_eps, double *grad_mu) {
  double *traj = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(steps + 1));
  double *lam = (double *)malloc(sizeof(double) * (size_t)n);
  const double s = (double)DEFF_SCALE;
  for (int64_t i = 0; i < n; ++i) traj[i] = x0[i * batch + b];
  for (int32_t t = 0; t < steps; ++t) {
    const int32_t *perm = perms + (size_t)t * n;
    const double *xo = traj + (size_t)t * n;
    double *xn = traj + (size_t)(t + 1) * n;
    memcpy(xn, xo, sizeof(double) * (size_t)n);
    for (int64_t i = 0; i + 1 < n; i += 2) {
      int64_t a = perm[i], c = perm[i + 1];
      double xa = xo[a], xc = xo[c];
      double d = xc - xa;
      double g = sigmoid_d((eps - fabs(d)) / s);
      double delta = mu * g * d;
      xn[a] = xa + delta;
      xn[c] = xc - delta;
    }
  }
  const double *xf = traj + (size_t)steps * n;
  double m = 0.0, v = 0.0, h[ORC_HIST_BINS];
  for (int64_t i = 0; i < n; ++i) m += xf[i];
  m /= (double)n;
  for (int64_t i = 0; i < n; ++i) {
    double d = xf[i] - m;
    v += d * d;
  }
  v /= (double)n;
  for (int k = 0; k < ORC_HIST_BINS; ++k) h[k] = 0.0;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < ORC_HIST_BINS; ++k) {
      double xi = xf[i];
      h[k] += sigmoid_d((xi - (double)k / ORC_HIST_BINS) * HIST_TAU_INV) -
              sigmoid_d((xi - (double)(k + 1) / ORC_HIST_BINS) * HIST_TAU_INV);
    }
  double loss = (m - target[0]) * (m - target[0]) + (v - target[1]) * (v - target[1]);
  double gh[ORC_HIST_BINS];
  for (int k = 0; k < ORC_HIST_BINS; ++k) {
    h[k] /= (double)n;
    loss += (h[k] - target[2 + k]) * (h[k] - target[2 + k]);
    gh[k] = 2.0 * (h[k] - target[2 + k]) / (double)n;
  }
  loss_out[b] = loss;
  double gm = 2.0 * (m - target[0]) / (double)n;
  double gv = 2.0 * (v - target[1]) * 2.0 / (double)n;
  for (int64_t i = 0; i < n; ++i) {
    double xi = xf[i];
    double g = gm + gv * (xi - m);
    for (int k = 0; k < ORC_HIST_BINS; ++k) {
      double sl = sigmoid_d((xi - (double)k / ORC_HIST_BINS) * HIST_TAU_INV);
      double sh = sigmoid_d((xi - (double)(k + 1) / ORC_HIST_BINS) * HIST_TAU_INV);
      g += gh[k] * HIST_TAU_INV * (sl * (1.0 - sl) - sh * (1.0 - sh));
    }
    lam[i] = g;
  }
  double ge = 0.0, gmu = 0.0;
  for (int32_t t = steps - 1; t >= 0; --t) {
    const int32_t *perm = perms + (size_t)t * n;
    const double *xo = traj + (size_t)t * n;
    for (int64_t i = 0; i + 1 < n; i += 2) {
      int64_t a = perm[i], c = perm[i + 1];
      double xa = xo[a], xc = xo[c];
      double d = xc - xa;
      double ad = fabs(d);
      double g = sigmoid_d((eps - ad) / s);
      double gp = g * (1.0 - g) / s;
      double sgn = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
      double fd = mu * (g - d * sgn * gp);
      double fe = mu * d * gp;
      double fm = g * d;
      double la = lam[a], lc = lam[c];
      double w = la - lc;
      lam[a] = la - w * fd;
      lam[c] = lc + w * fd;
      ge += w * fe;
      gmu += w * fm;
    }
  }
  grad_eps[b] = ge;
  grad_mu[b] = gmu;
  free(lam);
  free(traj);
}

void orc_deffuant_grad_threads(int64_t n, int32_t batch, const double *eps,
                               const double *mu, const float *x0,
                               const uint32_t rng0[2], int32_t steps, int mode,
                               const double *target, double *loss_out,
                               double *grad_eps, double *grad_mu, int32_t threads) {
  int32_t *perms = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * (size_t)steps);
  uint32_t rng[2] = {rng0[0], rng0[1]};
  for (int32_t t = 0; t < steps; ++t) deffuant_step_pairs(rng, n, mode, perms + (size_t)t * n);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int32_t b = 0; b < batch; ++b)
    deffuant_grad_column(n, batch, b, eps[b], mu[b], x0, perms, steps, target, loss_out,
                         grad_eps, grad_mu);
  free(perms);
}
