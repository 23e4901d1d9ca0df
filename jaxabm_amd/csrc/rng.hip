// rng.hip -- JAX-stream threefry PRNG on the device, behind the C-ABI.
//
// Replaces the jax.random primitives the reference calls per step:
// PRNGKey / split / uniform / randint / normal / bernoulli / permutation
// (jaxabm/model.py:46,129,156,164,183; jaxabm/agent.py:115,156;
// jaxabm/agentpy.py:509-513).  Keys live in device memory as uint32[2] so the
// calls can sit inside captured step graphs.  `*_each` variants are the vmap
// of a scalar draw over a [n,2] key array (what AgentCollection.update does
// with split(key, N), jaxabm/agent.py:156-177).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return kErrHip;
}

__device__ __forceinline__ Key load_key(const uint32_t *k) { return Key{k[0], k[1]}; }

// ---------------------------------------------------------------- kernels
__global__ void k_threefry(const uint32_t *__restrict__ key,
                           const uint32_t *__restrict__ x0,
                           const uint32_t *__restrict__ x1,
                           uint32_t *__restrict__ y0, uint32_t *__restrict__ y1,
                           int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t a, b;
  threefry(key[0], key[1], x0[i], x1[i], a, b);
  y0[i] = a;
  y1[i] = b;
}

__global__ void k_split(const uint32_t *__restrict__ key, int64_t num,
                        uint32_t *__restrict__ out, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  Key k = load_key(key);
  Key r = split_at(k, num, i, mode);
  out[2 * i] = r.k0;
  out[2 * i + 1] = r.k1;
}

// Each thread produces the elements one threefry block yields: in original
// mode lane i gives elements i and h+i, in partitionable mode one element.
enum Dist { kBits = 0, kUniform = 1, kNormal = 2, kBernoulli = 3 };

template <int D, typename T>
__device__ __forceinline__ T finish(uint32_t b, float lo, float span, float p) {
  if constexpr (D == kBits) {
    return (T)b;
  } else if constexpr (D == kUniform) {
    return (T)scale_uniform(unit_float(b), lo, span);
  } else if constexpr (D == kNormal) {
    return (T)normal_from_bits(b);
  } else {
    return (T)(unit_float(b) < p ? 1 : 0);
  }
}

template <int D, typename T>
__global__ void k_draw(const uint32_t *__restrict__ key, int64_t n,
                       T *__restrict__ out, float lo, float span, float p,
                       int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Key k = load_key(key);
  if (mode == kModePartitionable) {
    if (i >= n) return;
    uint32_t y0, y1;
    threefry(k.k0, k.k1, (uint32_t)((uint64_t)i >> 32), (uint32_t)i, y0, y1);
    out[i] = finish<D, T>(y0 ^ y1, lo, span, p);
    return;
  }
  const int64_t h = (n + 1) >> 1;
  if (i >= h) return;
  const int64_t j = h + i;
  uint32_t y0, y1;
  threefry(k.k0, k.k1, (uint32_t)i, j < n ? (uint32_t)j : 0u, y0, y1);
  out[i] = finish<D, T>(y0, lo, span, p);
  if (j < n) out[j] = finish<D, T>(y1, lo, span, p);
}

__global__ void k_randint(const uint32_t *__restrict__ key, int64_t n,
                          int32_t *__restrict__ out, int32_t lo, uint32_t span,
                          int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Key k1, k2;
  split2(load_key(key), k1, k2, mode);
  uint32_t hb = bits_at(k1, n, i, mode);
  uint32_t lb = bits_at(k2, n, i, mode);
  out[i] = (int32_t)((uint32_t)lo + randint_offset(hb, lb, span));
}

// vmapped scalar draws over keys[n][2]
__global__ void k_split_each(const uint32_t *__restrict__ keys, int64_t n,
                             int64_t num, uint32_t *__restrict__ out,
                             int mode) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * num) return;
  int64_t i = t / num, j = t - i * num;
  Key r = split_at(Key{keys[2 * i], keys[2 * i + 1]}, num, j, mode);
  out[2 * t] = r.k0;
  out[2 * t + 1] = r.k1;
}

template <int D, typename T>
__global__ void k_draw_each(const uint32_t *__restrict__ keys, int64_t n,
                            int64_t per, T *__restrict__ out, float lo,
                            float span, float p, int mode) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * per) return;
  int64_t i = t / per, j = t - i * per;
  uint32_t b = bits_at(Key{keys[2 * i], keys[2 * i + 1]}, per, j, mode);
  out[t] = finish<D, T>(b, lo, span, p);
}

__global__ void k_randint_each(const uint32_t *__restrict__ keys, int64_t n,
                               int64_t per, int32_t *__restrict__ out,
                               int32_t lo, uint32_t span, int mode) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * per) return;
  int64_t i = t / per, j = t - i * per;
  Key k1, k2;
  split2(Key{keys[2 * i], keys[2 * i + 1]}, k1, k2, mode);
  uint32_t hb = bits_at(k1, per, j, mode);
  uint32_t lb = bits_at(k2, per, j, mode);
  out[t] = (int32_t)((uint32_t)lo + randint_offset(hb, lb, span));
}

// gamma / loggamma / beta elements: out[t], t = i * per + j, from keys[i] and
// the element's parameters a[t] (and b[t] for beta).  Rejection loops end per
// lane (gamma_one's cap); 64-thread blocks keep a slow lane's wave short.
__global__ void __launch_bounds__(64)
    k_gamma_each(const uint32_t *__restrict__ keys, int64_t n, int64_t per,
                 const float *__restrict__ a, const float *__restrict__ b,
                 int kind, float *__restrict__ out, int mode) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * per) return;
  const int64_t i = t / per, j = t - i * per;
  const Key k{keys[2 * i], keys[2 * i + 1]};
  out[t] = kind == ABM_BETA ? beta_draw(k, per, j, a[t], b[t], mode)
                            : gamma_draw(k, per, j, a[t], kind == ABM_LOGGAMMA, mode);
}

static inline unsigned grid_for(int64_t n, int bs) {
  return (unsigned)((n + bs - 1) / bs);
}

template <int D, typename T>
static int launch_draw(const char *name, const uint32_t *key, int64_t n,
                       T *out, float lo, float span, float p, int mode,
                       void *stream) {
  if (n < 0 || (mode != 0 && mode != 1)) {
    set_error(std::string(name) + ": bad n or mode");
    return kErrInvalid;
  }
  if (n == 0) return kOk;
  int64_t lanes = mode == kModePartitionable ? n : (n + 1) / 2;
  hipLaunchKernelGGL((k_draw<D, T>), dim3(grid_for(lanes, 256)), dim3(256), 0,
                     (hipStream_t)stream, key, n, out, lo, span, p, mode);
  ABM_LAUNCH_CHECK(name);
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

const char *abm_last_error(void) { return g_last_error.c_str(); }

int abm_version(void) { return JAXABM_AMD_ABI_VERSION; }

int abm_threefry2x32(const uint32_t *key, const uint32_t *x0,
                     const uint32_t *x1, uint32_t *y0, uint32_t *y1,
                     int64_t n, void *stream) {
  ABM_REQUIRE(n >= 0, "abm_threefry2x32: n must be >= 0");
  if (n == 0) return kOk;
  ABM_REQUIRE(key && x0 && x1 && y0 && y1, "abm_threefry2x32: null pointer");
  hipLaunchKernelGGL(k_threefry, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, key, x0, x1, y0, y1, n);
  ABM_LAUNCH_CHECK("abm_threefry2x32");
  return kOk;
}

int abm_split(const uint32_t *key, int64_t num, uint32_t *out, int mode,
              void *stream) {
  ABM_REQUIRE(num >= 0, "abm_split: num must be >= 0");
  ABM_REQUIRE(mode == 0 || mode == 1, "abm_split: mode must be 0 or 1");
  if (num == 0) return kOk;
  hipLaunchKernelGGL(k_split, dim3(grid_for(num, 256)), dim3(256), 0,
                     (hipStream_t)stream, key, num, out, mode);
  ABM_LAUNCH_CHECK("abm_split");
  return kOk;
}

int abm_random_bits32(const uint32_t *key, int64_t n, uint32_t *out, int mode,
                      void *stream) {
  return launch_draw<kBits, uint32_t>("abm_random_bits32", key, n, out, 0.f,
                                      0.f, 0.f, mode, stream);
}

int abm_uniform_f32(const uint32_t *key, int64_t n, float lo, float hi,
                    float *out, int mode, void *stream) {
  volatile float span = hi - lo;
  return launch_draw<kUniform, float>("abm_uniform_f32", key, n, out, lo, span,
                                      0.f, mode, stream);
}

int abm_normal_f32(const uint32_t *key, int64_t n, float *out, int mode,
                   void *stream) {
  return launch_draw<kNormal, float>("abm_normal_f32", key, n, out, 0.f, 0.f,
                                     0.f, mode, stream);
}

int abm_bernoulli_u8(const uint32_t *key, int64_t n, float p, uint8_t *out,
                     int mode, void *stream) {
  return launch_draw<kBernoulli, uint8_t>("abm_bernoulli_u8", key, n, out, 0.f,
                                          0.f, p, mode, stream);
}

int abm_randint_i32(const uint32_t *key, int64_t n, int32_t lo, int32_t hi,
                    int32_t *out, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && (mode == 0 || mode == 1), "abm_randint_i32: bad n or mode");
  if (n == 0) return kOk;
  uint32_t span = hi <= lo ? 1u : (uint32_t)((int64_t)hi - (int64_t)lo);
  hipLaunchKernelGGL(k_randint, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, key, n, out, lo, span, mode);
  ABM_LAUNCH_CHECK("abm_randint_i32");
  return kOk;
}

int abm_split_each(const uint32_t *keys, int64_t n, int64_t num, uint32_t *out,
                   int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && num >= 0 && (mode == 0 || mode == 1),
              "abm_split_each: bad n, num or mode");
  if (n == 0 || num == 0) return kOk;
  hipLaunchKernelGGL(k_split_each, dim3(grid_for(n * num, 256)), dim3(256), 0,
                     (hipStream_t)stream, keys, n, num, out, mode);
  ABM_LAUNCH_CHECK("abm_split_each");
  return kOk;
}

int abm_draw_each(const uint32_t *keys, int64_t n, int64_t per, int dist,
                  float lo, float hi, void *out, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && per >= 0 && (mode == 0 || mode == 1),
              "abm_draw_each: bad n, per or mode");
  if (n == 0 || per == 0) return kOk;
  unsigned g = grid_for(n * per, 256);
  hipStream_t s = (hipStream_t)stream;
  volatile float span = hi - lo;
  switch (dist) {
    case kBits:
      hipLaunchKernelGGL((k_draw_each<kBits, uint32_t>), dim3(g), dim3(256), 0,
                         s, keys, n, per, (uint32_t *)out, 0.f, 0.f, 0.f, mode);
      break;
    case kUniform:
      hipLaunchKernelGGL((k_draw_each<kUniform, float>), dim3(g), dim3(256), 0,
                         s, keys, n, per, (float *)out, lo, span, 0.f, mode);
      break;
    case kNormal:
      hipLaunchKernelGGL((k_draw_each<kNormal, float>), dim3(g), dim3(256), 0,
                         s, keys, n, per, (float *)out, 0.f, 0.f, 0.f, mode);
      break;
    case kBernoulli:
      hipLaunchKernelGGL((k_draw_each<kBernoulli, uint8_t>), dim3(g), dim3(256),
                         0, s, keys, n, per, (uint8_t *)out, 0.f, 0.f, lo, mode);
      break;
    default:
      set_error("abm_draw_each: dist must be 0..3");
      return kErrInvalid;
  }
  ABM_LAUNCH_CHECK("abm_draw_each");
  return kOk;
}

int abm_randint_each(const uint32_t *keys, int64_t n, int64_t per, int32_t lo,
                     int32_t hi, int32_t *out, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && per >= 0 && (mode == 0 || mode == 1),
              "abm_randint_each: bad n, per or mode");
  if (n == 0 || per == 0) return kOk;
  uint32_t span = hi <= lo ? 1u : (uint32_t)((int64_t)hi - (int64_t)lo);
  hipLaunchKernelGGL(k_randint_each, dim3(grid_for(n * per, 256)), dim3(256),
                     0, (hipStream_t)stream, keys, n, per, out, lo, span, mode);
  ABM_LAUNCH_CHECK("abm_randint_each");
  return kOk;
}

int abm_gamma_each(const uint32_t *keys, int64_t n, int64_t per, const float *a,
                   const float *b, int kind, float *out, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && per >= 0 && (mode == 0 || mode == 1) &&
                  (kind == ABM_GAMMA || kind == ABM_LOGGAMMA || kind == ABM_BETA),
              "abm_gamma_each: bad n, per, kind or mode");
  if (n == 0 || per == 0) return kOk;
  ABM_REQUIRE(keys && a && out && (kind != ABM_BETA || b), "abm_gamma_each: null buffer");
  hipLaunchKernelGGL(k_gamma_each, dim3(grid_for(n * per, 64)), dim3(64), 0,
                     (hipStream_t)stream, keys, n, per, a, b, kind, out, mode);
  ABM_LAUNCH_CHECK("abm_gamma_each");
  return kOk;
}

// Host-side key arithmetic (no device involved): the step-key chains every
// program precomputes before its device-resident loop.
int abm_host_split(const uint32_t key[2], int64_t num, uint32_t *out,
                   int mode) {
  ABM_REQUIRE(key && out && num >= 0 && (mode == 0 || mode == 1),
              "abm_host_split: bad arguments");
  Key k{key[0], key[1]};
  for (int64_t i = 0; i < num; ++i) {
    Key r = split_at(k, num, i, mode);
    out[2 * i] = r.k0;
    out[2 * i + 1] = r.k1;
  }
  return kOk;
}

int abm_step_keys(uint32_t rng[2], int32_t steps, int32_t n_collections,
                  int32_t has_env_phase, int32_t split_env, int mode,
                  uint32_t *out) {
  ABM_REQUIRE(rng && out && steps >= 0 && n_collections >= 0 &&
                  (mode == 0 || mode == 1),
              "abm_step_keys: bad arguments");
  // Per step (jaxabm/model.py:156,164,183): rng, sk = split(rng);
  // per collection sk, ck = split(sk); then sk, uk = split(sk) when there is
  // an env phase.  Output row: [ck_0 .. ck_{C-1}, uk] (uk optionally split
  // once more into two keys, the SIR program's infection/recovery keys).
  const int32_t row = 2 * n_collections + (has_env_phase ? (split_env ? 4 : 2) : 0);
  Key r{rng[0], rng[1]};
  for (int32_t t = 0; t < steps; ++t) {
    Key nr, sk;
    split2(r, nr, sk, mode);
    r = nr;
    uint32_t *o = out + (size_t)t * row;
    for (int32_t c = 0; c < n_collections; ++c) {
      Key s2, ck;
      split2(sk, s2, ck, mode);
      sk = s2;
      o[2 * c] = ck.k0;
      o[2 * c + 1] = ck.k1;
    }
    if (has_env_phase) {
      Key s2, uk;
      split2(sk, s2, uk, mode);
      uint32_t *q = o + 2 * n_collections;
      if (split_env) {
        Key a, b;
        split2(uk, a, b, mode);
        q[0] = a.k0;
        q[1] = a.k1;
        q[2] = b.k0;
        q[3] = b.k1;
      } else {
        q[0] = uk.k0;
        q[1] = uk.k1;
      }
    }
  }
  rng[0] = r.k0;
  rng[1] = r.k1;
  return kOk;
}

}  // extern "C"
