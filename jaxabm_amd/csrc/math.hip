// math.hip -- the float32 / float64 exp, log, log1p and pow of abm_math.h,
// elementwise on the device.  They are the functions the draw and model
// kernels call (erf_inv inside jax.random.normal, the gamma / beta sampler,
// the Deffuant sigmoid gate), exposed so a caller can check them against any
// other evaluation of abm_math.h bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

__global__ void k_math_f32(int fn, const float *__restrict__ x, const float *__restrict__ y,
                           float *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  float r;
  switch (fn) {
    case ABM_MATH_EXP: r = abm_expf(v); break;
    case ABM_MATH_LOG: r = abm_logf(v); break;
    case ABM_MATH_LOG1P: r = abm_log1pf(v); break;
    default: r = abm_powf(v, y[i]); break;
  }
  out[i] = r;
}

__global__ void k_math_f64(int fn, const double *__restrict__ x, double *__restrict__ out,
                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fn == ABM_MATH_EXP ? abm_exp(x[i]) : abm_log(x[i]);
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_math_f32(int fn, const float *x, const float *y, float *out, int64_t n, void *stream) {
  ABM_REQUIRE(fn >= ABM_MATH_EXP && fn <= ABM_MATH_POW, "abm_math_f32: unknown function");
  ABM_REQUIRE(n >= 0 && n <= ((int64_t)1 << 38), "abm_math_f32: n out of range");
  if (n == 0) return kOk;
  ABM_REQUIRE(x && out && (fn != ABM_MATH_POW || y), "abm_math_f32: null buffer");
  const int threads = 256;
  k_math_f32<<<(unsigned)((n + threads - 1) / threads), threads, 0, (hipStream_t)stream>>>(
      fn, x, y, out, n);
  ABM_LAUNCH_CHECK("abm_math_f32");
  return kOk;
}

int abm_math_f64(int fn, const double *x, double *out, int64_t n, void *stream) {
  ABM_REQUIRE(fn == ABM_MATH_EXP || fn == ABM_MATH_LOG, "abm_math_f64: unknown function");
  ABM_REQUIRE(n >= 0 && n <= ((int64_t)1 << 38), "abm_math_f64: n out of range");
  if (n == 0) return kOk;
  ABM_REQUIRE(x && out, "abm_math_f64: null buffer");
  const int threads = 256;
  k_math_f64<<<(unsigned)((n + threads - 1) / threads), threads, 0, (hipStream_t)stream>>>(
      fn, x, out, n);
  ABM_LAUNCH_CHECK("abm_math_f64");
  return kOk;
}

}  // extern "C"
