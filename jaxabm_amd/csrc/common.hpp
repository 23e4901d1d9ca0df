// common.hpp -- shared host helpers of the agent-step engine (errors, launch
// checks, host-side thresholds) on top of the device arithmetic in
// abm_device.hpp (threefry2x32 and JAX's draw constructions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "abm_device.hpp"

namespace abm {

// Integer threshold for `uniform < p` on [0,1): with m = bits >> 9,
// u = m * 2^-23 exactly, so u < p  <=>  m < ceil(p * 2^23) (p in (0, 1]).
// Computed on the host in double: exact for float p.
inline uint32_t unit_threshold(float p) {
  if (!(p > 0.0f)) return 0u;
  if (p >= 1.0f) return 1u << 23;
  double t = (double)p * 8388608.0;
  double c = __builtin_ceil(t);
  return (uint32_t)c;
}

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

}  // namespace abm

#define ABM_CHECK_HIP(expr)                                     \
  do {                                                          \
    hipError_t _e = (expr);                                     \
    if (_e != hipSuccess) return ::abm::hip_fail(_e, #expr);    \
  } while (0)

#define ABM_REQUIRE(cond, msg)                        \
  do {                                                \
    if (!(cond)) {                                    \
      ::abm::set_error(msg);                          \
      return ::abm::kErrInvalid;                      \
    }                                                 \
  } while (0)

#define ABM_LAUNCH_CHECK(what)                                   \
  do {                                                           \
    hipError_t _e = hipGetLastError();                           \
    if (_e != hipSuccess) return ::abm::hip_fail(_e, what);      \
  } while (0)
