// csr.hip -- generic gather / scatter over a Network in CSR form.
//
// The reference answers neighbour queries with O(E) masks over its edge list
// (jaxabm/agentpy.py:584-615) and models aggregate over neighbours with
// jnp gathers and segment sums; here one launch reduces every row:
//   out[v] = sum over e in row v of x[col[e]]       (float32, float64 accumulation)
//   out[v] = number of e in row v with mask[col[e]] (int32, exact)
// Rows of up to kLightDeg entries are reduced by one thread; longer rows are
// queued in LDS and reduced by a whole 64-wide wavefront (segmented reduction
// with cross-lane shuffles), so BA hubs do not serialise a workgroup.
// Scatter-add: out[idx[i]] += val[i] with float / int32 atomics.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kCsrThreads = 256;
constexpr int kLightDeg = 32;
constexpr int kHeavyCap = 256;

template <typename Acc>
__device__ __forceinline__ Acc wave_sum(Acc v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// OP 0: sum of x (float) into float; OP 1: count of mask (uint8) into int32.
template <int OP>
__global__ void __launch_bounds__(kCsrThreads)
    k_csr_reduce(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                 const void *__restrict__ in, void *__restrict__ out) {
  __shared__ int32_t heavy[kHeavyCap];
  __shared__ int32_t nheavy;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float *x = (const float *)in;
  const uint8_t *mask = (const uint8_t *)in;
  if (tid == 0) nheavy = 0;
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * kCsrThreads + tid;
  if (v < n) {
    const int32_t e0 = row_ptr[v], e1 = row_ptr[v + 1];
    bool done = true;
    if (e1 - e0 > kLightDeg) {
      const int32_t slot = atomicAdd(&nheavy, 1);
      if (slot < kHeavyCap) {
        heavy[slot] = (int32_t)v;
        done = false;
      }
    }
    if (done) {
      if (OP == 0) {
        double s = 0.0;
        for (int32_t e = e0; e < e1; ++e) s += (double)x[col[e]];
        ((float *)out)[v] = (float)s;
      } else {
        int32_t c = 0;
        for (int32_t e = e0; e < e1; ++e) c += mask[col[e]] != 0;
        ((int32_t *)out)[v] = c;
      }
    }
  }
  __syncthreads();
  const int32_t nh = min(nheavy, kHeavyCap);
  for (int32_t h = wave; h < nh; h += kCsrThreads / 64) {
    const int32_t r = heavy[h];
    const int32_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
    if (OP == 0) {
      double s = 0.0;
      for (int32_t e = e0 + lane; e < e1; e += 64) s += (double)x[col[e]];
      s = wave_sum(s);
      if (lane == 0) ((float *)out)[r] = (float)s;
    } else {
      int32_t c = 0;
      for (int32_t e = e0 + lane; e < e1; e += 64) c += mask[col[e]] != 0;
      c = wave_sum(c);
      if (lane == 0) ((int32_t *)out)[r] = c;
    }
  }
}

template <typename T>
__global__ void k_scatter_add(int64_t m, const int32_t *__restrict__ idx,
                              const T *__restrict__ val, T *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) atomicAdd(&out[idx[i]], val[i]);
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_csr_gather_sum_f32(int64_t n, const int32_t *row_ptr, const int32_t *col, const float *x,
                           float *out, void *stream) {
  ABM_REQUIRE(n >= 0 && (n == 0 || (row_ptr && col && x && out)),
              "abm_csr_gather_sum_f32: bad arguments");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_csr_reduce<0>, dim3((unsigned)((n + kCsrThreads - 1) / kCsrThreads)),
                     dim3(kCsrThreads), 0, (hipStream_t)stream, n, row_ptr, col,
                     (const void *)x, (void *)out);
  ABM_LAUNCH_CHECK("abm_csr_gather_sum_f32");
  return kOk;
}

int abm_csr_count_u8(int64_t n, const int32_t *row_ptr, const int32_t *col, const uint8_t *mask,
                     int32_t *out, void *stream) {
  ABM_REQUIRE(n >= 0 && (n == 0 || (row_ptr && col && mask && out)),
              "abm_csr_count_u8: bad arguments");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_csr_reduce<1>, dim3((unsigned)((n + kCsrThreads - 1) / kCsrThreads)),
                     dim3(kCsrThreads), 0, (hipStream_t)stream, n, row_ptr, col,
                     (const void *)mask, (void *)out);
  ABM_LAUNCH_CHECK("abm_csr_count_u8");
  return kOk;
}

int abm_scatter_add_f32(int64_t m, const int32_t *idx, const float *val, float *out,
                        void *stream) {
  ABM_REQUIRE(m >= 0 && (m == 0 || (idx && val && out)), "abm_scatter_add_f32: bad arguments");
  if (m == 0) return kOk;
  hipLaunchKernelGGL(k_scatter_add<float>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, m, idx, val, out);
  ABM_LAUNCH_CHECK("abm_scatter_add_f32");
  return kOk;
}

int abm_scatter_add_i32(int64_t m, const int32_t *idx, const int32_t *val, int32_t *out,
                        void *stream) {
  ABM_REQUIRE(m >= 0 && (m == 0 || (idx && val && out)), "abm_scatter_add_i32: bad arguments");
  if (m == 0) return kOk;
  hipLaunchKernelGGL(k_scatter_add<int32_t>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, m, idx, val, out);
  ABM_LAUNCH_CHECK("abm_scatter_add_i32");
  return kOk;
}

}  // extern "C"
