// sir.hip -- network SIR step: synchronous S->I->R over an undirected CSR.
//
// Program semantics (reference schedule jaxabm/model.py:146-216, Network
// jaxabm/agentpy.py:530-615 rebuilt as CSR):
//   k_inf, k_rec = split(update_key)
//   u = uniform(k_inf, (nnz,)), r = uniform(k_rec, (n,))
//   S -> I  iff some CSR entry e of row v has old[col[e]] == I and u[e] < beta
//   I -> R  iff r[v] < gamma
// Counter-based draws let the kernel evaluate u[e] only for S-I entries and
// stop at the first infecting entry: the result is the same "any" as the dense
// formulation.  `u < p` is tested on the integer mantissa (m < ceil(p*2^23)),
// which is exact for the [0,1) uniform.
//
// Layout: state uint8[n] ping-pong (S=0, I=1, R=2), row_ptr int32[n+1],
// col int32[nnz] (rows ascending), heavy int32[] = rows with degree > 64,
// handled one 64-wide wavefront per row so hub rows do not serialise a lane.
// Counts (S, I, R) per step are reduced with wave ballots, block LDS and one
// atomic per block into counts[t*3 .. t*3+2].
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kSirThreads = 256;
constexpr int kHeavyDeg = 64;

__device__ __forceinline__ void count_classes(int c, int32_t *cnt) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint64_t m = __ballot(c == k);
    if (lane == 0 && m) atomicAdd(&cnt[k], (int32_t)__popcll(m));
  }
}

__global__ void __launch_bounds__(kSirThreads)
    k_sir_step(int64_t n, int64_t nnz, const int32_t *__restrict__ row_ptr,
               const int32_t *__restrict__ col,
               const int32_t *__restrict__ heavy, int32_t n_heavy,
               int32_t heavy_blocks, const uint8_t *__restrict__ in,
               uint8_t *__restrict__ out, const uint32_t *__restrict__ keys,
               uint32_t thr_beta, uint32_t thr_gamma, int mode,
               int32_t *__restrict__ counts) {
  __shared__ int32_t cnt[3];
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
  __syncthreads();
  const Key kinf{keys[0], keys[1]};
  const Key krec{keys[2], keys[3]};
  const int lane = threadIdx.x & 63;
  int c = -1;
  if ((int32_t)blockIdx.x < heavy_blocks) {
    // one wavefront per heavy row; only S rows need the neighbour scan
    const int32_t idx = blockIdx.x * (kSirThreads / 64) + (threadIdx.x >> 6);
    if (idx < n_heavy) {
      const int32_t v = heavy[idx];
      if (in[v] == 0) {
        const int64_t lo = row_ptr[v], hi = row_ptr[v + 1];
        bool infected = false;
        for (int64_t b = lo; b < hi; b += 64) {
          const int64_t e = b + lane;
          bool hit = false;
          if (e < hi) {
            const int32_t u = col[e];
            if (in[u] == 1) hit = (bits_at(kinf, nnz, e, mode) >> 9) < thr_beta;
          }
          if (__ballot(hit)) {
            infected = true;
            break;
          }
        }
        if (lane == 0) {
          out[v] = infected ? 1 : 0;
          c = infected ? 1 : 0;
        }
      }
    }
  } else {
    const int64_t v =
        (int64_t)(blockIdx.x - heavy_blocks) * kSirThreads + threadIdx.x;
    if (v < n) {
      uint8_t s = in[v];
      if (s == 0) {
        const int64_t lo = row_ptr[v], hi = row_ptr[v + 1];
        if (hi - lo <= kHeavyDeg) {
          for (int64_t e = lo; e < hi; ++e) {
            if (in[col[e]] == 1 &&
                (bits_at(kinf, nnz, e, mode) >> 9) < thr_beta) {
              s = 1;
              break;
            }
          }
          out[v] = s;
          c = s;
        }
      } else {
        if (s == 1 && (bits_at(krec, n, v, mode) >> 9) < thr_gamma) s = 2;
        out[v] = s;
        c = s;
      }
    }
  }
  count_classes(c, cnt);
  __syncthreads();
  if (threadIdx.x < 3 && cnt[threadIdx.x])
    atomicAdd(&counts[threadIdx.x], cnt[threadIdx.x]);
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_sir_seed(const int32_t *__restrict__ perm, int64_t n0,
                           uint8_t *__restrict__ state) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) state[perm[i]] = 1;
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream) {
  ABM_REQUIRE(n >= 0, "abm_fill_i32: n must be >= 0");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, (hipStream_t)stream, p, n, v);
  ABM_LAUNCH_CHECK("abm_fill_i32");
  return kOk;
}

uint32_t abm_unit_threshold(float p) { return unit_threshold(p); }

int abm_sir_seed(int64_t n, const int32_t *perm, int64_t n0, uint8_t *state,
                 void *stream) {
  ABM_REQUIRE(n > 0 && n0 >= 0 && n0 <= n, "abm_sir_seed: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, s, state, n, (uint8_t)0);
  if (n0 > 0)
    hipLaunchKernelGGL(k_sir_seed, dim3((unsigned)((n0 + 255) / 256)),
                       dim3(256), 0, s, perm, n0, state);
  ABM_LAUNCH_CHECK("abm_sir_seed");
  return kOk;
}

int abm_sir_run(const abm_sir_graph *g, uint8_t *state, uint8_t *scratch,
                const uint32_t *keys, int32_t steps, float beta, float gamma,
                int mode, int32_t *counts, void *stream) {
  ABM_REQUIRE(g && state && scratch && keys && counts,
              "abm_sir_run: null argument");
  ABM_REQUIRE(g->n > 0 && g->n < (1ll << 31) && g->nnz >= 0,
              "abm_sir_run: bad graph sizes");
  ABM_REQUIRE(g->n_heavy >= 0 && g->n_heavy <= g->n,
              "abm_sir_run: bad heavy-row count");
  ABM_REQUIRE(steps >= 0 && (mode == 0 || mode == 1),
              "abm_sir_run: bad steps or mode");
  if (steps == 0) return kOk;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t tb = unit_threshold(beta), tg = unit_threshold(gamma);
  const int32_t heavy_blocks =
      (int32_t)((g->n_heavy + (kSirThreads / 64) - 1) / (kSirThreads / 64));
  const int64_t light_blocks = (g->n + kSirThreads - 1) / kSirThreads;
  const int64_t blocks = heavy_blocks + light_blocks;
  ABM_REQUIRE(blocks < (1ll << 31), "abm_sir_run: grid too large");
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((3 * (int64_t)steps + 255) / 256)),
                     dim3(256), 0, s, counts, 3 * (int64_t)steps, 0);
  uint8_t *buf[2] = {state, scratch};
  for (int32_t t = 0; t < steps; ++t) {
    hipLaunchKernelGGL(k_sir_step, dim3((unsigned)blocks), dim3(kSirThreads), 0,
                       s, g->n, g->nnz, g->row_ptr, g->col, g->heavy,
                       (int32_t)g->n_heavy, heavy_blocks, buf[t & 1],
                       buf[(t + 1) & 1], keys + 4 * (int64_t)t, tb, tg, mode,
                       counts + 3 * (int64_t)t);
  }
  if (steps & 1) ABM_CHECK_HIP(hipMemcpyAsync(state, scratch, (size_t)g->n,
                                              hipMemcpyDeviceToDevice, s));
  ABM_LAUNCH_CHECK("abm_sir_run");
  return kOk;
}

int abm_sir_step(const abm_sir_graph *g, const uint8_t *in, uint8_t *out,
                 const uint32_t *keys4, float beta, float gamma, int mode,
                 int32_t *counts3, void *stream) {
  ABM_REQUIRE(g && in && out && keys4 && counts3 && in != out,
              "abm_sir_step: bad argument");
  ABM_REQUIRE(g->n > 0 && g->n < (1ll << 31) && g->nnz >= 0 &&
                  g->n_heavy >= 0 && g->n_heavy <= g->n &&
                  (mode == 0 || mode == 1),
              "abm_sir_step: bad graph sizes or mode");
  const int32_t heavy_blocks =
      (int32_t)((g->n_heavy + (kSirThreads / 64) - 1) / (kSirThreads / 64));
  const int64_t blocks = heavy_blocks + (g->n + kSirThreads - 1) / kSirThreads;
  hipLaunchKernelGGL(k_sir_step, dim3((unsigned)blocks), dim3(kSirThreads), 0,
                     (hipStream_t)stream, g->n, g->nnz, g->row_ptr, g->col,
                     g->heavy, (int32_t)g->n_heavy, heavy_blocks, in, out,
                     keys4, unit_threshold(beta), unit_threshold(gamma), mode,
                     counts3);
  ABM_LAUNCH_CHECK("abm_sir_step");
  return kOk;
}

int abm_sir_heavy_threshold(void) { return kHeavyDeg; }

}  // extern "C"
