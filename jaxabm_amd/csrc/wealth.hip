// wealth.hip -- wealth-exchange random walk: one workgroup runs a whole
// simulation of one instance; a launch runs a batch of instances.
//
// Program (reference API, jaxabm/model.py:118-216, jaxabm/agent.py:92-177):
//  init   agent keys = split(init_key, n); kx, ky = split(agent_key);
//         x = uniform(kx), y = uniform(ky), wealth = 1
//  step   agent phase (agent keys = split(ck, n)): k1, k2 = split(agent_key);
//           x = mod(x + 0.01 * normal(k1), 1), y = mod(y + 0.01 * normal(k2), 1)
//         env phase (uk): t = randint(uk, (n,), 0, n-1); t += (t >= i);
//           give = wealth > 0; wealth += bincount(t, give) - give
//  record mean(wealth) and Gini (examples/models/advanced_economic_model.py:
//         1798-1811): x = sort(wealth); S = sum(i * x_i) and T = sum(x) in
//         int32; G = f32(2 S) / f32(n T) - f32((n+1)/n) if T > 1e-6.
//
// The sort is replaced by a histogram over wealth values: with c_w agents of
// wealth w and b_w agents below, their ranks sum to c_w * b_w + c_w(c_w+1)/2,
// so S = sum_w w * ranksum_w (mod 2^32, the int32 wrap of the reference).
// Whole run in one launch: every phase boundary is a workgroup barrier.
// Scatter-adds are integer atomics (exact, order-free); the histogram and
// increments are cleared by the kernel after use.  An instance of up to
// kWealthLdsAgents agents runs out of LDS (x, y, wealth, increments and
// histogram, 20 bytes per agent, loaded at the start and x, y, wealth written
// back at the end), so a step costs LDS latencies instead of global memory
// round trips; larger instances run on their global buffers.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kWealthThreads = 1024;
// LDS per resident agent: x, y, wealth, increment, histogram bin
constexpr int64_t kWealthLdsBytesPerAgent = 20;
constexpr int64_t kWealthLdsAgents = 8000;  // 160,020 bytes with the extra bin

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  uint32_t s = 0;
  if (threadIdx.x < 64) {
    s = threadIdx.x < (blockDim.x >> 6) ? red[threadIdx.x] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  }
  __syncthreads();
  return s;  // valid in wave 0
}

// The whole run of one instance on the arrays x, y, w (state), inc and hist
// (zero on entry, left zero) -- global buffers or their LDS copies.
// The draws of agent i at one step: the position increments 0.01 * normal
// (rounded as the reference's product) and the exchange target.
struct WealthDraw {
  float mdx, mdy;
  uint32_t target;
};

__device__ __forceinline__ WealthDraw wealth_draw(const uint32_t *keys, int32_t t, int64_t n,
                                                  int64_t i, int mode) {
  const Key ck{keys[4 * t + 0], keys[4 * t + 1]};
  const Key uk{keys[4 * t + 2], keys[4 * t + 3]};
  Key kk1, kk2, k1, k2;
  split2(uk, kk1, kk2, mode);
  split2(split_at(ck, n, i, mode), k1, k2, mode);
  WealthDraw d;
  d.mdx = 0.01f * normal_from_bits(bits_scalar(k1, mode));
  d.mdy = 0.01f * normal_from_bits(bits_scalar(k2, mode));
  const uint32_t span = n > 1 ? (uint32_t)(n - 1) : 1u;
  uint32_t ti = randint_offset(bits_at(kk1, n, i, mode), bits_at(kk2, n, i, mode), span);
  d.target = ti + (ti >= (uint32_t)i ? 1u : 0u);
  return d;
}

// Every draw of a run, grid-wide, into draws[steps][3][n] (mdx, mdy bits,
// target): none depends on the state, so the sequential run only applies
// them.
__global__ void __launch_bounds__(256)
    k_wealth_draws(const abm_wealth_inst *__restrict__ insts, int mode) {
  const abm_wealth_inst in = insts[blockIdx.y];
  if (!in.draws) return;
  const int64_t n = in.n;
  const int64_t total = (int64_t)in.steps * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int32_t t = (int32_t)(e / n);
    const int64_t i = e - (int64_t)t * n;
    const WealthDraw d = wealth_draw(in.keys, t, n, i, mode);
    uint32_t *row = in.draws + (int64_t)t * 3 * n;
    row[i] = __float_as_uint(d.mdx);
    row[n + i] = __float_as_uint(d.mdy);
    row[2 * n + i] = d.target;
  }
}

__device__ __forceinline__ void wealth_steps(const abm_wealth_inst &in, float *gx, float *gy,
                                             int32_t *gw, int32_t *ginc, int32_t *ghist,
                                             uint32_t *red, uint32_t &carry, int mode) {
  const int64_t n = in.n;
  const int tid = threadIdx.x;
  for (int32_t t = 0; t < in.steps; ++t) {
    // agent phase + exchange targets
    for (int64_t i = tid; i < n; i += kWealthThreads) {
      WealthDraw d;
      if (in.draws) {
        const uint32_t *row = in.draws + (int64_t)t * 3 * n;
        d.mdx = __uint_as_float(row[i]);
        d.mdy = __uint_as_float(row[n + i]);
        d.target = row[2 * n + i];
      } else {
        d = wealth_draw(in.keys, t, n, i, mode);
      }
      gx[i] = mod1(gx[i] + d.mdx);
      gy[i] = mod1(gy[i] + d.mdy);
      if (gw[i] > 0) atomicAdd(&ginc[d.target], 1);
    }
    __syncthreads();
    // apply the exchange, histogram of wealth values
    uint32_t total = 0;
    for (int64_t i = tid; i < n; i += kWealthThreads) {
      const int32_t w0 = gw[i];
      const int32_t wi = w0 - (w0 > 0 ? 1 : 0) + ginc[i];
      gw[i] = wi;
      ginc[i] = 0;
      atomicAdd(&ghist[wi], 1);  // 0 <= wi <= n: wealth is conserved
      total += (uint32_t)wi;
    }
    const uint32_t T = block_sum_u32(total, red);
    __syncthreads();
    // S = sum_w w * (c_w b_w + c_w (c_w + 1) / 2): chunked exclusive scan of
    // the histogram, one bin per thread per chunk
    if (tid == 0) carry = 0;
    __syncthreads();
    uint32_t S_part = 0;
    for (int64_t base = 0; base <= n; base += kWealthThreads) {
      const int64_t w = base + tid;
      const uint32_t c = w <= n ? (uint32_t)ghist[w] : 0u;
      // inclusive wave scan of c, then across waves
      uint32_t inc = c;
      const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
      }
      if (lane == 63) red[wave] = inc;
      __syncthreads();
      uint32_t wave_off = 0;
      for (int q = 0; q < wave; ++q) wave_off += red[q];
      const uint32_t before = carry + wave_off + inc - c;
      if (c) {
        const uint64_t ranksum = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
        S_part += (uint32_t)w * (uint32_t)ranksum;
        ghist[w] = 0;
      }
      __syncthreads();
      if (tid == kWealthThreads - 1) carry = before + c;
      __syncthreads();
    }
    const uint32_t S = block_sum_u32(S_part, red);
    if (tid == 0) {
      in.mean_out[t] = (float)(int32_t)T / (float)n;
      const int32_t Ti = (int32_t)T;
      float g = 0.0f;
      if ((float)Ti > 1e-6f) {
        const int32_t num = (int32_t)(2u * S);
        const int32_t den = (int32_t)((uint32_t)n * T);
        const float q = (float)num / (float)den;
        const float cst = (float)((double)(n + 1) / (double)n);
        g = q - cst;
      }
      in.gini_out[t] = g;
    }
    __syncthreads();
  }
}

// n + 1 <= kWealthThreads with precomputed draws: one agent and one
// histogram bin per thread, x, y and wealth in registers for the whole run,
// increments and two histograms (steps alternate) in LDS; two barriers per
// step.  Wealth is conserved, so the total T is summed once.  The rank sum S
// of step t is taken by wave 0 alone from histogram t & 1 while the other
// waves go on to step t + 1; that histogram's bins are cleared by their
// owners after the next step's first barrier, when wave 0 has read them.
__device__ __forceinline__ void wealth_steps_small(const abm_wealth_inst &in, int32_t *linc2,
                                                   int32_t *lhist2, uint32_t *red_t) {
  const int64_t n = in.n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool agent = tid < n;
  float x = agent ? in.x[tid] : 0.0f, y = agent ? in.y[tid] : 0.0f;
  int32_t w = agent ? in.w[tid] : 0;
  if (tid <= n) {
    linc2[tid] = 0;
    linc2[kWealthThreads + tid] = 0;
    lhist2[tid] = 0;
    lhist2[kWealthThreads + tid] = 0;
  }
  // T: the total wealth, the same at every step
  uint32_t tw = agent ? (uint32_t)w : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tw += __shfl_xor(tw, off);
  if (lane == 0) red_t[wave] = tw;
  __syncthreads();
  uint32_t T = 0;
  for (int q = 0; q < kWealthThreads / 64; ++q) T += red_t[q];
  const float mean = (float)(int32_t)T / (float)n;
  const float cst = (float)((double)(n + 1) / (double)n);
  const int32_t steps = in.steps;
  // Mean and Gini of step s from its histogram (wave 0), clearing the bins
  // it reads: rank sum 64 bins a round until all n agents are counted (one
  // round while every wealth is below 64); the nonzero bins all lie in the
  // rounds read.
  auto record = [&](int32_t s) {
    int32_t *hist = lhist2 + (s & 1) * kWealthThreads;
    uint32_t sp = 0, counted = 0;
    for (int64_t b0 = 0; counted < (uint32_t)n && b0 <= n; b0 += 64) {
      const int64_t wv = b0 + lane;
      const uint32_t c = wv <= n ? (uint32_t)hist[wv] : 0u;
      if (wv <= n) hist[wv] = 0;
      uint32_t inc = c;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
      }
      const uint32_t before = counted + inc - c;
      if (c) {
        const uint64_t ranksum = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
        sp += (uint32_t)wv * (uint32_t)ranksum;
      }
      counted += __shfl(inc, 63);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sp += __shfl_xor(sp, off);
    if (lane == 0) {
      in.mean_out[s] = mean;
      const int32_t Ti = (int32_t)T;
      float g = 0.0f;
      if ((float)Ti > 1e-6f) {
        const int32_t num = (int32_t)(2u * sp);
        const int32_t den = (int32_t)((uint32_t)n * T);
        const float q = (float)num / (float)den;
        g = q - cst;
      }
      in.gini_out[s] = g;
    }
  };
  // Step t's agent phase (walk, one exchange unit into linc[t & 1]) runs in
  // the same phase as step t - 1's exchange (which reads and clears
  // linc[(t - 1) & 1]), one barrier per step; wave 0 records step t - 1
  // after the next barrier, while the others go on.  Each agent's own draw
  // words of step t + 1 are loaded a step ahead.
  uint32_t dx = 0u, dy = 0u, dt = 0u;
  if (agent && steps > 0) {
    dx = in.draws[tid];
    dy = in.draws[n + tid];
    dt = in.draws[2 * n + tid];
  }
  auto agent_phase = [&](int32_t t) {
    const float mdx = __uint_as_float(dx), mdy = __uint_as_float(dy);
    const uint32_t target = dt;
    if (t + 1 < steps) {
      const uint32_t *row = in.draws + (int64_t)(t + 1) * 3 * n;
      dx = row[tid];
      dy = row[n + tid];
      dt = row[2 * n + tid];
    }
    x = mod1(x + mdx);
    y = mod1(y + mdy);
    if (w > 0) atomicAdd(&linc2[(t & 1) * kWealthThreads + target], 1);
  };
  if (agent && steps > 0) agent_phase(0);
  for (int32_t t = 0; t < steps; ++t) {
    __syncthreads();  // step t's exchange units are in; step t - 1's histogram is complete
    if (wave == 0 && t > 0) record(t - 1);
    if (agent) {
      int32_t *linc = linc2 + (t & 1) * kWealthThreads;
      w = w - (w > 0 ? 1 : 0) + linc[tid];
      linc[tid] = 0;
      atomicAdd(&lhist2[(t & 1) * kWealthThreads + w], 1);  // 0 <= w <= n: conserved
      if (t + 1 < steps) agent_phase(t + 1);
    }
  }
  if (steps > 0) {
    __syncthreads();
    if (wave == 0) record(steps - 1);
  }
  if (agent) {
    in.x[tid] = x;
    in.y[tid] = y;
    in.w[tid] = w;
  }
}

// kLds: the launch carries dynamic LDS for max_lds agents; instances with
// n <= max_lds run out of it, others on their global buffers.
template <bool kLds>
__global__ void __launch_bounds__(kWealthThreads)
    k_wealth_run(const abm_wealth_inst *__restrict__ insts, int64_t max_lds, int mode) {
  extern __shared__ int32_t lds_agents[];
  __shared__ uint32_t red[16];
  __shared__ uint32_t carry;
  __shared__ uint32_t red_t[kWealthThreads / 64], red_s[kWealthThreads / 64];
  const abm_wealth_inst in = insts[blockIdx.x];
  const int64_t n = in.n;
  const int tid = threadIdx.x;
  if (in.draws && n + 1 <= kWealthThreads) {
    // two increment arrays and two histograms: the dynamic LDS (at least 4 x
    // kWealthThreads words in this launch), or a static array where the
    // launch has none
    if constexpr (kLds) {
      wealth_steps_small(in, lds_agents, lds_agents + 2 * kWealthThreads, red_t);
    } else {
      __shared__ int32_t small_lds[4 * kWealthThreads];
      wealth_steps_small(in, small_lds, small_lds + 2 * kWealthThreads, red_t);
    }
  } else if (kLds && n <= max_lds) {
    float *lx = reinterpret_cast<float *>(lds_agents);
    float *ly = lx + n;
    int32_t *lw = reinterpret_cast<int32_t *>(ly + n);
    int32_t *linc = lw + n;
    int32_t *lhist = linc + n;
    for (int64_t i = tid; i < n; i += kWealthThreads) {
      lx[i] = in.x[i];
      ly[i] = in.y[i];
      lw[i] = in.w[i];
      linc[i] = 0;
      lhist[i] = 0;
    }
    if (tid == 0) lhist[n] = 0;
    __syncthreads();
    wealth_steps(in, lx, ly, lw, linc, lhist, red, carry, mode);
    for (int64_t i = tid; i < n; i += kWealthThreads) {
      in.x[i] = lx[i];
      in.y[i] = ly[i];
      in.w[i] = lw[i];
    }
  } else {
    wealth_steps(in, in.x, in.y, in.w, in.inc, in.hist, red, carry, mode);
  }
}

// ---- one large instance across the whole GPU (abm_wealth_run_grid): two
// launches per step.  K1(t): every workgroup draws step t for its agents and
// applies the draws (positions; one exchange atomic per giving agent), and
// workgroup 0 first records step t-1 from the histogram.  K2(t): the
// exchange is applied and the wealth histogram built.  Wealth stays bunched
// on a few small values, so the histogram is gathered without hot words:
// per-workgroup LDS bins, flushed with one atomic per bin in use, the
// values below kGridHot into one of kGridCopies copies (chosen by workgroup)
// and the rest into the instance's histogram.  The total T is conserved, so
// it is summed once, at step 0; the record stops once it has counted all n
// agents, so no running maximum is kept.  The same integer and IEEE float
// operations as the one-workgroup run, so the same bits.
constexpr int kGridThreads = 256;
constexpr int kGridBins = 1024;     // LDS histogram bins per workgroup
constexpr int64_t kGridMaxBlocks = 1024;
constexpr int kGridHot = 64;        // values with their own histogram copies
constexpr int kGridCopies = 16;
static_assert(ABM_WEALTH_GRID_ACC_WORDS == 2 + kGridCopies * kGridHot,
              "acc layout: T, a spare word, then the hot-bin copies");

// Mean and Gini of step t from the histogram and the total; clears the bins
// it reads (one workgroup).  Bins are read 256 a round until all n agents
// are counted, which covers every nonzero bin.
__device__ void wealth_grid_record(const abm_wealth_inst &in, int32_t t, uint32_t *acc,
                                   uint32_t *red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n = in.n;
  const uint32_t T = acc[0];
  uint32_t *hot = acc + 2;
  __shared__ uint32_t carry;
  if (tid == 0) carry = 0;
  __syncthreads();
  uint32_t S_part = 0;
  for (int64_t base = 0; base <= n; base += kGridThreads) {
    const int64_t w = base + tid;
    uint32_t c = 0;
    if (w < kGridHot) {
      uint32_t v[kGridCopies];
#pragma unroll
      for (int k = 0; k < kGridCopies; ++k) v[k] = hot[k * kGridHot + w];
#pragma unroll
      for (int k = 0; k < kGridCopies; ++k) {
        c += v[k];
        if (v[k]) hot[k * kGridHot + w] = 0;
      }
    } else if (w <= n) {
      c = (uint32_t)in.hist[w];
      if (c) in.hist[w] = 0;
    }
    uint32_t inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0;
    for (int q = 0; q < wave; ++q) wave_off += red[q];
    const uint32_t before = carry + wave_off + inc - c;
    if (c) {
      const uint64_t ranksum = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
      S_part += (uint32_t)w * (uint32_t)ranksum;
    }
    __syncthreads();
    if (tid == kGridThreads - 1) carry = before + c;
    __syncthreads();
    if ((int64_t)carry >= n) break;  // block-uniform
  }
  const uint32_t S = block_sum_u32(S_part, red);
  if (tid == 0) {
    in.mean_out[t] = (float)(int32_t)T / (float)n;
    const int32_t Ti = (int32_t)T;
    float g = 0.0f;
    if ((float)Ti > 1e-6f) {
      const int32_t num = (int32_t)(2u * S);
      const int32_t den = (int32_t)((uint32_t)n * T);
      const float q = (float)num / (float)den;
      const float cst = (float)((double)(n + 1) / (double)n);
      g = q - cst;
    }
    in.gini_out[t] = g;
  }
}

__global__ void __launch_bounds__(kGridThreads)
    k_wealth_grid_agents(abm_wealth_inst in, int32_t t, uint32_t *acc, int mode) {
  __shared__ uint32_t red[kGridThreads / 64];
  if (blockIdx.x == 0 && t > 0) wealth_grid_record(in, t - 1, acc, red);
  if (t >= in.steps) return;
  const int64_t n = in.n;
  const int64_t stride = (int64_t)gridDim.x * kGridThreads;
  for (int64_t i = (int64_t)blockIdx.x * kGridThreads + threadIdx.x; i < n; i += stride) {
    const WealthDraw d = wealth_draw(in.keys, t, n, i, mode);
    in.x[i] = mod1(in.x[i] + d.mdx);
    in.y[i] = mod1(in.y[i] + d.mdy);
    if (in.w[i] > 0) atomicAdd(&in.inc[d.target], 1);
  }
}

__global__ void __launch_bounds__(kGridThreads)
    k_wealth_grid_apply(abm_wealth_inst in, uint32_t *acc, int32_t t) {
  __shared__ uint32_t red[kGridThreads / 64];
  __shared__ uint32_t bins[kGridBins];
  for (int b = threadIdx.x; b < kGridBins; b += kGridThreads) bins[b] = 0;
  __syncthreads();
  const int64_t n = in.n;
  const int64_t stride = (int64_t)gridDim.x * kGridThreads;
  uint32_t total = 0;
  for (int64_t i = (int64_t)blockIdx.x * kGridThreads + threadIdx.x; i < n; i += stride) {
    const int32_t w0 = in.w[i];
    const int32_t wi = w0 - (w0 > 0 ? 1 : 0) + in.inc[i];
    in.w[i] = wi;
    in.inc[i] = 0;
    // 0 <= wi <= n: wealth is conserved
    if (wi < kGridBins)
      atomicAdd(&bins[wi], 1u);
    else
      atomicAdd(&in.hist[wi], 1);
    total += (uint32_t)wi;
  }
  __syncthreads();
  uint32_t *hot = acc + 2 + (int64_t)(blockIdx.x % kGridCopies) * kGridHot;
  for (int b = threadIdx.x; b < kGridBins; b += kGridThreads) {
    const uint32_t c = bins[b];
    if (!c) continue;
    if (b < kGridHot)
      atomicAdd(&hot[b], c);
    else
      atomicAdd(&in.hist[b], (int32_t)c);
  }
  if (t == 0) {  // the total, conserved from here on
    const uint32_t T = block_sum_u32(total, red);
    if (threadIdx.x == 0) atomicAdd(&acc[0], T);
  }
}

__global__ void k_wealth_zero_acc(uint32_t *acc) {
  for (int i = threadIdx.x; i < ABM_WEALTH_GRID_ACC_WORDS; i += blockDim.x) acc[i] = 0;
}

__global__ void k_wealth_init(const uint32_t *__restrict__ init_key, int64_t n,
                              float *__restrict__ x, float *__restrict__ y,
                              int32_t *__restrict__ w, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Key kx, ky;
  split2(split_at(Key{init_key[0], init_key[1]}, n, i, mode), kx, ky, mode);
  x[i] = unit_float(bits_scalar(kx, mode));
  y[i] = unit_float(bits_scalar(ky, mode));
  w[i] = 1;
}

// The LDS-resident run needs more than 64 KB of dynamic LDS: the limit is
// raised once per device, under a lock.
constexpr int kMaxDevices = 64;
std::mutex g_wealth_attr_mu;
bool g_wealth_attr_set[kMaxDevices];

int wealth_lds_attr() {
  int dev = -1;
  ABM_CHECK_HIP(hipGetDevice(&dev));
  ABM_REQUIRE(dev >= 0 && dev < kMaxDevices, "wealth: device index out of range");
  std::lock_guard<std::mutex> lock(g_wealth_attr_mu);
  if (!g_wealth_attr_set[dev]) {
    ABM_CHECK_HIP(hipFuncSetAttribute((const void *)k_wealth_run<true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(kWealthLdsAgents * kWealthLdsBytesPerAgent + 4)));
    g_wealth_attr_set[dev] = true;
  }
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_wealth_init(const uint32_t *init_key, int64_t n, float *x, float *y, int32_t *w,
                    int mode, void *stream) {
  ABM_REQUIRE(init_key && x && y && w && n > 1 && n < (1ll << 31) &&
                  (mode == 0 || mode == 1),
              "abm_wealth_init: bad arguments");
  hipLaunchKernelGGL(k_wealth_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, init_key, n, x, y, w, mode);
  ABM_LAUNCH_CHECK("abm_wealth_init");
  return kOk;
}

int abm_wealth_run_grid(const abm_wealth_inst *inst, uint32_t *acc, int mode, void *stream) {
  ABM_REQUIRE(inst && acc && (mode == 0 || mode == 1), "abm_wealth_run_grid: bad arguments");
  const abm_wealth_inst in = *inst;
  ABM_REQUIRE(in.n >= 2 && in.n < (1ll << 31) && in.steps >= 0 && in.x && in.y && in.w &&
                  in.inc && in.hist && in.keys && in.mean_out && in.gini_out,
              "abm_wealth_run_grid: incomplete instance");
  hipStream_t s = (hipStream_t)stream;
  if (in.steps == 0) return kOk;
  hipLaunchKernelGGL(k_wealth_zero_acc, dim3(1), dim3(256), 0, s, acc);
  int64_t blocks = (in.n + kGridThreads - 1) / kGridThreads;
  if (blocks > kGridMaxBlocks) blocks = kGridMaxBlocks;
  for (int32_t t = 0; t <= in.steps; ++t) {
    hipLaunchKernelGGL(k_wealth_grid_agents, dim3((unsigned)blocks), dim3(kGridThreads), 0, s,
                       in, t, acc, mode);
    if (t < in.steps)
      hipLaunchKernelGGL(k_wealth_grid_apply, dim3((unsigned)blocks), dim3(kGridThreads), 0, s,
                         in, acc, t);
  }
  ABM_LAUNCH_CHECK("abm_wealth_run_grid");
  return kOk;
}

int abm_wealth_run_batch(const abm_wealth_inst *insts, int32_t count, int64_t max_n, int mode,
                         void *stream) {
  ABM_REQUIRE(insts && count >= 0 && max_n >= 0 && (mode == 0 || mode == 1),
              "abm_wealth_run_batch: bad arguments");
  if (count == 0) return kOk;
  hipLaunchKernelGGL(k_wealth_draws, dim3(512u, (unsigned)count), dim3(256), 0,
                     (hipStream_t)stream, insts, mode);
  if (max_n > 0 && max_n <= kWealthLdsAgents) {
    if (int rc = wealth_lds_attr()) return rc;
    size_t dyn = (size_t)(max_n * kWealthLdsBytesPerAgent + 4);
    // an instance of n < kWealthThreads keeps its increments and two
    // histograms there
    if (dyn < 4 * kWealthThreads * sizeof(int32_t)) dyn = 4 * kWealthThreads * sizeof(int32_t);
    hipLaunchKernelGGL(k_wealth_run<true>, dim3((unsigned)count), dim3(kWealthThreads), dyn,
                       (hipStream_t)stream, insts, max_n, mode);
  } else {
    hipLaunchKernelGGL(k_wealth_run<false>, dim3((unsigned)count), dim3(kWealthThreads), 0,
                       (hipStream_t)stream, insts, (int64_t)0, mode);
  }
  ABM_LAUNCH_CHECK("abm_wealth_run_batch");
  return kOk;
}

}  // extern "C"
