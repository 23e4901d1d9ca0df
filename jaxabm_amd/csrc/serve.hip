// serve.hip -- continuous-batching simulation server: wealth-exchange and
// network-SIR requests of uneven sizes share one GPU.
//
// Replaces the reference's serial runners (jaxabm/analysis.py:97-165,
// jaxabm/utils.py:128-176, jaxabm/agentpy.py:1116-1152), which run one model
// after another.  A request is prepared on a host worker thread (SIR graph,
// initial infections, per-step key tables, all into pinned staging memory),
// admitted when an instance slot and room in the device arena (the HBM
// budget) are free, uploaded on a copy stream, and then advanced together
// with every other active instance in ticks of K steps:
//   resident path  one workgroup owns a small instance and runs its K steps
//                  in one launch, the working set in LDS (SIR: both state
//                  copies; wealth: packed 16-bit incoming counts and two
//                  alternating wealth histograms)
//   stepwise path  a large instance is split into agent ranges; each step is
//                  one launch over the ranges of every stepwise instance
//                  (SIR: double-buffered state; wealth: step t's agent phase
//                  and scatter, step t - 1's apply and histogram and step
//                  t - 2's record, with double-buffered increments)
// Instances join and leave at tick boundaries (continuous batching).  Every
// instance runs the same arithmetic whatever else is on the GPU, so its
// results are identical alone and in a mix.
//
// Programs (as the single-instance engines and the oracle define them):
//  wealth  agent keys = split(ck, n); k1, k2 = split(agent_key);
//          x = mod(x + 0.01 normal(k1), 1), y likewise with k2;
//          t = randint(uk, (n,), 0, n-1), t += t >= i; every agent with
//          wealth > 0 gives one unit to agent t; record mean and Gini
//  SIR     k_inf, k_rec = split(uk); S -> I if some CSR entry e of the row has
//          an infected neighbour and uniform(k_inf, (nnz,))[e] < beta;
//          I -> R if uniform(k_rec, (n,))[v] < gamma; record (S, I, R)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {
namespace serve {

enum : int32_t { kWealth = ABM_SERVE_WEALTH, kSir = ABM_SERVE_SIR };

constexpr int kResThreads = 1024;
constexpr int kResWaves = kResThreads / 64;
constexpr int kStepThreads = 256;
constexpr int kStepWaves = kStepThreads / 64;
constexpr int64_t kResidentMax = 65536;        // agents of a resident instance
constexpr int kHistLds = 2048;                 // wealth bins per LDS histogram (two, alternating)
constexpr size_t kResLds = 2 * kResidentMax + 2 * kHistLds * 4;  // 144 KiB
constexpr int kHeavyDeg = 64;                  // rows scanned by a whole wavefront
constexpr int kResHeavyCap = 1024;
constexpr int kStepHeavyCap = 256;
constexpr int kStepBins = 256;                 // per-block wealth bins (stepwise apply)
// Stepwise SIR skips an S agent whose neighbours cannot be infected: an agent
// infected after step s (newly, or on its refresh phase (s + v) % H == 0)
// writes mark = s + H into each neighbour; at step t an S agent scans its row
// only if (mark - t) mod 2^16 <= H.  An agent infected at the start of step t
// wrote a mark in some step of [t - H, t - 1] with a value in [t, t + H - 1],
// and later writes store larger values, so no S agent with an infected
// neighbour is skipped (stale marks only cost a scan).
constexpr uint32_t kMarkHorizon = 16;
constexpr uint16_t kMarkFar = 0x8000;          // initial mark: far from every step

struct Inst {
  int32_t kind, steps, t0, n;
  int64_t nnz;
  uint32_t thr_beta, thr_gamma;
  const uint32_t *keys;  // [steps][4]: SIR (k_inf, k_rec); wealth (ck, uk)
  const int32_t *row_ptr, *col;
  uint8_t *state0, *state1;  // SIR: resident home copy / stepwise double buffer
  uint16_t *mark;            // SIR: neighbour-infection marks (stepwise path)
  int32_t *counts;           // SIR [steps][3]
  float *x, *y;
  // wealth: w[n]; the stepwise path keeps two copies of the exchange
  // increments inc[2][n] and of the histogram hist[2][n+1], one per step
  // parity, and the total in acc[0] (the resident path uses copy 0 of hist)
  int32_t *w, *inc, *hist, *acc;
  float *mean, *gini;             // wealth [steps]
};

struct Range {
  int32_t inst, begin, end, kind;
};

// ---------------------------------------------------------------- block helpers
template <int THREADS>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int w = 0; w < THREADS / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Rank sum of the sorted wealth from its histogram: with c_w agents of
// wealth w and b_w below, their ranks sum to c_w b_w + c_w (c_w + 1) / 2, so
// S = sum_w w (c_w b_w + c_w (c_w + 1) / 2) mod 2^32.  Bins below lbins are in
// LDS (lh), the rest in global memory (gh).  Clears the bins it reads.
// Returns this thread's share of S.
template <int THREADS>
__device__ uint32_t rank_sum_and_clear(uint32_t *lh, int32_t lbins, int32_t *gh, int32_t maxw,
                                       uint32_t *red, uint32_t *carry) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) *carry = 0;
  __syncthreads();
  uint32_t S = 0;
  for (int32_t base = 0; base <= maxw; base += THREADS) {
    const int32_t w = base + tid;
    uint32_t c = 0;
    if (w <= maxw) c = w < lbins ? lh[w] : (uint32_t)gh[w];
    uint32_t inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    uint32_t before = *carry;
    for (int q = 0; q < wave; ++q) before += red[q];
    before += inc - c;
    if (c) {
      const uint64_t rs = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
      S += (uint32_t)w * (uint32_t)rs;
      if (w < lbins)
        lh[w] = 0;
      else
        gh[w] = 0;
    }
    __syncthreads();
    if (tid == THREADS - 1) *carry = before + c;
    __syncthreads();
  }
  return S;
}

// The same rank sum over a global histogram gh[0 .. n], read THREADS bins a
// round until all n agents are counted (every nonzero bin lies in the
// rounds read), so no largest wealth is needed.  Clears the bins it reads.
template <int THREADS>
__device__ uint32_t rank_sum_until_n(int32_t *gh, int64_t n, uint32_t *red, uint32_t *carry) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) *carry = 0;
  __syncthreads();
  uint32_t S = 0;
  for (int64_t base = 0; base <= n; base += THREADS) {
    const int64_t w = base + tid;
    const uint32_t c = w <= n ? (uint32_t)gh[w] : 0u;
    uint32_t inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) red[wave] = inc;
    __syncthreads();
    uint32_t before = *carry;
    for (int q = 0; q < wave; ++q) before += red[q];
    before += inc - c;
    if (c) {
      const uint64_t rs = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
      S += (uint32_t)w * (uint32_t)rs;
      gh[w] = 0;
    }
    __syncthreads();
    if (tid == THREADS - 1) *carry = before + c;
    __syncthreads();
    if ((int64_t)*carry >= n) break;  // block-uniform
  }
  return S;
}

// mean and Gini (examples/models/advanced_economic_model.py:1798-1811) from
// the int32 totals: G = f32(2 S) / f32(n T) - f32((n+1)/n) if T > 1e-6.
__device__ void wealth_record(int64_t n, uint32_t T, uint32_t S, float *mean, float *gini,
                              int32_t t) {
  mean[t] = (float)(int32_t)T / (float)n;
  const int32_t Ti = (int32_t)T;
  float g = 0.0f;
  if ((float)Ti > 1e-6f) {
    const int32_t num = (int32_t)(2u * S);
    const int32_t den = (int32_t)((uint32_t)n * T);
    const float q = (float)num / (float)den;
    const float cst = (float)((double)(n + 1) / (double)n);
    g = q - cst;
  }
  gini[t] = g;
}

// Agent phase of one wealth agent; returns the exchange target or -1.
__device__ __forceinline__ int32_t wealth_agent(const Inst &in, int64_t i, Key ck, Key kk1,
                                                Key kk2, uint32_t span, int mode) {
  Key k1, k2;
  split2(split_at(ck, in.n, i, mode), k1, k2, mode);
  const float dx = normal_from_bits(bits_scalar(k1, mode));
  const float dy = normal_from_bits(bits_scalar(k2, mode));
  const float sx = in.x[i] + 0.01f * dx;
  const float sy = in.y[i] + 0.01f * dy;
  in.x[i] = mod1(sx);
  in.y[i] = mod1(sy);
  if (in.w[i] <= 0) return -1;
  uint32_t ti = randint_offset(bits_at(kk1, in.n, i, mode), bits_at(kk2, in.n, i, mode), span);
  ti += ti >= (uint32_t)i ? 1u : 0u;
  return (int32_t)ti;
}

// Next SIR state of a light row (scanned by one thread).
__device__ __forceinline__ uint8_t sir_light(const Inst &in, const uint8_t *old, int64_t v,
                                             uint8_t s, int32_t e0, int32_t e1, Key kinf,
                                             Key krec, int mode) {
  if (s == 0) {
    for (int32_t e = e0; e < e1; ++e)
      if (old[in.col[e]] == 1 && (bits_at(kinf, in.nnz, e, mode) >> 9) < in.thr_beta) return 1;
    return 0;
  }
  if (s == 1) return (bits_at(krec, in.n, v, mode) >> 9) < in.thr_gamma ? 2 : 1;
  return s;
}

// Infection test of a heavy row by one wavefront (any infecting entry
// decides, so the scan order does not matter).
__device__ __forceinline__ bool sir_heavy_hit(const Inst &in, const uint8_t *old, int32_t e0,
                                              int32_t e1, Key kinf, int mode) {
  const int lane = threadIdx.x & 63;
  for (int32_t base = e0; base < e1; base += 64) {
    const int32_t e = base + lane;
    const bool mine = e < e1 && old[in.col[e]] == 1 &&
                      (bits_at(kinf, in.nnz, e, mode) >> 9) < in.thr_beta;
    if (__ballot(mine)) return true;
  }
  return false;
}

// ---------------------------------------------------------------- resident path
// Two barriers per step: after the light rows (heavy-row list complete), and
// after the heavy rows and the counts.  The S/I/R counts and the heavy-row
// count live in LDS, double-buffered by step parity: a step's buffer is
// reset during the step before it, after every thread has read it.
__device__ void sir_resident(const Inst &in, int32_t t_end, int mode, uint8_t *lds,
                             int32_t *heavy, int32_t *nheavy2, uint32_t *cnt2) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n = in.n;
  // a step after one that left no infected agent is the identity: its
  // counts are the previous step's, and the state (not part of the results)
  // need not be touched
  if (in.t0 > 0 && in.counts[3 * (in.t0 - 1) + 1] == 0) {
    for (int64_t i = tid; i < 3 * (int64_t)(t_end - in.t0); i += kResThreads)
      in.counts[3 * in.t0 + i] = in.counts[3 * (in.t0 - 1) + i % 3];
    return;
  }
  uint8_t *cur = lds, *nxt = lds + kResidentMax;
  for (int64_t i = tid; i < n; i += kResThreads) cur[i] = in.state0[i];
  if (tid == 0) nheavy2[in.t0 & 1] = 0;
  if (tid < 3) cnt2[3 * (in.t0 & 1) + tid] = 0;
  __syncthreads();
  for (int32_t t = in.t0; t < t_end; ++t) {
    const Key kinf{in.keys[4 * t], in.keys[4 * t + 1]};
    const Key krec{in.keys[4 * t + 2], in.keys[4 * t + 3]};
    int32_t *nheavy = nheavy2 + (t & 1);
    uint32_t *cnt = cnt2 + 3 * (t & 1);
    uint32_t cs = 0, ci = 0, cr = 0;
    for (int64_t v = tid; v < n; v += kResThreads) {
      uint8_t s = cur[v];
      const int32_t e0 = in.row_ptr[v], e1 = in.row_ptr[v + 1];
      if (s == 0 && e1 - e0 > kHeavyDeg) {
        const int32_t slot = atomicAdd(nheavy, 1);
        if (slot < kResHeavyCap) {
          heavy[slot] = (int32_t)v;
          continue;
        }
      }
      s = sir_light(in, cur, v, s, e0, e1, kinf, krec, mode);
      nxt[v] = s;
      cs += s == 0;
      ci += s == 1;
      cr += s == 2;
    }
    __syncthreads();
    const int32_t nh = min(*nheavy, kResHeavyCap);
    for (int32_t h = wave; h < nh; h += kResWaves) {
      const int32_t v = heavy[h];
      const bool hit = sir_heavy_hit(in, cur, in.row_ptr[v], in.row_ptr[v + 1], kinf, mode);
      if (lane == 0) {
        nxt[v] = hit ? 1 : 0;
        ci += hit;
        cs += !hit;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      cs += __shfl_xor(cs, off);
      ci += __shfl_xor(ci, off);
      cr += __shfl_xor(cr, off);
    }
    if (lane == 0) {
      if (cs) atomicAdd(&cnt[0], cs);
      if (ci) atomicAdd(&cnt[1], ci);
      if (cr) atomicAdd(&cnt[2], cr);
    }
    if (tid == 0) nheavy2[(t + 1) & 1] = 0;  // read last before this step's first barrier
    if (tid < 3) cnt2[3 * ((t + 1) & 1) + tid] = 0;
    __syncthreads();
    const uint32_t S = cnt[0], I = cnt[1], Rc = cnt[2];
    if (tid == 0) {
      in.counts[3 * t] = (int32_t)S;
      in.counts[3 * t + 1] = (int32_t)I;
      in.counts[3 * t + 2] = (int32_t)Rc;
    }
    uint8_t *tmp = cur;
    cur = nxt;
    nxt = tmp;
    if (I == 0) {  // block-uniform: every later step is the identity
      for (int64_t i = tid; i < 3 * (int64_t)(t_end - 1 - t); i += kResThreads)
        in.counts[3 * (t + 1) + i] = (int32_t)(i % 3 == 0 ? S : (i % 3 == 1 ? 0u : Rc));
      break;
    }
  }
  __syncthreads();
  for (int64_t i = tid; i < n; i += kResThreads) in.state0[i] = cur[i];
}

// Two barriers per step: after the exchange increments and after the
// histogram.  Wealth is conserved, so the total T is summed once per launch;
// the rank sum S of step t is taken by wave 0 alone from histogram t & 1 (64
// bins a round until all n agents are counted) while the other waves go on to
// step t + 1.  That histogram's LDS bins are cleared by their owners after
// the next step's first barrier; values of kHistLds or more count in the
// instance's global bins, which wave 0 clears as it reads them.  The packed
// increments of agents 2j and 2j + 1 are read and cleared by one thread
// before the second barrier.
__device__ void wealth_resident(const Inst &in, int32_t t_end, int mode, uint8_t *lds,
                                uint32_t *red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n = in.n;
  uint32_t *inc2 = reinterpret_cast<uint32_t *>(lds);  // two 16-bit counts per word
  uint32_t *hist2 = reinterpret_cast<uint32_t *>(lds + 2 * kResidentMax);  // 2 x kHistLds
  const int64_t words = (n + 1) >> 1;
  const uint32_t span = (uint32_t)(n - 1);
  for (int64_t i = tid; i < words; i += kResThreads) inc2[i] = 0;
  for (int i = tid; i < 2 * kHistLds; i += kResThreads) hist2[i] = 0;
  uint32_t tw = 0;
  for (int64_t i = tid; i < n; i += kResThreads) tw += (uint32_t)in.w[i];
  const uint32_t T = block_sum<kResThreads>(tw, red);  // its barriers order the clears too
  for (int32_t t = in.t0; t < t_end; ++t) {
    uint32_t *hist = hist2 + (t & 1) * kHistLds;
    uint32_t *prev = hist2 + ((t + 1) & 1) * kHistLds;
    const Key ck{in.keys[4 * t], in.keys[4 * t + 1]};
    const Key uk{in.keys[4 * t + 2], in.keys[4 * t + 3]};
    Key kk1, kk2;
    split2(uk, kk1, kk2, mode);
    for (int64_t i = tid; i < n; i += kResThreads) {
      const int32_t ti = wealth_agent(in, i, ck, kk1, kk2, span, mode);
      // an agent receives at most n - 1 < 2^16 units per step: no carry
      if (ti >= 0) atomicAdd(&inc2[ti >> 1], 1u << ((ti & 1) * 16));
    }
    __syncthreads();
    for (int b = tid; b < kHistLds; b += kResThreads) prev[b] = 0;
    for (int64_t j = tid; j < words; j += kResThreads) {
      const uint32_t word = inc2[j];
      inc2[j] = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t i = 2 * j + h;
        if (i >= n) break;
        const int32_t got = (int32_t)((word >> (16 * h)) & 0xFFFFu);
        const int32_t w0 = in.w[i];
        const int32_t wi = w0 - (w0 > 0 ? 1 : 0) + got;
        in.w[i] = wi;
        if (wi < kHistLds)
          atomicAdd(&hist[wi], 1u);
        else
          atomicAdd(&in.hist[wi], 1);  // 0 <= wi <= n: wealth is conserved
      }
    }
    __syncthreads();
    if (wave == 0) {
      uint32_t sp = 0, counted = 0;
      for (int64_t b0 = 0; counted < (uint32_t)n && b0 <= n; b0 += 64) {
        const int64_t wv = b0 + lane;
        uint32_t c = 0;
        if (wv <= n) {
          if (wv < kHistLds) {
            c = hist[wv];
          } else {
            c = (uint32_t)in.hist[wv];
            if (c) in.hist[wv] = 0;
          }
        }
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t y = __shfl_up(incl, off);
          if (lane >= off) incl += y;
        }
        const uint32_t before = counted + incl - c;
        if (c) {
          const uint64_t rs = (uint64_t)c * before + (uint64_t)c * (c + 1) / 2;
          sp += (uint32_t)wv * (uint32_t)rs;
        }
        counted += __shfl(incl, 63);
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) sp += __shfl_xor(sp, off);
      if (lane == 0) wealth_record(n, T, sp, in.mean, in.gini, t);
    }
  }
}

__global__ void __launch_bounds__(kResThreads)
    k_serve_resident(const Inst *__restrict__ insts, const int32_t *__restrict__ list, int32_t K,
                     int mode) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int32_t heavy[kResHeavyCap];
  __shared__ int32_t nheavy2[2];
  __shared__ uint32_t cnt2[6];
  __shared__ uint32_t red[kResWaves];
  const Inst in = insts[list[blockIdx.x]];
  const int32_t t_end = min(in.t0 + K, in.steps);
  if (in.t0 >= t_end) return;
  if (in.kind == kSir)
    sir_resident(in, t_end, mode, lds, heavy, nheavy2, cnt2);
  else
    wealth_resident(in, t_end, mode, lds, red);
}

// ---------------------------------------------------------------- stepwise path
// Mean and Gini of wealth step t (one workgroup): rank sum of the step's
// histogram copy (cleared for step t + 2) and the total, which wealth
// conservation keeps at acc[0] from step 0 on.
__device__ void wealth_stepwise_record(const Inst &in, int32_t t, uint32_t *red,
                                       uint32_t *carry) {
  int32_t *hist = in.hist + (int64_t)(t & 1) * (in.n + 1);
  const uint32_t T = (uint32_t)in.acc[0];
  uint32_t S = rank_sum_until_n<kStepThreads>(hist, in.n, red, carry);
  S = block_sum<kStepThreads>(S, red);
  if (threadIdx.x == 0) wealth_record(in.n, T, S, in.mean, in.gini, t);
}

// Wealth sub-step t of range r (see k_serve_phase1).  The exchange of step
// t - 1 and the agent phase of step t touch each agent from the same thread;
// step t's increments go to the other inc copy, so no agent reads an
// increment of the step being drawn.
__device__ void wealth_stepwise(const Inst &in, const Range &r, int32_t t, int mode,
                                uint32_t *red, uint32_t *carry, uint32_t *bins) {
  const int tid = threadIdx.x;
  const int64_t n = in.n;
  if (r.begin == 0 && t >= 2) wealth_stepwise_record(in, t - 2, red, carry);  // block-uniform
  const bool apply = t >= 1 && t - 1 < in.steps;  // block-uniform
  const bool draw = t < in.steps;
  int32_t *inc_old = in.inc + (int64_t)((t + 1) & 1) * n;  // step t - 1's increments
  int32_t *inc_new = in.inc + (int64_t)(t & 1) * n;
  int32_t *hist = in.hist + (int64_t)((t + 1) & 1) * (n + 1);  // step t - 1's histogram
  if (apply) {
    bins[tid] = 0;
    __syncthreads();
  }
  Key ck{0u, 0u}, kk1{0u, 0u}, kk2{0u, 0u};
  if (draw) {
    ck = Key{in.keys[4 * t], in.keys[4 * t + 1]};
    split2(Key{in.keys[4 * t + 2], in.keys[4 * t + 3]}, kk1, kk2, mode);
  }
  const uint32_t span = (uint32_t)(n - 1);
  uint32_t total = 0;
  for (int64_t i = r.begin + tid; i < r.end; i += kStepThreads) {
    if (apply) {
      const int32_t w0 = in.w[i];
      const int32_t wi = w0 - (w0 > 0 ? 1 : 0) + inc_old[i];
      in.w[i] = wi;
      inc_old[i] = 0;
      total += (uint32_t)wi;
      if (wi < kStepBins)
        atomicAdd(&bins[wi], 1u);
      else
        atomicAdd(&hist[wi], 1);  // 0 <= wi <= n: wealth is conserved
    }
    if (draw) {
      const int32_t ti = wealth_agent(in, i, ck, kk1, kk2, span, mode);
      if (ti >= 0) atomicAdd(&inc_new[ti], 1);
    }
  }
  if (apply) {
    __syncthreads();
    if (bins[tid]) atomicAdd(&hist[tid], (int32_t)bins[tid]);
    if (t == 1) {  // after step 0: the total, conserved from here on
      const uint32_t T = block_sum<kStepThreads>(total, red);
      if (tid == 0) atomicAdd(&in.acc[0], (int32_t)T);
    }
  }
}

// Phase 1 of sub-step s, the one launch per step: SIR step of the range, or
// for wealth (t = t0 + s): the record of step t - 2 (by the block of the
// instance's first range), the exchange of step t - 1 applied to the range's
// agents (wealth, histogram, total and maximum), then step t's agent phase
// (positions and exchange increments).  A wealth instance runs two sub-steps
// past its last step to apply and record it.
__global__ void __launch_bounds__(kStepThreads)
    k_serve_phase1(const Inst *__restrict__ insts, const Range *__restrict__ ranges, int32_t s,
                   int mode) {
  __shared__ int32_t heavy[kStepHeavyCap];
  __shared__ int32_t marking[kStepHeavyCap];
  __shared__ int32_t nheavy, nmark;
  __shared__ uint32_t red[kStepWaves];
  __shared__ uint32_t carry;
  __shared__ uint32_t bins[kStepBins];
  const Range r = ranges[blockIdx.x];
  const Inst in = insts[r.inst];
  const int32_t t = in.t0 + s;
  if (in.kind == kWealth) {
    if (t >= in.steps + 2) return;  // block-uniform
    wealth_stepwise(in, r, t, mode, red, &carry, bins);
    return;
  }
  if (t >= in.steps) return;  // block-uniform
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the step's keys and this thread's first agent (state, mark, row bounds)
  // are loaded together with the previous step's infected count
  const Key kinf{in.keys[4 * t], in.keys[4 * t + 1]};
  const Key krec{in.keys[4 * t + 2], in.keys[4 * t + 3]};
  const uint8_t *old = (t & 1) ? in.state1 : in.state0;
  uint8_t *nxt = (t & 1) ? in.state0 : in.state1;
  const int64_t v_first = r.begin + tid;
  const bool has_first = v_first < r.end;
  const uint8_t st_first = has_first ? old[v_first] : 0;
  const uint16_t mk_first = has_first ? in.mark[v_first] : 0;
  const int32_t e0_first = has_first ? in.row_ptr[v_first] : 0;
  const int32_t e1_first = has_first ? in.row_ptr[v_first + 1] : 0;
  // identity step (the previous one, computed by an earlier launch, left no
  // infected agent): only the counts carry over; the state, not part of the
  // results, is never read again
  if (t > 0 && in.counts[3 * (t - 1) + 1] == 0) {
    if (r.begin == 0 && tid < 3) in.counts[3 * t + tid] = in.counts[3 * (t - 1) + tid];
    return;
  }
  const uint16_t new_mark = (uint16_t)((uint32_t)t + kMarkHorizon);
  if (tid == 0) nheavy = nmark = 0;
  __syncthreads();
  uint32_t cs = 0, ci = 0, cr = 0;
  for (int64_t v = v_first; v < r.end; v += kStepThreads) {
    const bool first = v == v_first;
    const uint8_t st = first ? st_first : old[v];
    const uint16_t mk = first ? mk_first : in.mark[v];
    if (st == 0 && (uint16_t)(mk - (uint16_t)t) > kMarkHorizon) {
      nxt[v] = 0;  // no neighbour can be infected
      ++cs;
      continue;
    }
    const int32_t e0 = first ? e0_first : in.row_ptr[v];
    const int32_t e1 = first ? e1_first : in.row_ptr[v + 1];
    if (st == 0 && e1 - e0 > kHeavyDeg) {
      const int32_t slot = atomicAdd(&nheavy, 1);
      if (slot < kStepHeavyCap) {
        heavy[slot] = (int32_t)v;
        continue;
      }
    }
    const uint8_t ns = sir_light(in, old, v, st, e0, e1, kinf, krec, mode);
    nxt[v] = ns;
    cs += ns == 0;
    ci += ns == 1;
    cr += ns == 2;
    if (ns == 1 && (st == 0 || ((uint32_t)t + (uint32_t)v) % kMarkHorizon == 0)) {
      if (e1 - e0 > kHeavyDeg) {
        const int32_t slot = atomicAdd(&nmark, 1);
        if (slot < kStepHeavyCap) {
          marking[slot] = (int32_t)v;
          continue;
        }
      }
      for (int32_t e = e0; e < e1; ++e) in.mark[in.col[e]] = new_mark;
    }
  }
  __syncthreads();
  const int32_t nh = min(nheavy, kStepHeavyCap);
  for (int32_t h = wave; h < nh; h += kStepWaves) {
    const int32_t v = heavy[h];
    const int32_t e0 = in.row_ptr[v], e1 = in.row_ptr[v + 1];
    const bool hit = sir_heavy_hit(in, old, e0, e1, kinf, mode);
    if (lane == 0) {
      nxt[v] = hit ? 1 : 0;
      ci += hit;
      cs += !hit;
    }
    if (hit)  // wave-uniform: the newly infected hub marks its neighbours
      for (int32_t e = e0 + lane; e < e1; e += 64) in.mark[in.col[e]] = new_mark;
  }
  const int32_t nm = min(nmark, kStepHeavyCap);
  for (int32_t h = wave; h < nm; h += kStepWaves) {
    const int32_t v = marking[h];
    for (int32_t e = in.row_ptr[v] + lane; e < in.row_ptr[v + 1]; e += 64)
      in.mark[in.col[e]] = new_mark;
  }
  cs = block_sum<kStepThreads>(cs, red);
  ci = block_sum<kStepThreads>(ci, red);
  cr = block_sum<kStepThreads>(cr, red);
  if (tid == 0) {
    if (cs) atomicAdd(&in.counts[3 * t], (int32_t)cs);
    if (ci) atomicAdd(&in.counts[3 * t + 1], (int32_t)ci);
    if (cr) atomicAdd(&in.counts[3 * t + 2], (int32_t)cr);
  }
}

// ---------------------------------------------------------------- admission
__global__ void k_serve_init_wealth(Inst in, uint32_t ik0, uint32_t ik1, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < in.n) {
    Key kx, ky;
    split2(split_at(Key{ik0, ik1}, in.n, i, mode), kx, ky, mode);
    in.x[i] = unit_float(bits_scalar(kx, mode));
    in.y[i] = unit_float(bits_scalar(ky, mode));
    in.w[i] = 1;
    in.inc[i] = 0;
    in.inc[in.n + i] = 0;
  }
  if (i <= in.n) {
    in.hist[i] = 0;
    in.hist[in.n + 1 + i] = 0;
  }
  if (i < 4) in.acc[i] = 0;
}

__global__ void k_serve_init_sir(Inst in) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < in.n) {
    in.state0[i] = 0;
    in.mark[i] = kMarkFar;
  }
  if (i < 3 * (int64_t)in.steps) in.counts[i] = 0;
}

// The initial infections, and their marks (value H: at risk for steps 0 .. H,
// after which the refresh phases take over).
__global__ void k_serve_seed_sir(Inst in, const int32_t *__restrict__ infected, int32_t n0) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) {
    const int32_t v = infected[i];
    in.state0[v] = 1;
    for (int32_t e = in.row_ptr[v]; e < in.row_ptr[v + 1]; ++e)
      in.mark[in.col[e]] = (uint16_t)kMarkHorizon;
  }
}

// ---------------------------------------------------------------- host side
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Pinned {
  void *p = nullptr;
  size_t cap = 0;
};

// Pinned staging buffers in power-of-two size classes, reused (freeing pinned
// memory while kernels run can serialise the device).
class PinnedPool {
 public:
  hipError_t get(size_t bytes, Pinned *out) {
    size_t cap = (size_t)1 << 16;
    while (cap < bytes) cap <<= 1;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto &v = free_[cap];
      if (!v.empty()) {
        out->p = v.back();
        out->cap = cap;
        v.pop_back();
        return hipSuccess;
      }
    }
    out->cap = cap;
    return hipHostMalloc(&out->p, cap, hipHostMallocDefault);
  }
  void put(Pinned b) {
    if (!b.p) return;
    std::lock_guard<std::mutex> g(mu_);
    free_[b.cap].push_back(b.p);
  }
  ~PinnedPool() {
    for (auto &kv : free_)
      for (void *p : kv.second) (void)hipHostFree(p);
  }

 private:
  std::mutex mu_;
  std::map<size_t, std::vector<void *>> free_;
};

// First-fit allocator over the device arena (scheduler thread only).
class Arena {
 public:
  void init(size_t size) {
    free_.clear();
    free_[0] = size;
    size_ = size;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = align256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, o = it->first;
      free_.erase(it);
      if (rest) free_[o + bytes] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align256(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

enum : int32_t {
  kQueued = ABM_SERVE_QUEUED,
  kPrepared = ABM_SERVE_PREPARED,
  kUploading = ABM_SERVE_UPLOADING,
  kActive = ABM_SERVE_ACTIVE,
  kDone = ABM_SERVE_DONE,
  kFailed = ABM_SERVE_FAILED,
};

struct Req {
  int64_t ticket = 0;
  abm_request q{};
  int32_t status = kQueued;  // guarded by the server mutex
  double t_submit = 0, t_ready = 0, t_admit = 0, t_start = 0, t_done = 0;
  std::string error;
  std::vector<uint8_t> result;
  // preparation
  Pinned staging;
  size_t upload_bytes = 0, dev_bytes = 0, result_off = 0, result_bytes = 0;
  size_t off_keys = 0, off_rowptr = 0, off_col = 0, off_infected = 0, off_state0 = 0,
         off_state1 = 0, off_mark = 0, off_counts = 0, off_x = 0, off_y = 0, off_w = 0, off_inc = 0,
         off_hist = 0, off_acc = 0, off_mean = 0, off_gini = 0;
  int64_t nnz = 0;
  int32_t n0 = 0;
  uint32_t ik[2] = {0, 0};
  bool resident = false;
  std::vector<int32_t> ranges;  // stepwise agent range bounds
  // device (scheduler thread)
  size_t dev_off = 0;
  Inst d{};
  hipEvent_t up_ev = nullptr;
  int32_t t0 = 0;
};

struct Tick {
  hipEvent_t ev_r = nullptr, ev_w = nullptr;
  std::vector<Req *> finishing;
};

class Server {
 public:
  explicit Server(const abm_serve_config &c) : cfg_(c) {}
  int start();
  ~Server();
  int submit(const abm_request &q, int64_t *ticket);
  int poll(int64_t ticket, int32_t *status);
  int wait(int64_t ticket, double timeout, int32_t *status);
  int collect(int64_t ticket, void *out, size_t bytes, abm_ticket_info *info);
  int drain(double timeout);
  void stats(abm_serve_stats *st);

 private:
  void prep_loop();
  void prepare(Req *r);
  void sched_loop();
  bool sched_once();
  int launch_tick();
  bool retire_tick(bool block);
  void admit();
  void activate_uploads();
  void fail(Req *r, const std::string &msg);
  int ensure_tables(size_t n_inst, size_t n1);

  abm_serve_config cfg_;
  int device_ = 0;
  std::mutex mu_;
  std::condition_variable cv_prep_, cv_sched_, cv_done_;
  std::deque<Req *> prep_q_, ready_q_;
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs_;
  int64_t next_ticket_ = 1;
  int64_t outstanding_ = 0;  // submitted and not yet done/failed
  bool stop_ = false;
  std::vector<std::thread> workers_;
  std::thread sched_;
  std::string sched_error_;

  // scheduler thread state
  std::vector<Req *> uploading_, active_;
  std::deque<Tick> inflight_;
  std::vector<hipEvent_t> free_events_;
  Arena arena_;
  char *dev_base_ = nullptr;
  hipStream_t sR_ = nullptr, sW_ = nullptr, sC_ = nullptr, sD_ = nullptr;
  hipEvent_t ev_tab_ = nullptr;
  PinnedPool pool_;
  static constexpr int kSets = 4;
  struct TableSet {
    void *host = nullptr, *dev = nullptr;
    size_t bytes = 0;
  } sets_[kSets];
  int64_t tick_idx_ = 0;
  // statistics
  std::atomic<int64_t> n_ticks_{0}, n_done_{0}, agent_steps_{0};
  std::atomic<int64_t> peak_active_{0}, peak_bytes_{0};
};

hipEvent_t make_event() {
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

int Server::start() {
  ABM_CHECK_HIP(hipGetDevice(&device_));
  ABM_CHECK_HIP(hipStreamCreateWithFlags(&sR_, hipStreamNonBlocking));
  ABM_CHECK_HIP(hipStreamCreateWithFlags(&sW_, hipStreamNonBlocking));
  ABM_CHECK_HIP(hipStreamCreateWithFlags(&sC_, hipStreamNonBlocking));
  ABM_CHECK_HIP(hipStreamCreateWithFlags(&sD_, hipStreamNonBlocking));
  ev_tab_ = make_event();
  if (!ev_tab_) return hip_fail(hipErrorOutOfMemory, "abm_serve_create: event");
  void *base = nullptr;
  ABM_CHECK_HIP(hipMalloc(&base, cfg_.hbm_budget));
  dev_base_ = (char *)base;
  arena_.init(cfg_.hbm_budget);
  ABM_CHECK_HIP(hipFuncSetAttribute((const void *)k_serve_resident,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kResLds));
  if (int rc = ensure_tables(cfg_.max_active, 4096)) return rc;
  for (int i = 0; i < cfg_.prep_threads; ++i) workers_.emplace_back([this] { prep_loop(); });
  sched_ = std::thread([this] { sched_loop(); });
  return kOk;
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_prep_.notify_all();
  cv_sched_.notify_all();
  for (auto &t : workers_) t.join();
  if (sched_.joinable()) sched_.join();
  if (sW_) (void)hipStreamSynchronize(sW_);
  if (sR_) (void)hipStreamSynchronize(sR_);
  if (sC_) (void)hipStreamSynchronize(sC_);
  for (auto &kv : reqs_) {
    Req *r = kv.second.get();
    pool_.put(r->staging);
    if (r->up_ev) (void)hipEventDestroy(r->up_ev);
  }
  for (auto &t : inflight_) {
    free_events_.push_back(t.ev_r);
    free_events_.push_back(t.ev_w);
  }
  for (hipEvent_t e : free_events_) (void)hipEventDestroy(e);
  if (ev_tab_) (void)hipEventDestroy(ev_tab_);
  for (auto &s : sets_) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
  }
  if (dev_base_) (void)hipFree(dev_base_);
  for (hipStream_t s : {sR_, sW_, sC_, sD_})
    if (s) (void)hipStreamDestroy(s);
}

// Table sets: pinned host / device pairs, each at least `bytes`; a tick
// packs Inst[A] | Range[n1] | int32 resident[] at 256-byte aligned offsets
// and copies only what it uses.
int Server::ensure_tables(size_t n_inst, size_t n1) {
  const size_t need = align256(sizeof(Inst) * n_inst) + align256(sizeof(Range) * n1) +
                      align256(4 * n_inst);
  if (need <= sets_[0].bytes) return kOk;
  // the sets may be in use by queued ticks: drain them first
  for (hipStream_t s : {sR_, sW_}) ABM_CHECK_HIP(hipStreamSynchronize(s));
  const size_t bytes = std::max(need * 2, (size_t)1 << 20);
  for (auto &s : sets_) {
    if (s.host) ABM_CHECK_HIP(hipHostFree(s.host));
    if (s.dev) ABM_CHECK_HIP(hipFree(s.dev));
    s.host = s.dev = nullptr;
    s.bytes = 0;
    ABM_CHECK_HIP(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
    ABM_CHECK_HIP(hipMalloc(&s.dev, bytes));
    s.bytes = bytes;
  }
  return kOk;
}

int Server::submit(const abm_request &q, int64_t *ticket) {
  auto r = std::make_unique<Req>();
  r->q = q;
  r->t_submit = now_s();
  Req *raw = r.get();
  {
    std::lock_guard<std::mutex> g(mu_);
    if (!sched_error_.empty()) {
      set_error("abm_serve_submit: server stopped: " + sched_error_);
      return kErrInvalid;
    }
    raw->ticket = next_ticket_++;
    *ticket = raw->ticket;
    reqs_[raw->ticket] = std::move(r);
    prep_q_.push_back(raw);
    ++outstanding_;
  }
  cv_prep_.notify_one();
  return kOk;
}

void Server::fail(Req *r, const std::string &msg) {
  {
    std::lock_guard<std::mutex> g(mu_);
    r->status = kFailed;
    r->error = msg;
    r->t_done = now_s();
    --outstanding_;
  }
  cv_done_.notify_all();
}

void Server::prep_loop() {
  if (hipSetDevice(device_) != hipSuccess) return;
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_prep_.wait(lk, [this] { return stop_ || !prep_q_.empty(); });
      if (stop_) return;
      r = prep_q_.front();
      prep_q_.pop_front();
    }
    prepare(r);
  }
}

// permutation(key, n)[:n0] on the host: ceil(3 ln n / ln(2^32 - 1)) rounds
// of a stable sort of x by fresh 32-bit keys.  All rounds but the last are
// stable LSD radix sorts; the last only needs its n0 smallest (key, position)
// pairs, so it is a selection followed by a sort of those n0.
std::vector<int32_t> host_permutation_prefix(Key key, int64_t n, int32_t n0, int mode) {
  std::vector<int32_t> x(n), x2(n);
  std::vector<uint32_t> k(n), k2(n);
  for (int64_t i = 0; i < n; ++i) x[i] = (int32_t)i;
  const int rounds = (int)__builtin_ceil(3.0 * __builtin_log((double)(n > 1 ? n : 1)) /
                                         __builtin_log(4294967295.0));
  const int64_t take = std::min<int64_t>(n0, n);
  for (int r = 0; r < rounds; ++r) {
    Key nk, sub;
    split2(key, nk, sub, mode);
    key = nk;
    for (int64_t i = 0; i < n; ++i) k[i] = bits_at(sub, n, i, mode);
    if (r + 1 == rounds) {
      std::vector<uint64_t> kp((size_t)n);
      for (int64_t i = 0; i < n; ++i) kp[i] = ((uint64_t)k[i] << 32) | (uint64_t)i;
      if (take < n) std::nth_element(kp.begin(), kp.begin() + take, kp.end());
      std::sort(kp.begin(), kp.begin() + take);
      std::vector<int32_t> out((size_t)take);
      for (int64_t i = 0; i < take; ++i) out[i] = x[(size_t)(kp[i] & 0xFFFFFFFFu)];
      return out;
    }
    for (int shift = 0; shift < 32; shift += 8) {
      size_t cnt[257] = {0};
      for (int64_t i = 0; i < n; ++i) ++cnt[((k[i] >> shift) & 255u) + 1];
      for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
      for (int64_t i = 0; i < n; ++i) {
        const size_t p = cnt[(k[i] >> shift) & 255u]++;
        k2[p] = k[i];
        x2[p] = x[i];
      }
      k.swap(k2);
      x.swap(x2);
    }
  }
  x.resize((size_t)take);
  return x;
}

void Server::prepare(Req *r) {
  const abm_request &q = r->q;
  const int mode = cfg_.mode;
  if (q.n < 2 || q.n > cfg_.max_agents || q.steps < 1 || q.steps > 100000 ||
      (q.kind != kSir && q.kind != kWealth) ||
      (q.kind == kSir && (q.m < 1 || q.n_infected < 0 || !(q.beta >= 0.0f) || !(q.gamma >= 0.0f)))) {
    fail(r, "invalid request");
    return;
  }
  const int64_t n = q.n, steps = q.steps;
  Key rng, ik;
  split2(Key{0u, (uint32_t)(q.seed & 0xFFFFFFFFll)}, rng, ik, mode);
  r->ik[0] = ik.k0;
  r->ik[1] = ik.k1;
  uint32_t rk[2] = {rng.k0, rng.k1};
  const bool sir = q.kind == kSir;
  std::vector<uint32_t> tab((size_t)steps * (sir ? 6 : 4));
  abm_step_keys(rk, (int32_t)steps, 1, 1, sir ? 1 : 0, mode, tab.data());
  std::vector<int32_t> row_ptr, col, infected;
  size_t off = 0;
  auto place = [&off](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  r->off_keys = place(16 * (size_t)steps);
  const int64_t E = sir ? abm_ba_num_edges(n, q.m) : 0;
  if (sir) {
    r->nnz = 2 * E;
    r->n0 = (int32_t)std::min<int64_t>(q.n_infected, n);
    r->off_rowptr = place(4 * ((size_t)n + 1));
    r->off_col = place(4 * (size_t)r->nnz);
    r->off_infected = place(4 * (size_t)std::max<int32_t>(r->n0, 1));
    r->upload_bytes = off;
    r->off_state0 = place((size_t)n);
    r->off_state1 = place((size_t)n);
    r->off_mark = place(2 * (size_t)n);
    r->off_counts = place(12 * (size_t)steps);
    r->result_off = r->off_counts;
    r->result_bytes = 12 * (size_t)steps;
    r->resident = n <= std::min<int64_t>(cfg_.sir_resident_max, kResidentMax);
  } else {
    r->upload_bytes = off;
    r->off_x = place(4 * (size_t)n);
    r->off_y = place(4 * (size_t)n);
    r->off_w = place(4 * (size_t)n);
    r->off_inc = place(8 * (size_t)n);           // two copies (step parity)
    r->off_hist = place(8 * ((size_t)n + 1));
    r->off_acc = place(16);
    r->off_mean = place(8 * (size_t)steps);  // mean[steps] then gini[steps]
    r->off_gini = r->off_mean + 4 * (size_t)steps;
    r->result_off = r->off_mean;
    r->result_bytes = 8 * (size_t)steps;
    r->resident = n <= std::min<int64_t>(cfg_.wealth_resident_max, kResidentMax);
    if (!r->resident) {
      const int64_t chunk = std::max<int64_t>(cfg_.wealth_chunk, 64);
      for (int64_t b = 0; b < n; b += chunk) r->ranges.push_back((int32_t)b);
      r->ranges.push_back((int32_t)n);
    }
  }
  r->dev_bytes = off;
  if (r->dev_bytes > cfg_.hbm_budget) {
    fail(r, "request exceeds the HBM budget");
    return;
  }
  if (sir) {
    const uint32_t gk[2] = {0u, (uint32_t)(q.graph_seed & 0xFFFFFFFFll)};
    std::vector<int32_t> src((size_t)E), dst((size_t)E);
    abm_ba_edges(n, q.m, gk, src.data(), dst.data());
    row_ptr.resize((size_t)n + 1);
    col.resize((size_t)(2 * E));
    if (abm_csr_undirected(n, E, src.data(), dst.data(), row_ptr.data(), col.data()) != 0) {
      fail(r, "graph construction failed");
      return;
    }
    infected = host_permutation_prefix(ik, n, r->n0, mode);
    if (!r->resident) {
      const int64_t target = std::max<int64_t>(cfg_.sir_chunk, 64);
      std::vector<int32_t> cp((size_t)n + 1);
      const int64_t nc = abm_node_chunks(n, row_ptr.data(), 64, 4, kHeavyDeg, 128, target,
                                         1 << 20, cp.data());
      cp.resize((size_t)nc + 1);
      r->ranges = std::move(cp);
    }
  }
  if (pool_.get(r->upload_bytes, &r->staging) != hipSuccess || !(r->up_ev = make_event())) {
    fail(r, "pinned staging allocation failed");
    return;
  }
  char *h = (char *)r->staging.p;
  uint32_t *keys = (uint32_t *)(h + r->off_keys);
  for (int64_t t = 0; t < steps; ++t) {
    const uint32_t *row = tab.data() + (size_t)t * (sir ? 6 : 4);
    std::memcpy(keys + 4 * t, row + (sir ? 2 : 0), 16);
  }
  if (sir) {
    std::memcpy(h + r->off_rowptr, row_ptr.data(), 4 * row_ptr.size());
    std::memcpy(h + r->off_col, col.data(), 4 * col.size());
    if (r->n0) std::memcpy(h + r->off_infected, infected.data(), 4 * infected.size());
  }
  {
    std::lock_guard<std::mutex> g(mu_);
    r->status = kPrepared;
    r->t_ready = now_s();
    ready_q_.push_back(r);
  }
  cv_sched_.notify_one();
}

void Server::admit() {
  std::vector<Req *> picked;
  {
    std::lock_guard<std::mutex> g(mu_);
    const size_t slots = (size_t)cfg_.max_active + (size_t)cfg_.prefetch;
    size_t look = 0;
    for (auto it = ready_q_.begin();
         it != ready_q_.end() && look < 16 &&
         active_.size() + uploading_.size() + picked.size() < slots;
         ++look) {
      Req *r = *it;
      size_t off;
      if (!arena_.alloc(r->dev_bytes, &off)) {
        ++it;
        if (look == 0 && cfg_.strict_fifo) break;
        continue;
      }
      r->dev_off = off;
      r->status = kUploading;
      r->t_admit = now_s();
      picked.push_back(r);
      it = ready_q_.erase(it);
    }
  }
  peak_bytes_ = std::max<int64_t>(peak_bytes_, (int64_t)arena_.used());
  for (Req *r : picked) {
    char *base = dev_base_ + r->dev_off;
    Inst &d = r->d;
    d = Inst{};
    d.kind = r->q.kind;
    d.steps = r->q.steps;
    d.n = (int32_t)r->q.n;
    d.keys = (const uint32_t *)(base + r->off_keys);
    if (d.kind == kSir) {
      d.nnz = r->nnz;
      d.thr_beta = unit_threshold(r->q.beta);
      d.thr_gamma = unit_threshold(r->q.gamma);
      d.row_ptr = (const int32_t *)(base + r->off_rowptr);
      d.col = (const int32_t *)(base + r->off_col);
      d.state0 = (uint8_t *)(base + r->off_state0);
      d.state1 = (uint8_t *)(base + r->off_state1);
      d.mark = (uint16_t *)(base + r->off_mark);
      d.counts = (int32_t *)(base + r->off_counts);
    } else {
      d.x = (float *)(base + r->off_x);
      d.y = (float *)(base + r->off_y);
      d.w = (int32_t *)(base + r->off_w);
      d.inc = (int32_t *)(base + r->off_inc);
      d.hist = (int32_t *)(base + r->off_hist);
      d.acc = (int32_t *)(base + r->off_acc);
      d.mean = (float *)(base + r->off_mean);
      d.gini = (float *)(base + r->off_gini);
    }
    hipError_t e = hipMemcpyAsync(base, r->staging.p, r->upload_bytes, hipMemcpyHostToDevice, sC_);
    if (e == hipSuccess) e = hipEventRecord(r->up_ev, sC_);
    if (e != hipSuccess) {
      arena_.release(r->dev_off, r->dev_bytes);
      fail(r, std::string("upload failed: ") + hipGetErrorString(e));
      continue;
    }
    uploading_.push_back(r);
  }
}

void Server::activate_uploads() {
  for (size_t i = 0; i < uploading_.size();) {
    Req *r = uploading_[i];
    if (hipEventQuery(r->up_ev) != hipSuccess) {
      ++i;
      continue;
    }
    if ((int)active_.size() >= cfg_.max_active) break;
    pool_.put(r->staging);
    r->staging = Pinned{};
    const Inst &d = r->d;
    if (d.kind == kSir) {
      const int64_t len = std::max<int64_t>(d.n, 3 * (int64_t)d.steps);
      hipLaunchKernelGGL(k_serve_init_sir, dim3((unsigned)((len + 255) / 256)), dim3(256), 0,
                         sW_, d);
      if (r->n0)
        hipLaunchKernelGGL(k_serve_seed_sir, dim3((unsigned)((r->n0 + 255) / 256)), dim3(256),
                           0, sW_, d,
                           (const int32_t *)(dev_base_ + r->dev_off + r->off_infected), r->n0);
    } else {
      hipLaunchKernelGGL(k_serve_init_wealth, dim3((unsigned)((d.n + 256) / 256)), dim3(256), 0,
                         sW_, d, r->ik[0], r->ik[1], cfg_.mode);
    }
    r->t0 = 0;
    {
      std::lock_guard<std::mutex> g(mu_);
      r->status = kActive;
      r->t_start = now_s();
    }
    active_.push_back(r);
    uploading_.erase(uploading_.begin() + (long)i);
  }
  peak_active_ = std::max<int64_t>(peak_active_, (int64_t)active_.size());
}

int Server::launch_tick() {
  const int32_t K = cfg_.steps_per_tick;
  size_t n1 = 0;
  for (Req *r : active_)
    if (!r->resident) n1 += r->ranges.size() - 1;
  if (int rc = ensure_tables(active_.size(), n1)) return rc;
  TableSet &ts = sets_[tick_idx_ % kSets];
  const size_t o1 = align256(sizeof(Inst) * active_.size());
  const size_t o2 = o1 + align256(sizeof(Range) * n1);
  const size_t used = o2 + align256(4 * active_.size());
  char *h = (char *)ts.host;
  Inst *hi = (Inst *)h;
  Range *h1 = (Range *)(h + o1);
  int32_t *hres = (int32_t *)(h + o2);
  const char *dv = (const char *)ts.dev;
  const Inst *di = (const Inst *)dv;
  const Range *d1 = (const Range *)(dv + o1);
  const int32_t *dres = (const int32_t *)(dv + o2);
  int32_t nres = 0;
  size_t i1 = 0;
  Tick tick;
  for (size_t i = 0; i < active_.size(); ++i) {
    Req *r = active_[i];
    hi[i] = r->d;
    hi[i].t0 = r->t0;
    if (r->resident) {
      hres[nres++] = (int32_t)i;
    } else {
      for (size_t c = 0; c + 1 < r->ranges.size(); ++c)
        h1[i1++] = Range{(int32_t)i, r->ranges[c], r->ranges[c + 1], r->d.kind};
    }
    const int32_t run = std::max<int32_t>(0, std::min<int32_t>(K, r->d.steps - r->t0));
    agent_steps_ += (int64_t)run * r->d.n;
    r->t0 += K;
    // a stepwise wealth instance applies and records its last step in the
    // two sub-steps after it
    const int32_t end = r->d.steps + (!r->resident && r->d.kind == kWealth ? 2 : 0);
    if (r->t0 >= end) tick.finishing.push_back(r);
  }
  ABM_CHECK_HIP(hipMemcpyAsync(ts.dev, ts.host, used, hipMemcpyHostToDevice, sW_));
  ABM_CHECK_HIP(hipEventRecord(ev_tab_, sW_));
  ABM_CHECK_HIP(hipStreamWaitEvent(sR_, ev_tab_, 0));
  if (nres)
    hipLaunchKernelGGL(k_serve_resident, dim3((unsigned)nres), dim3(kResThreads), kResLds, sR_,
                       di, dres, K, cfg_.mode);
  for (int32_t s = 0; i1 && s < K; ++s)
    hipLaunchKernelGGL(k_serve_phase1, dim3((unsigned)i1), dim3(kStepThreads), 0, sW_, di, d1, s,
                       cfg_.mode);
  ABM_LAUNCH_CHECK("serve tick");
  auto take = [this]() {
    if (!free_events_.empty()) {
      hipEvent_t e = free_events_.back();
      free_events_.pop_back();
      return e;
    }
    return make_event();
  };
  tick.ev_r = take();
  tick.ev_w = take();
  if (!tick.ev_r || !tick.ev_w) return hip_fail(hipErrorOutOfMemory, "serve tick events");
  ABM_CHECK_HIP(hipEventRecord(tick.ev_r, sR_));
  ABM_CHECK_HIP(hipEventRecord(tick.ev_w, sW_));
  for (Req *r : tick.finishing)
    active_.erase(std::find(active_.begin(), active_.end(), r));
  inflight_.push_back(std::move(tick));
  ++tick_idx_;
  ++n_ticks_;
  return kOk;
}

bool Server::retire_tick(bool block) {
  if (inflight_.empty()) return false;
  Tick &t = inflight_.front();
  if (block) {
    (void)hipEventSynchronize(t.ev_r);
    (void)hipEventSynchronize(t.ev_w);
  } else if (hipEventQuery(t.ev_r) != hipSuccess || hipEventQuery(t.ev_w) != hipSuccess) {
    return false;
  }
  for (Req *r : t.finishing) {
    std::vector<uint8_t> out(r->result_bytes);
    hipError_t e = hipMemcpyAsync(out.data(), dev_base_ + r->dev_off + r->result_off,
                                  r->result_bytes, hipMemcpyDeviceToHost, sD_);
    if (e == hipSuccess) e = hipStreamSynchronize(sD_);
    arena_.release(r->dev_off, r->dev_bytes);
    if (e != hipSuccess) {
      fail(r, std::string("result copy failed: ") + hipGetErrorString(e));
      continue;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      r->result = std::move(out);
      r->status = kDone;
      r->t_done = now_s();
      --outstanding_;
    }
    ++n_done_;
  }
  free_events_.push_back(t.ev_r);
  free_events_.push_back(t.ev_w);
  inflight_.pop_front();
  cv_done_.notify_all();
  return true;
}

bool Server::sched_once() {
  bool progress = false;
  while (retire_tick(false)) progress = true;
  const size_t before = uploading_.size() + active_.size();
  admit();
  activate_uploads();
  if (uploading_.size() + active_.size() != before) progress = true;
  if (!active_.empty() && (int)inflight_.size() < cfg_.max_inflight) {
    const int rc = launch_tick();
    if (rc) {
      std::lock_guard<std::mutex> g(mu_);
      sched_error_ = abm_last_error();
      stop_ = true;
      return false;
    }
    progress = true;
  }
  return progress;
}

void Server::sched_loop() {
  if (hipSetDevice(device_) != hipSuccess) return;
  for (;;) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (stop_) break;
    }
    if (sched_once()) continue;
    if (!inflight_.empty() && (int)inflight_.size() >= cfg_.max_inflight) {
      retire_tick(true);
      continue;
    }
    if (!inflight_.empty() || !uploading_.empty()) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      continue;
    }
    std::unique_lock<std::mutex> lk(mu_);
    cv_sched_.wait_for(lk, std::chrono::milliseconds(2),
                       [this] { return stop_ || !ready_q_.empty(); });
  }
  while (!inflight_.empty()) retire_tick(true);
  // whatever did not finish fails
  std::lock_guard<std::mutex> g(mu_);
  for (auto &kv : reqs_) {
    Req *r = kv.second.get();
    if (r->status != kDone && r->status != kFailed) {
      r->status = kFailed;
      r->error = sched_error_.empty() ? "server stopped" : sched_error_;
      --outstanding_;
    }
  }
  cv_done_.notify_all();
}

int Server::poll(int64_t ticket, int32_t *status) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) {
    set_error("abm_serve_poll: unknown ticket");
    return kErrInvalid;
  }
  *status = it->second->status;
  return kOk;
}

int Server::wait(int64_t ticket, double timeout, int32_t *status) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) {
    set_error("abm_serve_wait: unknown ticket");
    return kErrInvalid;
  }
  Req *r = it->second.get();
  cv_done_.wait_for(lk, std::chrono::duration<double>(timeout),
                    [r] { return r->status == kDone || r->status == kFailed; });
  *status = r->status;
  return kOk;
}

int Server::collect(int64_t ticket, void *out, size_t bytes, abm_ticket_info *info) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) {
    set_error("abm_serve_collect: unknown ticket");
    return kErrInvalid;
  }
  Req *r = it->second.get();
  if (info) {
    *info = abm_ticket_info{};
    info->status = r->status;
    info->kind = r->q.kind;
    info->steps = r->q.steps;
    info->resident = r->resident ? 1 : 0;
    info->n = r->q.n;
    info->nnz = r->nnz;
    info->device_bytes = (int64_t)r->dev_bytes;
    info->result_bytes = (int64_t)r->result_bytes;
    info->t_submit = r->t_submit;
    info->t_ready = r->t_ready;
    info->t_admit = r->t_admit;
    info->t_start = r->t_start;
    info->t_done = r->t_done;
  }
  if (r->status == kFailed) {
    set_error("abm_serve_collect: request failed: " + r->error);
    return kErrInvalid;
  }
  if (!out) return kOk;
  if (r->status != kDone) {
    set_error("abm_serve_collect: request not done");
    return kErrInvalid;
  }
  if (bytes < r->result.size()) {
    set_error("abm_serve_collect: output buffer too small");
    return kErrInvalid;
  }
  std::memcpy(out, r->result.data(), r->result.size());
  return kOk;
}

int Server::drain(double timeout) {
  std::unique_lock<std::mutex> lk(mu_);
  const bool ok = cv_done_.wait_for(lk, std::chrono::duration<double>(timeout),
                                    [this] { return outstanding_ == 0; });
  if (!ok) {
    set_error("abm_serve_drain: timed out");
    return kErrInvalid;
  }
  if (!sched_error_.empty()) {
    set_error("abm_serve_drain: " + sched_error_);
    return kErrInvalid;
  }
  return kOk;
}

void Server::stats(abm_serve_stats *st) {
  *st = abm_serve_stats{};
  st->ticks = n_ticks_;
  st->completed = n_done_;
  st->agent_steps = agent_steps_;
  st->peak_active = peak_active_;
  st->peak_device_bytes = peak_bytes_;
  std::lock_guard<std::mutex> g(mu_);
  st->outstanding = outstanding_;
}

}  // namespace serve
}  // namespace abm

using namespace abm;
using abm::serve::Server;

extern "C" {

double abm_serve_clock(void) { return serve::now_s(); }

void abm_serve_default_config(abm_serve_config *cfg) {
  if (!cfg) return;
  *cfg = abm_serve_config{};
  cfg->hbm_budget = (uint64_t)8 << 30;
  cfg->max_agents = 1 << 22;
  cfg->max_active = 64;
  cfg->prefetch = 2;
  cfg->max_inflight = 2;
  cfg->steps_per_tick = 16;
  cfg->prep_threads = 12;
  cfg->mode = kModePartitionable;
  cfg->sir_chunk = 16384;
  cfg->wealth_chunk = 2048;
  cfg->sir_resident_max = 8192;
  cfg->wealth_resident_max = 2048;
  cfg->strict_fifo = 0;
}

int abm_serve_create(const abm_serve_config *cfg, void **server) {
  ABM_REQUIRE(cfg && server, "abm_serve_create: null argument");
  ABM_REQUIRE(cfg->hbm_budget >= ((uint64_t)1 << 20) && cfg->max_active >= 1 &&
                  cfg->max_active <= 4096 && cfg->prefetch >= 0 && cfg->max_inflight >= 1 &&
                  cfg->max_inflight <= 3 && cfg->steps_per_tick >= 1 &&
                  cfg->steps_per_tick <= 4096 && cfg->prep_threads >= 1 &&
                  cfg->prep_threads <= 64 && (cfg->mode == 0 || cfg->mode == 1) &&
                  cfg->max_agents >= 2 && cfg->max_agents < (1ll << 30),
              "abm_serve_create: bad configuration");
  *server = nullptr;
  auto s = std::make_unique<Server>(*cfg);
  if (int rc = s->start()) return rc;
  *server = s.release();
  return kOk;
}

int abm_serve_submit(void *server, const abm_request *req, int64_t *ticket) {
  ABM_REQUIRE(server && req && ticket, "abm_serve_submit: null argument");
  return ((Server *)server)->submit(*req, ticket);
}

int abm_serve_poll(void *server, int64_t ticket, int32_t *status) {
  ABM_REQUIRE(server && status, "abm_serve_poll: null argument");
  return ((Server *)server)->poll(ticket, status);
}

int abm_serve_wait(void *server, int64_t ticket, double timeout_s, int32_t *status) {
  ABM_REQUIRE(server && status && timeout_s >= 0, "abm_serve_wait: bad argument");
  return ((Server *)server)->wait(ticket, timeout_s, status);
}

int abm_serve_collect(void *server, int64_t ticket, void *out, size_t out_bytes,
                      abm_ticket_info *info) {
  ABM_REQUIRE(server, "abm_serve_collect: null server");
  return ((Server *)server)->collect(ticket, out, out_bytes, info);
}

int abm_serve_drain(void *server, double timeout_s) {
  ABM_REQUIRE(server && timeout_s >= 0, "abm_serve_drain: bad argument");
  return ((Server *)server)->drain(timeout_s);
}

int abm_serve_get_stats(void *server, abm_serve_stats *st) {
  ABM_REQUIRE(server && st, "abm_serve_get_stats: null argument");
  ((Server *)server)->stats(st);
  return kOk;
}

int abm_serve_destroy(void *server) {
  delete (Server *)server;
  return kOk;
}

}  // extern "C"
