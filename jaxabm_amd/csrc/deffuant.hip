// deffuant.hip -- differentiable Deffuant bounded-confidence model: batched
// forward step, final statistics and loss, and the hand-written adjoint with
// checkpointed recomputation.
//
// Program (reference API: one collection, env-phase pairing with the update
// key, jaxabm/model.py:146-216; loss and gradient set-up
// jaxabm/analysis.py:489-531):
//  init   agent keys = split(init_key, n); x_i = uniform(agent_key_i), the
//         same value for every instance of the batch
//  step   perm = permutation(uk, n), shared by the batch (common random
//         numbers); pairs (perm[2i], perm[2i+1]); d = x_c - x_a;
//         g = 1 / (1 + exp(-(eps - |d|) / 0.02)); delta = (mu g) d;
//         x_a += delta; x_c -= delta (an odd agent out keeps its opinion)
//  loss   stats = [mean, variance, 32-bin soft histogram with edges k/32,
//         bin k = mean sigmoid(128 (x - k/32)) - sigmoid(128 (x - (k+1)/32))];
//         L = sum (stats - target)^2 per instance, in float64
//
// Layout (HBM): x[n][B], agent-major and instance-minor, so a pair reads and
// writes two contiguous rows of B floats: every access is a coalesced row
// segment whatever the pairing.  In the step kernels the 64 lanes of a
// wavefront span VEC-wide column groups and the waves span pairs.
//
// Adjoint of one step, per pair and instance (lam = dL/dx after the step):
//   g' = g (1 - g) / 0.02
//   fd = d delta/dd   = mu (g - |d| g')     w = lam_a - lam_c
//   fe = d delta/deps = mu d g'             lam_a -= w fd, lam_c += w fd
//   fm = d delta/dmu  = g d                 dL/deps += w fe, dL/dmu += w fm
// The parameter sums are float64; each workgroup row of the grid owns one
// accumulator slot for the whole sweep (same slot every step), and the slots
// are summed in a fixed order, so the gradients are deterministic.
// Checkpoints: the forward keeps the state at every ckpt_every-th step; the
// backward recomputes each segment's states from its checkpoint, then sweeps
// the segment in reverse.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr float kGateScale = 0.02f;
constexpr int kDefThreads = 256;
constexpr int kDefWaves = kDefThreads / 64;
constexpr int kStatLen = 2 + 32;
constexpr int kHistBins = 32;
constexpr double kHistSlope = 128.0;

template <int VEC>
__device__ __forceinline__ void load_cols(const float *p, float (&v)[VEC]) {
  if constexpr (VEC == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  } else {
    v[0] = *p;
  }
}

template <int VEC>
__device__ __forceinline__ void store_cols(float *p, const float (&v)[VEC]) {
  if constexpr (VEC == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    *p = v[0];
  }
}

__device__ __forceinline__ float gate(float eps, float d) {
  const float z = (eps - fabsf(d)) / kGateScale;
  const float e = abm_expf(-z);
  const float den = 1.0f + e;
  return 1.0f / den;
}

// Pairs each wavefront holds in flight per loop iteration of the step kernel:
// eight 1 KB row loads outstanding per wave (measured: 4 beats 2 and 8).
constexpr int kDefStepPairs = 4;

// One step for the whole batch.  src == dst updates in place (pairs are
// disjoint); otherwise every row of dst is written (the odd agent is copied).
template <int VEC>
__global__ void __launch_bounds__(kDefThreads)
    k_deffuant_step(const float *src, float *dst, int64_t n, int32_t B,
                    const int32_t *__restrict__ perm, const float *__restrict__ eps,
                    const float *__restrict__ mu) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = blockIdx.x * 64 + lane;
  if (q >= B / VEC) return;  // no barriers below
  const int col = q * VEC;
  float e[VEC], m[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    e[v] = eps[col + v];
    m[v] = mu[col + v];
  }
  const int64_t npairs = n >> 1;
  const int64_t stride = (int64_t)gridDim.y * kDefWaves;
  for (int64_t p0 = (int64_t)blockIdx.y * kDefWaves + wave; p0 < npairs;
       p0 += kDefStepPairs * stride) {
    // pairs p0, p0 + stride, ...; one past the end re-reads pair p0 and
    // stores nothing
    int64_t a[kDefStepPairs], c[kDefStepPairs];
    bool live[kDefStepPairs];
#pragma unroll
    for (int i = 0; i < kDefStepPairs; ++i) {
      const int64_t p = p0 + i * stride;
      live[i] = p < npairs;
      const int64_t pp = live[i] ? p : p0;
      a[i] = perm[2 * pp];
      c[i] = perm[2 * pp + 1];
    }
    float xa[kDefStepPairs][VEC], xc[kDefStepPairs][VEC];
#pragma unroll
    for (int i = 0; i < kDefStepPairs; ++i) {
      load_cols<VEC>(src + a[i] * B + col, xa[i]);
      load_cols<VEC>(src + c[i] * B + col, xc[i]);
    }
#pragma unroll
    for (int i = 0; i < kDefStepPairs; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const float d = xc[i][v] - xa[i][v];
        const float delta = (m[v] * gate(e[v], d)) * d;
        xa[i][v] = xa[i][v] + delta;
        xc[i][v] = xc[i][v] - delta;
      }
#pragma unroll
    for (int i = 0; i < kDefStepPairs; ++i)
      if (live[i]) {
        store_cols<VEC>(dst + a[i] * B + col, xa[i]);
        store_cols<VEC>(dst + c[i] * B + col, xc[i]);
      }
  }
  if ((n & 1) && src != dst && blockIdx.y == 0 && wave == 0) {
    const int64_t r = perm[n - 1];
    float x[VEC];
    load_cols<VEC>(src + r * B + col, x);
    store_cols<VEC>(dst + r * B + col, x);
  }
}

// Reverse sweep of one step: x_pre is the state before the step, lam holds
// dL/dx after it and is updated in place to dL/dx before it.
template <int VEC>
__global__ void __launch_bounds__(kDefThreads)
    k_deffuant_adjoint(const float *__restrict__ x_pre, float *__restrict__ lam, int64_t n,
                       int32_t B, const int32_t *__restrict__ perm,
                       const float *__restrict__ eps, const float *__restrict__ mu,
                       double *__restrict__ acc) {
  __shared__ double red[2][kDefWaves][64 * VEC];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = blockIdx.x * 64 + lane;
  const bool active = q < B / VEC;
  const int col = q * VEC;
  double ge[VEC], gm[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) ge[v] = gm[v] = 0.0;
  if (active) {
    float e[VEC], m[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      e[v] = eps[col + v];
      m[v] = mu[col + v];
    }
    const int64_t npairs = n >> 1;
    const int64_t stride = (int64_t)gridDim.y * kDefWaves;
    for (int64_t p = (int64_t)blockIdx.y * kDefWaves + wave; p < npairs; p += stride) {
      const int64_t a = perm[2 * p], c = perm[2 * p + 1];
      float xa[VEC], xc[VEC], la[VEC], lc[VEC];
      load_cols<VEC>(x_pre + a * B + col, xa);
      load_cols<VEC>(x_pre + c * B + col, xc);
      load_cols<VEC>(lam + a * B + col, la);
      load_cols<VEC>(lam + c * B + col, lc);
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const float d = xc[v] - xa[v];
        const float g = gate(e[v], d);
        const float gp = g * (1.0f - g) / kGateScale;
        const float fd = m[v] * (g - fabsf(d) * gp);
        const float fe = m[v] * d * gp;
        const float fm = g * d;
        const float w = la[v] - lc[v];
        la[v] = la[v] - w * fd;
        lc[v] = lc[v] + w * fd;
        ge[v] += (double)(w * fe);
        gm[v] += (double)(w * fm);
      }
      store_cols<VEC>(lam + a * B + col, la);
      store_cols<VEC>(lam + c * B + col, lc);
    }
  }
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    red[0][wave][lane * VEC + v] = ge[v];
    red[1][wave][lane * VEC + v] = gm[v];
  }
  __syncthreads();
  if (wave == 0 && active) {
    double *slot = acc + (int64_t)blockIdx.y * 2 * B;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      double se = 0.0, sm = 0.0;
      for (int w = 0; w < kDefWaves; ++w) {
        se += red[0][w][lane * VEC + v];
        sm += red[1][w][lane * VEC + v];
      }
      slot[col + v] += se;
      slot[B + col + v] += sm;
    }
  }
}

// 1 / (1 + t) for t >= 0 in float64: the hardware reciprocal estimate
// (v_rcp_f64) refined by two Newton-Raphson steps r += r (1 - d r), a few
// FMAs instead of the IEEE division sequence; accurate to float64 rounding
// (a few ulps), far inside the statistics' 1e-10 tolerance.
__device__ __forceinline__ double inv_one_plus(double t) {
  const double d = 1.0 + t;
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// Per-chunk float64 partial sums of one instance column:
// [sum x, sum (x - 1/2)^2, hist_0 .. hist_31] -> partial[chunk][34][B].
// sigmoid(128 (x - k/32)) = 1 / (1 + t e^(4k)) with t = e^(-128 x).
__global__ void __launch_bounds__(kDefThreads)
    k_deffuant_stats(const float *__restrict__ x, int64_t n, int32_t B, int64_t rows,
                     double *__restrict__ partial) {
  const int b = blockIdx.x * kDefThreads + threadIdx.x;
  if (b >= B) return;
  const int64_t i0 = (int64_t)blockIdx.y * rows;
  const int64_t i1 = i0 + rows < n ? i0 + rows : n;
  const double e4 = exp(4.0);
  double s1 = 0.0, s2 = 0.0, h[kHistBins];
#pragma unroll
  for (int k = 0; k < kHistBins; ++k) h[k] = 0.0;
  for (int64_t i = i0; i < i1; ++i) {
    const double xi = (double)x[i * B + b];
    s1 += xi;
    s2 += (xi - 0.5) * (xi - 0.5);
    double t = exp(-kHistSlope * xi);
    double prev = inv_one_plus(t);
#pragma unroll
    for (int k = 0; k < kHistBins; ++k) {
      t *= e4;
      const double cur = inv_one_plus(t);
      h[k] += prev - cur;
      prev = cur;
    }
  }
  double *out = partial + (int64_t)blockIdx.y * kStatLen * B + b;
  out[0] = s1;
  out[B] = s2;
#pragma unroll
  for (int k = 0; k < kHistBins; ++k) out[(int64_t)(2 + k) * B] = h[k];
}

// out[j][b] = sum over r of in[r][j][b], for j = blockIdx.y; 64 columns per
// block, 4 row groups combined in a fixed order.
__global__ void __launch_bounds__(256)
    k_sum_rows(const double *__restrict__ in, int64_t rows, int32_t J, int32_t B,
               double *__restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int b = blockIdx.x * 64 + lane;
  const int j = blockIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (b < B) {
    int64_t r = g;
    for (; r + 4 < rows; r += 8) {
      s0 += in[(r * J + j) * B + b];
      s1 += in[((r + 4) * J + j) * B + b];
    }
    for (; r < rows; r += 4) s0 += in[(r * J + j) * B + b];
  }
  red[g][lane] = s0 + s1;
  __syncthreads();
  if (g == 0 && b < B) out[(int64_t)j * B + b] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// stats[b][34], loss[b] and the coefficients of dL/dx (coef[34][B]:
// dL/dmean / n, dL/dvar * 2 / n, dL/dhist_k / n) from the column sums.
// weights (optional, [34]): L = sum_j w_j (stats_j - target_j)^2.
__global__ void k_deffuant_loss(const double *__restrict__ sums, int64_t n, int32_t B,
                                const double *__restrict__ target,
                                const double *__restrict__ weights, double *__restrict__ stats,
                                double *__restrict__ loss, double *__restrict__ coef) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double dn = (double)n;
  const double m = sums[b] / dn;
  const double v = sums[B + b] / dn - (m - 0.5) * (m - 0.5);
  const double w0 = weights ? weights[0] : 1.0, w1 = weights ? weights[1] : 1.0;
  double L = w0 * (m - target[0]) * (m - target[0]) + w1 * (v - target[1]) * (v - target[1]);
  stats[(int64_t)b * kStatLen] = m;
  stats[(int64_t)b * kStatLen + 1] = v;
  coef[b] = w0 * 2.0 * (m - target[0]) / dn;
  coef[B + b] = w1 * 2.0 * (v - target[1]) * 2.0 / dn;
  for (int k = 0; k < kHistBins; ++k) {
    const double hk = sums[(int64_t)(2 + k) * B + b] / dn;
    const double wk = weights ? weights[2 + k] : 1.0;
    stats[(int64_t)b * kStatLen + 2 + k] = hk;
    L += wk * (hk - target[2 + k]) * (hk - target[2 + k]);
    coef[(int64_t)(2 + k) * B + b] = wk * 2.0 * (hk - target[2 + k]) / dn;
  }
  loss[b] = L;
}

// lam[i][b] = dL/dx_i for instance b from the final state.
__global__ void k_deffuant_lambda(const float *__restrict__ x, int64_t n, int32_t B,
                                  const double *__restrict__ stats,
                                  const double *__restrict__ coef, float *__restrict__ lam) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * B) return;
  const int b = (int)(idx % B);
  const double xi = (double)x[idx];
  const double e4 = exp(4.0);
  double g = coef[b] + coef[B + b] * (xi - stats[(int64_t)b * kStatLen]);
  double t = exp(-kHistSlope * xi);
  double s = inv_one_plus(t);
  double ds_prev = s * (1.0 - s);
  for (int k = 0; k < kHistBins; ++k) {
    t *= e4;
    s = inv_one_plus(t);
    const double ds = s * (1.0 - s);
    g += coef[(int64_t)(2 + k) * B + b] * kHistSlope * (ds_prev - ds);
    ds_prev = ds;
  }
  lam[idx] = (float)g;
}

__global__ void k_deffuant_init(const uint32_t *__restrict__ init_key, int64_t n, int32_t B,
                                float *__restrict__ x, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = unit_float(bits_scalar(split_at(Key{init_key[0], init_key[1]}, n, i, mode), mode));
  for (int32_t b = 0; b < B; ++b) x[i * B + b] = v;
}

__global__ void k_fill_f64(double *__restrict__ p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// VEC = 4 when every row segment is 16-byte aligned.
inline bool use_vec4(int32_t B, const void *a, const void *b) {
  return (B % 4) == 0 && aligned16(a) && aligned16(b);
}

void launch_step(const float *src, float *dst, int64_t n, int32_t B, const int32_t *perm,
                 const float *eps, const float *mu, int32_t pair_blocks, hipStream_t s) {
  if (use_vec4(B, src, dst)) {
    const dim3 g(nblocks(B / 4, 64), (unsigned)pair_blocks);
    hipLaunchKernelGGL(k_deffuant_step<4>, g, dim3(kDefThreads), 0, s, src, dst, n, B, perm,
                       eps, mu);
  } else {
    const dim3 g(nblocks(B, 64), (unsigned)pair_blocks);
    hipLaunchKernelGGL(k_deffuant_step<1>, g, dim3(kDefThreads), 0, s, src, dst, n, B, perm,
                       eps, mu);
  }
}

void launch_adjoint(const float *x_pre, float *lam, int64_t n, int32_t B, const int32_t *perm,
                    const float *eps, const float *mu, int32_t pair_blocks, double *acc,
                    hipStream_t s) {
  if (use_vec4(B, x_pre, lam)) {
    const dim3 g(nblocks(B / 4, 64), (unsigned)pair_blocks);
    hipLaunchKernelGGL(k_deffuant_adjoint<4>, g, dim3(kDefThreads), 0, s, x_pre, lam, n, B,
                       perm, eps, mu, acc);
  } else {
    const dim3 g(nblocks(B, 64), (unsigned)pair_blocks);
    hipLaunchKernelGGL(k_deffuant_adjoint<1>, g, dim3(kDefThreads), 0, s, x_pre, lam, n, B,
                       perm, eps, mu, acc);
  }
}

void launch_loss(const abm_deffuant_problem *p, const float *x, hipStream_t s) {
  const int64_t rows = (p->n + p->stat_chunks - 1) / p->stat_chunks;
  const int64_t chunks = (p->n + rows - 1) / rows;
  hipLaunchKernelGGL(k_deffuant_stats, dim3(nblocks(p->batch, kDefThreads), (unsigned)chunks),
                     dim3(kDefThreads), 0, s, x, p->n, p->batch, rows, p->partial);
  hipLaunchKernelGGL(k_sum_rows, dim3(nblocks(p->batch, 64), (unsigned)kStatLen), dim3(256), 0,
                     s, (const double *)p->partial, chunks, kStatLen, p->batch, p->sums);
  hipLaunchKernelGGL(k_deffuant_loss, dim3(nblocks(p->batch, 256)), dim3(256), 0, s,
                     (const double *)p->sums, p->n, p->batch, p->target, p->weights, p->stats,
                     p->loss, p->coef);
}

int check_problem(const abm_deffuant_problem *p, const char *who, bool backward) {
  bool ok = p && p->n >= 2 && p->n < (1ll << 31) && p->batch >= 1 && p->steps >= 0 &&
            p->ckpt_every >= 1 && p->pair_blocks >= 1 && p->pair_blocks <= 65535 &&
            p->stat_chunks >= 1 && p->stat_chunks <= 65535 && p->n * (int64_t)p->batch < (1ll << 40) &&
            p->perms && p->eps && p->mu && p->work && p->ckpt && p->partial && p->sums &&
            p->coef && p->target && p->stats && p->loss;
  if (ok && backward)
    ok = p->lam && p->acc && p->grads && (p->ckpt_every == 1 || p->traj);
  if (!ok) {
    set_error(std::string(who) + ": bad problem descriptor");
    return kErrInvalid;
  }
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_deffuant_init(const uint32_t *init_key, int64_t n, int32_t batch, float *x, int mode,
                      void *stream) {
  ABM_REQUIRE(init_key && x && n >= 1 && n < (1ll << 31) && batch >= 1 &&
                  (mode == 0 || mode == 1),
              "abm_deffuant_init: bad arguments");
  hipLaunchKernelGGL(k_deffuant_init, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     init_key, n, batch, x, mode);
  ABM_LAUNCH_CHECK("abm_deffuant_init");
  return kOk;
}

int abm_deffuant_step(const float *src, float *dst, int64_t n, int32_t batch,
                      const int32_t *perm, const float *eps, const float *mu,
                      int32_t pair_blocks, void *stream) {
  ABM_REQUIRE(src && dst && perm && eps && mu && n >= 1 && n < (1ll << 31) && batch >= 1 &&
                  pair_blocks >= 1 && pair_blocks <= 65535,
              "abm_deffuant_step: bad arguments");
  launch_step(src, dst, n, batch, perm, eps, mu, pair_blocks, (hipStream_t)stream);
  ABM_LAUNCH_CHECK("abm_deffuant_step");
  return kOk;
}

int abm_deffuant_forward(const abm_deffuant_problem *p, void *stream) {
  if (int rc = check_problem(p, "abm_deffuant_forward", false)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t nB = p->n * (int64_t)p->batch;
  const int32_t k = p->ckpt_every;
  // first step of the last segment; with the tape flag its steps write traj
  // out of place (the backward then skips recomputing it)
  const int32_t t_last = p->steps > 0 ? (p->steps - 1) / k * k : 0;
  const bool tape = (p->flags & ABM_DEFFUANT_TAPE_LAST) && p->traj && k > 1;
  const float *src = p->ckpt;
  for (int32_t t = 0; t < p->steps; ++t) {
    const bool to_ckpt = ((t + 1) % k) == 0 && t + 1 < p->steps;
    float *dst = p->work;
    if (to_ckpt)
      dst = p->ckpt + (int64_t)((t + 1) / k) * nB;
    else if (tape && t >= t_last && t + 1 < p->steps)
      dst = p->traj + (int64_t)(t - t_last) * nB;
    launch_step(src, dst, p->n, p->batch, p->perms + (int64_t)t * p->n, p->eps, p->mu,
                p->pair_blocks, s);
    src = dst;
  }
  launch_loss(p, p->steps > 0 ? p->work : p->ckpt, s);
  ABM_LAUNCH_CHECK("abm_deffuant_forward");
  return kOk;
}

int abm_deffuant_backward(const abm_deffuant_problem *p, void *stream) {
  if (int rc = check_problem(p, "abm_deffuant_backward", true)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = p->n, nB = n * (int64_t)p->batch;
  const int32_t k = p->ckpt_every;
  const float *final_x = p->steps > 0 ? p->work : p->ckpt;
  hipLaunchKernelGGL(k_deffuant_lambda, dim3(nblocks(nB, 256)), dim3(256), 0, s, final_x, n,
                     p->batch, (const double *)p->stats, (const double *)p->coef, p->lam);
  const int64_t acc_len = (int64_t)p->pair_blocks * 2 * p->batch;
  hipLaunchKernelGGL(k_fill_f64, dim3(nblocks(acc_len, 256)), dim3(256), 0, s, p->acc, acc_len,
                     0.0);
  const int32_t nseg = (p->steps + k - 1) / k;
  for (int32_t j = nseg - 1; j >= 0; --j) {
    const int32_t t0 = j * k, t1 = t0 + k < p->steps ? t0 + k : p->steps;
    const float *ck = p->ckpt + (int64_t)j * nB;
    const float *src = ck;
    const bool taped = j == nseg - 1 && (p->flags & ABM_DEFFUANT_TAPE_LAST);
    for (int32_t u = 0; !taped && u + 1 < t1 - t0; ++u) {
      float *dst = p->traj + (int64_t)u * nB;
      launch_step(src, dst, n, p->batch, p->perms + (int64_t)(t0 + u) * n, p->eps, p->mu,
                  p->pair_blocks, s);
      src = dst;
    }
    for (int32_t t = t1 - 1; t >= t0; --t) {
      const float *pre = t == t0 ? ck : p->traj + (int64_t)(t - t0 - 1) * nB;
      launch_adjoint(pre, p->lam, n, p->batch, p->perms + (int64_t)t * n, p->eps, p->mu,
                     p->pair_blocks, p->acc, s);
    }
  }
  hipLaunchKernelGGL(k_sum_rows, dim3(nblocks(p->batch, 64), 2u), dim3(256), 0, s,
                     (const double *)p->acc, (int64_t)p->pair_blocks, 2, p->batch, p->grads);
  ABM_LAUNCH_CHECK("abm_deffuant_backward");
  return kOk;
}

}  // extern "C"
