// reduce.hip -- record-time reductions over one agent variable: state
// counts, mean, Gini and the soft histogram.
//
// Reference: the metric reductions of JaxModel.step / agentpy
// Model.compute_metrics and AgentCollection.aggregate (jaxabm/model.py:
// 194-200, jaxabm/agentpy.py:926-942, jaxabm/agent.py:198-211) and the Gini
// definition of examples/models/advanced_economic_model.py:1798-1811.  The
// programs fuse these into their step kernels; these entry points serve any
// model's metrics.  Inputs and outputs stay on the device, nothing waits on
// the host, and outputs and scratch are cleared by fill kernels (no memset),
// so a captured step can hold them.
//
//  counts  integer atomics: exact and order-free
//  mean    int32: exact int64 sum, then f32(sum) / f32(n) as the reference's
//          jnp.mean of an int32 vector; float32: float64 block sums added in a
//          fixed order (deterministic), rounded once to float32
//  gini    sort (radix), then S = sum (i+1) x_(i) and T = sum x in uint32
//          (the reference's int32 wrap; addition mod 2^32 is order-free, so
//          atomics are exact); G = f32(2S) / f32(nT) - f32((n+1)/n), 0 when
//          T <= 1e-6
//  soft    h_k = mean_i [sigmoid(s (x_i - k/B)) - sigmoid(s (x_i - (k+1)/B))],
//  hist    float64, per-block partials added in a fixed order
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

int radix_sort_pairs(uint32_t *keys, int32_t *vals, int64_t n, void *ws, hipStream_t s,
                     int64_t segs);
size_t sort_workspace_bytes(int64_t n, int64_t segs);

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;
constexpr int kSoftMaxBins = 256;

inline unsigned red_blocks(int64_t n) {
  int64_t b = (n + kRedThreads - 1) / kRedThreads;
  if (b < 1) b = 1;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  return (unsigned)b;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kRedThreads)
    k_counts_u8(const uint8_t *__restrict__ s, int64_t n, int32_t k, int32_t *__restrict__ counts) {
  __shared__ int32_t h[256];
  for (int b = threadIdx.x; b < 256; b += kRedThreads) h[b] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kRedThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride)
    atomicAdd(&h[s[i]], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < k; b += kRedThreads)
    if (h[b]) atomicAdd(&counts[b], h[b]);
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kRedThreads / 64; ++q) s += red[q];  // fixed order
  return s;  // valid in thread 0
}

// partial[block] = sum of the block's grid-stride elements, as float64 (or
// as int64 for int32 input, exact)
template <typename In, typename Acc>
__global__ void __launch_bounds__(kRedThreads)
    k_partial_sum(const In *__restrict__ x, int64_t n, Acc *__restrict__ partial) {
  __shared__ Acc red[kRedThreads / 64];
  Acc v = 0;
  const int64_t stride = (int64_t)gridDim.x * kRedThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride)
    v += (Acc)x[i];
  const Acc s = block_sum(v, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// One workgroup of 1024 threads reduces the whole array (n up to
// kSmallMax): thread t accumulates elements t, t + 1024, ... in order, then
// the wavefront sums (fixed shuffle order) are added in wave order.  One
// launch, no workspace.  OUT: 0 float32 sum, 1 float32 mean, 2 int32 sum
// (wrapping), 3 float32 mean of int32 (f32(exact sum) / f32(n)).
constexpr int64_t kSmallMax = 65536;

// One 1024-thread workgroup reduces x[0, n): thread t sums x[t], x[t+1024],
// ... in Acc, the wavefront sums by xor shuffles, then thread 0 adds the 16
// wavefront sums in order.  Every whole-array sum and mean of up to kSmallMax
// values runs exactly this sequence, alone or batched.
template <typename In, typename Acc, int OUT>
__device__ __forceinline__ void reduce_one_block(const In *__restrict__ x, int64_t n,
                                                 void *__restrict__ out, Acc *red) {
  Acc v = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) v += (Acc)x[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc s = 0;
    for (int w = 0; w < 16; ++w) s += red[w];
    if constexpr (OUT == 0) *(float *)out = (float)s;
    else if constexpr (OUT == 1) *(float *)out = (float)((double)s / (double)n);
    else if constexpr (OUT == 2) *(int32_t *)out = (int32_t)(uint32_t)(uint64_t)s;
    else *(float *)out = (float)s / (float)n;
  }
}

template <typename In, typename Acc, int OUT>
__global__ void __launch_bounds__(1024)
    k_reduce_one_block(const In *__restrict__ x, int64_t n, void *__restrict__ out) {
  __shared__ Acc red[16];
  reduce_one_block<In, Acc, OUT>(x, n, out, red);
}

// Several whole-array reductions in one launch: workgroup b runs item b
// (items travel by value, so a captured launch keeps them).
struct ReduceBatch {
  abm_reduce_item items[ABM_REDUCE_BATCH_MAX];
};

__global__ void __launch_bounds__(1024) k_reduce_batch(ReduceBatch batch) {
  __shared__ double redf[16];
  __shared__ int64_t redi[16];
  const abm_reduce_item it = batch.items[blockIdx.x];
  switch (it.op) {  // block-uniform
    case ABM_REDUCE_SUM_F32:
      reduce_one_block<float, double, 0>((const float *)it.x, it.n, it.out, redf);
      break;
    case ABM_REDUCE_MEAN_F32:
      reduce_one_block<float, double, 1>((const float *)it.x, it.n, it.out, redf);
      break;
    case ABM_REDUCE_SUM_I32:
      reduce_one_block<int32_t, int64_t, 2>((const int32_t *)it.x, it.n, it.out, redi);
      break;
    case ABM_REDUCE_MEAN_I32:
      reduce_one_block<int32_t, int64_t, 3>((const int32_t *)it.x, it.n, it.out, redi);
      break;
    default:
      reduce_one_block<uint8_t, int64_t, 2>((const uint8_t *)it.x, it.n, it.out, redi);
      break;
  }
}

// Final pass of the multi-block sums: partial[b] added in block order.
template <typename Acc, int OUT>
__global__ void k_reduce_final(const Acc *__restrict__ partial, int nb, int64_t n,
                               void *__restrict__ out) {
  Acc s = 0;
  for (int b = 0; b < nb; ++b) s += partial[b];
  if constexpr (OUT == 0) *(float *)out = (float)s;
  else if constexpr (OUT == 2) *(int32_t *)out = (int32_t)(uint32_t)(uint64_t)s;
}

__global__ void k_mean_f32_final(const double *__restrict__ partial, int nb, int64_t n,
                                 float *__restrict__ out) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[b];
  *out = (float)(s / (double)n);
}

__global__ void k_mean_i32_final(const int64_t *__restrict__ partial, int nb, int64_t n,
                                 float *__restrict__ out) {
  int64_t s = 0;
  for (int b = 0; b < nb; ++b) s += partial[b];
  *out = (float)s / (float)n;
}

// int32 -> uint32 keys in ascending signed order
__global__ void k_gini_keys(const int32_t *__restrict__ x, int64_t n, uint32_t *__restrict__ keys,
                            int32_t *__restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)x[i] ^ 0x80000000u;
  vals[i] = 0;
}

// st[0] += sum (i+1) x_(i), st[1] += sum x (mod 2^32) over sorted keys
__global__ void __launch_bounds__(kRedThreads)
    k_gini_sums(const uint32_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ st) {
  __shared__ uint32_t red[kRedThreads / 64];
  uint32_t s = 0, t = 0;
  const int64_t stride = (int64_t)gridDim.x * kRedThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride) {
    const uint32_t v = keys[i] ^ 0x80000000u;
    s += (uint32_t)(i + 1) * v;
    t += v;
  }
  const uint32_t S = block_sum(s, red);
  __syncthreads();
  const uint32_t T = block_sum(t, red);
  if (threadIdx.x == 0) {
    atomicAdd(&st[0], S);
    atomicAdd(&st[1], T);
  }
}

__global__ void k_gini_final(const uint32_t *__restrict__ st, int64_t n, float *__restrict__ out) {
  const uint32_t S = st[0], T = st[1];
  const int32_t Ti = (int32_t)T;
  float g = 0.0f;
  if ((float)Ti > 1e-6f) {
    const int32_t num = (int32_t)(2u * S);
    const int32_t den = (int32_t)((uint32_t)n * T);
    const float q = (float)num / (float)den;
    const float c = (float)((double)(n + 1) / (double)n);
    g = q - c;
  }
  *out = g;
}

// partial[block][k] = sum over the block's elements of
// sigmoid(s (x - k/B)) - sigmoid(s (x - (k+1)/B))
__global__ void __launch_bounds__(kRedThreads)
    k_soft_hist(const float *__restrict__ x, int64_t n, int32_t bins, double slope,
                double *__restrict__ partial) {
  __shared__ double red[kRedThreads / 64][kSoftMaxBins];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double h[8];
  for (int k0 = 0; k0 < bins; k0 += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kRedThreads;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride) {
      const double xi = (double)x[i];
      double prev = 1.0 / (1.0 + exp(-slope * (xi - (double)k0 / bins)));
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + j;
        if (k < bins) {
          const double cur = 1.0 / (1.0 + exp(-slope * (xi - (double)(k + 1) / bins)));
          h[j] += prev - cur;
          prev = cur;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      double v = h[j];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0 && k0 + j < bins) red[wave][k0 + j] = v;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bins; k += kRedThreads) {
    double s = 0.0;
    for (int q = 0; q < kRedThreads / 64; ++q) s += red[q][k];  // fixed order
    partial[(int64_t)blockIdx.x * bins + k] = s;
  }
}

__global__ void k_soft_hist_final(const double *__restrict__ partial, int nb, int32_t bins,
                                  int64_t n, double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= bins) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * bins + k];
  out[k] = s / (double)n;
}

// ---- scratch of abm_gini_i32_noalloc: a file-scope device array, one copy
// per GPU for the life of the process, sized for ABM_GINI_NOALLOC_MAX_N
// values (abm_reduce_workspace of that n, checked at set-up).  Host code
// never uses the array's name as an address: the address of each device's
// copy comes from hipGetSymbolAddress, taken once per device with that
// device current.
constexpr size_t kNoAllocScratchBytes = 5505024;  // 5.25 MiB
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// The current device's copy of `scratch` (set up on its first use), or an
// error status.
int noalloc_scratch(void **out) {
  int dev = -1;
  ABM_CHECK_HIP(hipGetDevice(&dev));
  ABM_REQUIRE(dev >= 0 && dev < kMaxDevices, "abm_gini_i32_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    ABM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ABM_REQUIRE(p && ((uintptr_t)p & 255) == 0, "abm_gini_i32_noalloc: scratch misaligned");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

// Whole-array sum: one block up to kSmallMax elements, else block partials
// (k_partial_sum) and a final pass.
template <typename In, typename Acc, int OUT>
int sum_all(const In *x, int64_t n, void *out, void *ws, size_t ws_bytes, hipStream_t s,
            const char *who) {
  if (!(x && out && n >= 1)) {
    set_error(std::string(who) + ": bad arguments");
    return kErrInvalid;
  }
  if (n <= kSmallMax) {
    hipLaunchKernelGGL((k_reduce_one_block<In, Acc, OUT>), dim3(1), dim3(1024), 0, s, x, n, out);
  } else {
    if (!ws || ws_bytes < (size_t)kRedMaxBlocks * 8) {
      set_error(std::string(who) + ": workspace too small");
      return kErrInvalid;
    }
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL((k_partial_sum<In, Acc>), dim3(nb), dim3(kRedThreads), 0, s, x, n,
                       (Acc *)ws);
    hipLaunchKernelGGL((k_reduce_final<Acc, OUT>), dim3(1), dim3(1), 0, s, (const Acc *)ws,
                       (int)nb, n, out);
  }
  ABM_LAUNCH_CHECK(who);
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_reduce_counts_u8(const uint8_t *states, int64_t n, int32_t k, int32_t *counts,
                         void *stream) {
  ABM_REQUIRE((states || n == 0) && counts && n >= 0 && k >= 1 && k <= 256,
              "abm_reduce_counts_u8: bad arguments (1 <= k <= 256)");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, (uint32_t *)counts, (int64_t)k, 0u);
  if (n > 0)
    hipLaunchKernelGGL(k_counts_u8, dim3(red_blocks(n)), dim3(kRedThreads), 0, s, states, n, k,
                       counts);
  ABM_LAUNCH_CHECK("abm_reduce_counts_u8");
  return kOk;
}

size_t abm_reduce_workspace(int64_t n) {
  // gini: keys + values + the radix sort's workspace; means and the soft
  // histogram use less
  if (n < 1) n = 1;
  const size_t gini = (size_t)n * 8 + sort_workspace_bytes(n, 1) + 256;
  const size_t soft = (size_t)kRedMaxBlocks * kSoftMaxBins * 8;
  return (gini > soft ? gini : soft) + 256;
}

int abm_mean_f32(const float *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream) {
  ABM_REQUIRE(x && out && n >= 1, "abm_mean_f32: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (n <= kSmallMax) {
    hipLaunchKernelGGL((k_reduce_one_block<float, double, 1>), dim3(1), dim3(1024), 0, s, x, n,
                       (void *)out);
    ABM_LAUNCH_CHECK("abm_mean_f32");
    return kOk;
  }
  ABM_REQUIRE(ws && ws_bytes >= (size_t)kRedMaxBlocks * 8, "abm_mean_f32: workspace too small");
  const unsigned nb = red_blocks(n);
  hipLaunchKernelGGL((k_partial_sum<float, double>), dim3(nb), dim3(kRedThreads), 0, s, x, n,
                     (double *)ws);
  hipLaunchKernelGGL(k_mean_f32_final, dim3(1), dim3(1), 0, s, (const double *)ws, (int)nb, n,
                     out);
  ABM_LAUNCH_CHECK("abm_mean_f32");
  return kOk;
}

int abm_mean_i32(const int32_t *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream) {
  ABM_REQUIRE(x && out && n >= 1, "abm_mean_i32: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (n <= kSmallMax) {
    hipLaunchKernelGGL((k_reduce_one_block<int32_t, int64_t, 3>), dim3(1), dim3(1024), 0, s, x, n,
                       (void *)out);
    ABM_LAUNCH_CHECK("abm_mean_i32");
    return kOk;
  }
  ABM_REQUIRE(ws && ws_bytes >= (size_t)kRedMaxBlocks * 8, "abm_mean_i32: workspace too small");
  const unsigned nb = red_blocks(n);
  hipLaunchKernelGGL((k_partial_sum<int32_t, int64_t>), dim3(nb), dim3(kRedThreads), 0, s, x, n,
                     (int64_t *)ws);
  hipLaunchKernelGGL(k_mean_i32_final, dim3(1), dim3(1), 0, s, (const int64_t *)ws, (int)nb, n,
                     out);
  ABM_LAUNCH_CHECK("abm_mean_i32");
  return kOk;
}

int abm_reduce_batch(const abm_reduce_item *items, int32_t count, void *stream) {
  ABM_REQUIRE(items && count >= 1 && count <= ABM_REDUCE_BATCH_MAX,
              "abm_reduce_batch: bad arguments (1 <= count <= ABM_REDUCE_BATCH_MAX)");
  ReduceBatch batch;
  for (int32_t i = 0; i < count; ++i) {
    const abm_reduce_item &it = items[i];
    ABM_REQUIRE(it.x && it.out && it.n >= 1 && it.n <= kSmallMax && it.op >= ABM_REDUCE_SUM_F32 &&
                    it.op <= ABM_REDUCE_SUM_U8,
                "abm_reduce_batch: bad item (1 <= n <= 65536, a known op)");
    batch.items[i] = it;
  }
  hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)count), dim3(1024), 0, (hipStream_t)stream,
                     batch);
  ABM_LAUNCH_CHECK("abm_reduce_batch");
  return kOk;
}

int abm_sum_f32(const float *x, int64_t n, float *out, void *ws, size_t ws_bytes, void *stream) {
  return sum_all<float, double, 0>(x, n, out, ws, ws_bytes, (hipStream_t)stream, "abm_sum_f32");
}

int abm_sum_i32(const int32_t *x, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
                void *stream) {
  return sum_all<int32_t, int64_t, 2>(x, n, out, ws, ws_bytes, (hipStream_t)stream,
                                      "abm_sum_i32");
}

int abm_sum_u8(const uint8_t *x, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
               void *stream) {
  return sum_all<uint8_t, int64_t, 2>(x, n, out, ws, ws_bytes, (hipStream_t)stream,
                                      "abm_sum_u8");
}

int abm_gini_i32(const int32_t *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream) {
  ABM_REQUIRE(x && out && ws && n >= 1 && n < (1ll << 31) && ws_bytes >= abm_reduce_workspace(n),
              "abm_gini_i32: bad arguments or workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char *p = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  uint32_t *st = (uint32_t *)p;
  p += 256;
  uint32_t *keys = (uint32_t *)p;
  p += (size_t)n * 4;
  int32_t *vals = (int32_t *)p;
  p += (size_t)n * 4;
  void *sort_ws = (void *)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, st, (int64_t)2, 0u);
  hipLaunchKernelGGL(k_gini_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, keys,
                     vals);
  if (int rc = radix_sort_pairs(keys, vals, n, sort_ws, s, 1)) return rc;
  hipLaunchKernelGGL(k_gini_sums, dim3(red_blocks(n)), dim3(kRedThreads), 0, s, keys, n, st);
  hipLaunchKernelGGL(k_gini_final, dim3(1), dim3(1), 0, s, (const uint32_t *)st, n, out);
  ABM_LAUNCH_CHECK("abm_gini_i32");
  return kOk;
}

int abm_gini_i32_noalloc(const int32_t *x, int64_t n, float *out, void *stream) {
  ABM_REQUIRE(x && out && n >= 1, "abm_gini_i32_noalloc: bad arguments");
  ABM_REQUIRE(n <= ABM_GINI_NOALLOC_MAX_N,
              "abm_gini_i32_noalloc: n exceeds ABM_GINI_NOALLOC_MAX_N (262144); use abm_gini_i32 "
              "with a workspace");
  const size_t need = abm_reduce_workspace(ABM_GINI_NOALLOC_MAX_N);
  ABM_REQUIRE(need + 256 <= kNoAllocScratchBytes, "abm_gini_i32_noalloc: scratch too small");
  void *ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  // the same kernels as abm_gini_i32; its counters are cleared by a fill
  // kernel inside, so nothing carries over from the previous call
  return abm_gini_i32(x, n, out, ws, kNoAllocScratchBytes, stream);
}

int abm_soft_hist_f32(const float *x, int64_t n, int32_t bins, double slope, double *out,
                      void *ws, size_t ws_bytes, void *stream) {
  ABM_REQUIRE(x && out && ws && n >= 1 && bins >= 1 && bins <= kSoftMaxBins &&
                  ws_bytes >= (size_t)kRedMaxBlocks * bins * 8,
              "abm_soft_hist_f32: bad arguments (1 <= bins <= 256)");
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  hipLaunchKernelGGL(k_soft_hist, dim3(nb), dim3(kRedThreads), 0, s, x, n, bins, slope,
                     (double *)ws);
  hipLaunchKernelGGL(k_soft_hist_final, dim3((unsigned)((bins + 63) / 64)), dim3(64), 0, s,
                     (const double *)ws, (int)nb, bins, n, out);
  ABM_LAUNCH_CHECK("abm_soft_hist_f32");
  return kOk;
}

}  // extern "C"
