// sort.hip -- stable LSD radix sort of (uint32 key, int32 value) pairs and
// jax.random.permutation built on it.
//
// jax.random.permutation(key, n) (reference call sites
// jaxabm/analysis.py:87 and the Deffuant/SIR programs' pairing and seeding)
// runs ceil(3 ln n / ln(2^32-1)) rounds of: key, sub = split(key);
// x = sort_key_val(bits(sub, (n,)), x) with a STABLE sort, so ties among the
// 32-bit keys keep their previous order.  Ties are likely at n = 65,536
// (birthday bound), so the sort must be stable: an LSD radix sort whose
// per-tile ranks follow the input order is.
//
// Per 8-bit digit pass: (1) tile histograms, (2) one exclusive scan over
// digit-major / tile-minor counts, (3) stable scatter, ranking within a
// 64-wide wavefront by ballot matching on the digit bits.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;  // items per thread per tile, in rounds
constexpr int kTile = kSortThreads * kSortRounds;
constexpr int kWaves = kSortThreads / 64;

__global__ void __launch_bounds__(kSortThreads)
    k_digit_hist(const uint32_t *__restrict__ keys, int64_t n, int shift,
                 uint32_t *__restrict__ hist, int tiles) {
  // segment blockIdx.y: keys[seg * n ..), hist[seg * 256 * tiles ..)
  keys += (int64_t)blockIdx.y * n;
  hist += (int64_t)blockIdx.y * 256 * tiles;
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kTile;
  for (int r = 0; r < kSortRounds; ++r) {
    int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of `len` counts by a single block per segment (blockIdx.y).
__global__ void __launch_bounds__(1024)
    k_scan_exclusive(uint32_t *__restrict__ data, int64_t len) {
  data += (int64_t)blockIdx.y * len;
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  int64_t per = (len + 1023) / 1024;
  int64_t lo = t * per, hi = lo + per < len ? lo + per : len;
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += data[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (int64_t i = lo; i < hi; ++i) {
    uint32_t v = data[i];
    data[i] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kSortThreads)
    k_digit_scatter(const uint32_t *__restrict__ keys_in,
                    const int32_t *__restrict__ vals_in, int64_t n, int shift,
                    const uint32_t *__restrict__ offsets, int tiles,
                    uint32_t *__restrict__ keys_out,
                    int32_t *__restrict__ vals_out) {
  {
    const int64_t seg = (int64_t)blockIdx.y * n;
    keys_in += seg;
    vals_in += seg;
    keys_out += seg;
    vals_out += seg;
    offsets += (int64_t)blockIdx.y * 256 * tiles;
  }
  __shared__ uint32_t base[256];
  __shared__ uint32_t wc[kWaves][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  base[t] = offsets[(int64_t)t * tiles + blockIdx.x];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tile0 = (int64_t)blockIdx.x * kTile;
  for (int r = 0; r < kSortRounds; ++r) {
    for (int w = 0; w < kWaves; ++w) wc[w][t] = 0;
    __syncthreads();
    const int64_t i = tile0 + (int64_t)r * kSortThreads + t;
    const bool valid = i < n;
    uint32_t k = valid ? keys_in[i] : 0u;
    int32_t v = valid ? vals_in[i] : 0;
    uint32_t d = (k >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      keys_out[pos] = k;
      vals_out[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kWaves; ++w) add += wc[w][t];
    base[t] += add;
    __syncthreads();
  }
}

// out[s * n + i] = i for segment s = blockIdx.y
__global__ void k_iota(int32_t *__restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[(int64_t)blockIdx.y * n + i] = (int32_t)i;
}

// Round r's sort keys: k_0 = key; k_{j+1}, sub_j = split(k_j); bits(sub_r).
// Segment s = blockIdx.y takes its key at key[s * key_stride].
__global__ void k_perm_round_bits(const uint32_t *__restrict__ key, int64_t key_stride,
                                  int round, int64_t n, uint32_t *__restrict__ out,
                                  int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key += (int64_t)blockIdx.y * key_stride;
  out += (int64_t)blockIdx.y * n;
  Key k{key[0], key[1]}, sub;
  for (int j = 0; j <= round; ++j) {
    Key nk;
    split2(k, nk, sub, mode);
    k = nk;
  }
  out[i] = bits_at(sub, n, i, mode);
}

size_t sort_workspace_bytes(int64_t n, int64_t segs = 1) {
  int64_t tiles = (n + kTile - 1) / kTile;
  // keys x2, vals x1 (the second value buffer is the caller's), hist
  size_t b = 0;
  b += (size_t)n * (size_t)segs * 4 * 3;
  b += (size_t)tiles * 256 * 4 * (size_t)segs;
  return (b + 255) & ~(size_t)255;
}

// Sort each of `segs` consecutive segments of n (key, value) pairs, keys[]
// destroyed, vals[] in place, using ws.
int radix_sort_pairs(uint32_t *keys, int32_t *vals, int64_t n, void *ws,
                     hipStream_t s, int64_t segs = 1) {
  if (n <= 1 || segs < 1) return kOk;
  int tiles = (int)((n + kTile - 1) / kTile);
  const int64_t total = n * segs;
  char *p = (char *)ws;
  uint32_t *k2 = (uint32_t *)p;
  p += (size_t)total * 4;
  int32_t *v2 = (int32_t *)p;
  p += (size_t)total * 4;
  p += (size_t)total * 4;  // reserved
  uint32_t *hist = (uint32_t *)p;
  uint32_t *ka = keys, *kb = k2;
  int32_t *va = vals, *vb = v2;
  const dim3 grid((unsigned)tiles, (unsigned)segs);
  for (int shift = 0; shift < 32; shift += 8) {
    hipLaunchKernelGGL(k_digit_hist, grid, dim3(kSortThreads), 0, s, ka, n, shift, hist,
                       tiles);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1, (unsigned)segs), dim3(1024), 0, s, hist,
                       (int64_t)tiles * 256);
    hipLaunchKernelGGL(k_digit_scatter, grid, dim3(kSortThreads), 0, s, ka, va, n, shift,
                       hist, tiles, kb, vb);
    uint32_t *tk = ka;
    ka = kb;
    kb = tk;
    int32_t *tv = va;
    va = vb;
    vb = tv;
  }
  // 4 passes: the result is back in (keys, vals)
  ABM_LAUNCH_CHECK("radix_sort_pairs");
  return kOk;
}

// permutation(key_s, n) for segs keys (key_s at keys[s * key_stride]), the
// results in out[s * n ..).
int permutation_segments(const uint32_t *keys, int64_t key_stride, int64_t segs, int64_t n,
                         int32_t *out, void *ws, int mode, hipStream_t s) {
  const dim3 g((unsigned)((n + 255) / 256), (unsigned)segs);
  hipLaunchKernelGGL(k_iota, g, dim3(256), 0, s, out, n);
  uint32_t *bits = (uint32_t *)((char *)ws + sort_workspace_bytes(n, segs));
  const int rounds = (int)__builtin_ceil(3.0 * __builtin_log((double)(n > 1 ? n : 1)) /
                                         __builtin_log(4294967295.0));
  for (int r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(k_perm_round_bits, g, dim3(256), 0, s, keys, key_stride, r, n, bits,
                       mode);
    int rc = radix_sort_pairs(bits, out, n, ws, s, segs);
    if (rc) return rc;
  }
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

size_t abm_permutation_workspace(int64_t n) {
  return sort_workspace_bytes(n) + (((size_t)n * 4 + 255) & ~(size_t)255);
}

int abm_permutation_rounds(int64_t n) {
  double nn = (double)(n > 1 ? n : 1);
  return (int)__builtin_ceil(3.0 * __builtin_log(nn) / __builtin_log(4294967295.0));
}

int abm_permutation(const uint32_t *key, int64_t n, int32_t *out, void *ws,
                    size_t ws_bytes, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && n < (1ll << 31), "abm_permutation: n out of range");
  ABM_REQUIRE(mode == 0 || mode == 1, "abm_permutation: bad mode");
  if (n == 0) return kOk;
  ABM_REQUIRE(key && out && ws && ws_bytes >= abm_permutation_workspace(n),
              "abm_permutation: workspace too small");
  int rc = permutation_segments(key, 0, 1, n, out, ws, mode, (hipStream_t)stream);
  if (rc) return rc;
  ABM_LAUNCH_CHECK("abm_permutation");
  return kOk;
}

size_t abm_permutation_batch_workspace(int64_t n, int32_t count) {
  if (n < 0 || count < 0) return 0;
  return sort_workspace_bytes(n, count) +
         (((size_t)n * (size_t)count * 4 + 255) & ~(size_t)255);
}

int abm_permutation_batch(const uint32_t *keys, int64_t key_stride, int32_t count, int64_t n,
                          int32_t *out, void *ws, size_t ws_bytes, int mode, void *stream) {
  ABM_REQUIRE(n >= 0 && n < (1ll << 31) && count >= 0 && count < 65536 &&
                  n * (int64_t)count < (1ll << 31) && key_stride >= 2,
              "abm_permutation_batch: sizes out of range");
  ABM_REQUIRE(mode == 0 || mode == 1, "abm_permutation_batch: bad mode");
  if (n == 0 || count == 0) return kOk;
  ABM_REQUIRE(keys && out && ws && ws_bytes >= abm_permutation_batch_workspace(n, count),
              "abm_permutation_batch: workspace too small");
  int rc = permutation_segments(keys, key_stride, count, n, out, ws, mode, (hipStream_t)stream);
  if (rc) return rc;
  ABM_LAUNCH_CHECK("abm_permutation_batch");
  return kOk;
}

int abm_sort_pairs_u32(uint32_t *keys, int32_t *vals, int64_t n, void *ws,
                       size_t ws_bytes, void *stream) {
  ABM_REQUIRE(n >= 0 && n < (1ll << 31), "abm_sort_pairs_u32: n out of range");
  if (n <= 1) return kOk;
  ABM_REQUIRE(ws && ws_bytes >= sort_workspace_bytes(n),
              "abm_sort_pairs_u32: workspace too small");
  return radix_sort_pairs(keys, vals, n, ws, (hipStream_t)stream);
}

}  // extern "C"
