// csr_build.hip -- a Network's CSR built on the device from its edge list.
//
// The reference stores a Network as an (E, 2) int32 edge list grown one
// jnp.concatenate per add_edge, and answers get_neighbors with O(E) masks
// (jaxabm/agentpy.py:545-615): the targets of the agent's entries in
// insertion order when directed, the sorted unique ids of both directions when
// undirected.  Here the whole list becomes CSR in a few launches:
//
//  1. entries (row, col) = (src, dst), and with ABM_CSR_SYMMETRIZE also
//     (dst, src) right after each one (every edge in both directions, as
//     abm_csr_undirected builds on the host);
//  2. a stable LSD radix sort by col, then a stable one by row: rows
//     ascending, neighbours ascending (with ABM_CSR_INSERTION_ORDER only the
//     sort by row, so a row keeps the entries' order);
//  3. with ABM_CSR_DEDUP, repeated (row, col) entries dropped: first-of-run
//     flags, an exclusive scan over tiles (tile sums, a scan of those, the
//     tile scans), and a compaction;
//  4. row_ptr from the sorted rows: entry i writes row_ptr[r] = i for every
//     row r in (row[i-1], row[i]], the last entry also the rows after it.
//
// The sorts are the engine's stable 8-bit-digit radix sort (sort.hip).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

int radix_sort_pairs(uint32_t *keys, int32_t *vals, int64_t n, void *ws, hipStream_t s,
                     int64_t segs);
size_t sort_workspace_bytes(int64_t n, int64_t segs);

constexpr int kBuildThreads = 256;
constexpr int kScanTile = 1024 * 4;  // entries per scan tile (4 per thread)

inline unsigned build_blocks(int64_t n) {
  return (unsigned)((n + kBuildThreads - 1) / kBuildThreads);
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Entry layout: sorted mode puts the col in `key` and the row in `val` (the
// first sort is by col); insertion order puts the row in `key`.  An endpoint
// outside [0, n) sets *err and is replaced by 0 so later passes stay in range.
__global__ void k_csr_entries(int64_t n, int64_t m, const int32_t *__restrict__ src,
                              const int32_t *__restrict__ dst, int sym, int by_row,
                              uint32_t *__restrict__ key, int32_t *__restrict__ val,
                              uint32_t *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int32_t s = src[i], d = dst[i];
  if (s < 0 || s >= n || d < 0 || d >= n) {
    atomicOr(err, 1u);
    s = d = 0;
  }
  const int64_t j = sym ? 2 * i : i;
  if (by_row) {
    key[j] = (uint32_t)s;
    val[j] = d;
    if (sym) {
      key[j + 1] = (uint32_t)d;
      val[j + 1] = s;
    }
  } else {
    key[j] = (uint32_t)d;
    val[j] = s;
    if (sym) {
      key[j + 1] = (uint32_t)s;
      val[j + 1] = d;
    }
  }
}

// flag[i] = 1 when entry i is the first of its (row, col) run (rows and cols
// sorted).
__global__ void k_csr_first(const uint32_t *__restrict__ row, const int32_t *__restrict__ col,
                            int64_t nnz, uint32_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  flag[i] = (i == 0 || row[i] != row[i - 1] || col[i] != col[i - 1]) ? 1u : 0u;
}

// Tile sums of flag[] (kScanTile entries per block of 1024 threads).
__global__ void __launch_bounds__(1024)
    k_tile_sums(const uint32_t *__restrict__ flag, int64_t len, uint32_t *__restrict__ sums) {
  __shared__ uint32_t red[16];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < len) s += flag[base + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; ++w) t += red[w];
    sums[blockIdx.x] = t;
  }
}

// Exclusive scan of the tile sums by one block (serial per thread chunk,
// then a scan of the chunk totals); sums[tiles] = the grand total.
__global__ void __launch_bounds__(1024)
    k_scan_tile_sums(uint32_t *__restrict__ sums, int64_t tiles) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (tiles + 1023) / 1024;
  const int64_t lo = t * per, hi = lo + per < tiles ? lo + per : tiles;
  uint32_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += sums[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (int64_t i = lo; i < hi; ++i) {
    const uint32_t v = sums[i];
    sums[i] = run;
    run += v;
  }
  if (t == 1023) sums[tiles] = part[1023];
}

// Compaction: entry i with flag 1 goes to position (tile offset + its
// exclusive rank within the tile).
__global__ void __launch_bounds__(1024)
    k_compact(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ tile_off,
              const uint32_t *__restrict__ row, const int32_t *__restrict__ col, int64_t len,
              uint32_t *__restrict__ row_out, int32_t *__restrict__ col_out) {
  __shared__ uint32_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + t * 4;
  uint32_t f[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[k] = base + k < len ? flag[base + k] : 0u;
    s += f[k];
  }
  // inclusive scan of s across the wave, then across waves
  uint32_t incl = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  uint32_t pos = tile_off[blockIdx.x] + before + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (f[k]) {
      row_out[pos] = row[base + k];
      col_out[pos] = col[base + k];
      ++pos;
    }
  }
}

// row_ptr from ascending rows: entry i sets row_ptr[r] = i for rows r in
// (row[i-1], row[i]]; the last entry also sets rows row[last]+1 .. n to nnz.
__global__ void k_row_ptr(const uint32_t *__restrict__ row, int64_t nnz, int64_t n,
                          int32_t *__restrict__ row_ptr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int64_t r = row[i];
  const int64_t prev = i ? (int64_t)row[i - 1] : -1;
  for (int64_t q = prev + 1; q <= r; ++q) row_ptr[q] = (int32_t)i;
  if (i == nnz - 1)
    for (int64_t q = r + 1; q <= n; ++q) row_ptr[q] = (int32_t)nnz;
}

__global__ void k_fill_rows(int32_t *__restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy_cols(const int32_t *__restrict__ in, int64_t n, int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Workspace layout for `len` entries: error word, two key and two value
// arrays, flags, tile sums, then the radix sort's own workspace.
struct BuildWs {
  uint32_t *err;
  uint32_t *key, *key2;
  int32_t *val, *val2;
  uint32_t *flag, *sums;
  void *sort;
  size_t bytes;
};

BuildWs build_layout(char *base, int64_t len) {
  BuildWs w;
  size_t off = 0;
  auto take = [&](size_t b) {
    char *p = base ? base + off : nullptr;
    off += align256(b);
    return p;
  };
  const int64_t tiles = (len + kScanTile - 1) / kScanTile;
  w.err = (uint32_t *)take(4);
  w.key = (uint32_t *)take((size_t)len * 4);
  w.key2 = (uint32_t *)take((size_t)len * 4);
  w.val = (int32_t *)take((size_t)len * 4);
  w.val2 = (int32_t *)take((size_t)len * 4);
  w.flag = (uint32_t *)take((size_t)len * 4);
  w.sums = (uint32_t *)take((size_t)(tiles + 1) * 4);
  w.sort = take(sort_workspace_bytes(len > 1 ? len : 1, 1));
  w.bytes = off;
  return w;
}

int64_t build_len(int64_t m, int32_t flags) {
  return (flags & ABM_CSR_SYMMETRIZE) ? 2 * m : m;
}

}  // namespace abm

using namespace abm;

extern "C" {

size_t abm_csr_from_edges_workspace(int64_t m, int32_t flags) {
  if (m < 0) return 0;
  return build_layout(nullptr, build_len(m, flags)).bytes + 256;
}

int abm_csr_from_edges(int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                       int32_t flags, int32_t *row_ptr, int32_t *col, int64_t *nnz, void *ws,
                       size_t ws_bytes, void *stream) {
  ABM_REQUIRE(n >= 0 && n < (1ll << 31) - 1 && m >= 0, "abm_csr_from_edges: bad sizes");
  ABM_REQUIRE((flags & ~(ABM_CSR_SYMMETRIZE | ABM_CSR_DEDUP | ABM_CSR_INSERTION_ORDER)) == 0,
              "abm_csr_from_edges: unknown flag");
  ABM_REQUIRE(!((flags & ABM_CSR_DEDUP) && (flags & ABM_CSR_INSERTION_ORDER)),
              "abm_csr_from_edges: ABM_CSR_DEDUP needs sorted rows (not ABM_CSR_INSERTION_ORDER)");
  const int64_t len = build_len(m, flags);
  ABM_REQUIRE(len < (1ll << 31), "abm_csr_from_edges: too many entries");
  ABM_REQUIRE(row_ptr && nnz && (m == 0 || (src && dst && col)),
              "abm_csr_from_edges: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (len == 0) {
    hipLaunchKernelGGL(k_fill_rows, dim3(build_blocks(n + 1)), dim3(kBuildThreads), 0, s, row_ptr,
                       n + 1, 0);
    ABM_LAUNCH_CHECK("abm_csr_from_edges");
    *nnz = 0;
    return kOk;
  }
  ABM_REQUIRE(ws && ws_bytes >= abm_csr_from_edges_workspace(m, flags),
              "abm_csr_from_edges: workspace too small");
  char *base = (char *)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  const BuildWs w = build_layout(base, len);
  const bool by_row = (flags & ABM_CSR_INSERTION_ORDER) != 0;
  const bool sym = (flags & ABM_CSR_SYMMETRIZE) != 0;
  hipLaunchKernelGGL(k_fill_rows, dim3(1), dim3(64), 0, s, (int32_t *)w.err, (int64_t)1, 0);
  hipLaunchKernelGGL(k_csr_entries, dim3(build_blocks(m)), dim3(kBuildThreads), 0, s, n, m, src,
                     dst, sym ? 1 : 0, by_row ? 1 : 0, w.key, w.val, w.err);
  // rows (as keys) and cols (as values) sorted by row, stably
  uint32_t *rows = w.key;
  int32_t *cols = w.val;
  if (!by_row) {
    // by col first, then the rows (the values) become the keys of a second
    // stable sort that carries the cols along
    if (int rc = radix_sort_pairs(w.key, w.val, len, w.sort, s, 1)) return rc;
    rows = reinterpret_cast<uint32_t *>(w.val);
    cols = reinterpret_cast<int32_t *>(w.key);
  }
  if (int rc = radix_sort_pairs(rows, cols, len, w.sort, s, 1)) return rc;
  int64_t count = len;
  if (flags & ABM_CSR_DEDUP) {
    const int64_t tiles = (len + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_csr_first, dim3(build_blocks(len)), dim3(kBuildThreads), 0, s, rows, cols,
                       len, w.flag);
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)tiles), dim3(1024), 0, s, w.flag, len, w.sums);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, s, w.sums, tiles);
    hipLaunchKernelGGL(k_compact, dim3((unsigned)tiles), dim3(1024), 0, s, w.flag, w.sums, rows,
                       cols, len, w.key2, col);
    uint32_t total = 0;
    ABM_CHECK_HIP(hipMemcpyAsync(&total, w.sums + tiles, 4, hipMemcpyDeviceToHost, s));
    ABM_CHECK_HIP(hipStreamSynchronize(s));
    count = total;
    rows = w.key2;
  } else {
    hipLaunchKernelGGL(k_copy_cols, dim3(build_blocks(len)), dim3(kBuildThreads), 0, s, cols, len,
                       col);
  }
  hipLaunchKernelGGL(k_fill_rows, dim3(build_blocks(n + 1)), dim3(kBuildThreads), 0, s, row_ptr,
                     n + 1, 0);
  if (count > 0)
    hipLaunchKernelGGL(k_row_ptr, dim3(build_blocks(count)), dim3(kBuildThreads), 0, s, rows,
                       count, n, row_ptr);
  uint32_t err = 0;
  ABM_CHECK_HIP(hipMemcpyAsync(&err, w.err, 4, hipMemcpyDeviceToHost, s));
  ABM_CHECK_HIP(hipStreamSynchronize(s));
  ABM_LAUNCH_CHECK("abm_csr_from_edges");
  if (err) {
    set_error("abm_csr_from_edges: edge endpoint out of range [0, n)");
    return kErrInvalid;
  }
  *nnz = count;
  return kOk;
}

}  // extern "C"
