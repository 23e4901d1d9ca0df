// grid.hip -- Moore-8 neighbourhood reductions over a Grid.
//
// The reference Grid keeps only its shape and periodic flag in the model
// environment (jaxabm/agentpy.py:465-527); models compute neighbourhoods in
// their own jnp code over a (rows, cols) array.  These kernels give any such
// model the Moore-8 reductions in one launch:
//
//   sum      out[r][c] = sum over the 8 offsets (dr, dc) != (0, 0) of
//            x[r+dr][c+dc]: float64 accumulation in the fixed order
//            (-1,-1) (-1,0) (-1,1) (0,-1) (0,1) (1,-1) (1,0) (1,1), rounded once
//   counts   same[r][c]     = neighbours nb with cells[nb] != 0 and
//                             cells[nb] == cells[r][c]
//            occupied[r][c] = neighbours nb with cells[nb] != 0
//
// Periodic grids wrap both axes (a torus, the offsets taken modulo the
// shape); bounded grids leave out the offsets that fall outside.  Each
// workgroup stages a 32 x 32 tile plus a one-cell halo in LDS (34 x 34), so
// every value is read from HBM about 1.13 times instead of 9.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kGT = 32;           // tile edge
constexpr int kGH = kGT + 2;      // with halo
constexpr int kGThreads = kGT * kGT;

// Cell (r, c) of a tile with halo, r in [-1, rows] and c in [-1, cols] for
// the cells some valid cell reads (a tile can reach past the grid's last
// row or column: those cells are never read and are not loaded).
__device__ __forceinline__ bool halo_cell(int64_t rows, int64_t cols, bool periodic, int64_t r,
                                          int64_t c, int64_t *idx) {
  if (r < -1 || r > rows || c < -1 || c > cols) return false;
  if (periodic) {
    r = r < 0 ? r + rows : (r >= rows ? r - rows : r);
    c = c < 0 ? c + cols : (c >= cols ? c - cols : c);
  } else if (r < 0 || r >= rows || c < 0 || c >= cols) {
    return false;
  }
  *idx = r * cols + c;
  return true;
}

__global__ void __launch_bounds__(kGThreads)
    k_moore_sum(int64_t rows, int64_t cols, int periodic, const float *__restrict__ x,
                float *__restrict__ out) {
  __shared__ float tile[kGH][kGH + 1];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kGT - 1, c0 = (int64_t)blockIdx.x * kGT - 1;
  for (int i = tid; i < kGH * kGH; i += kGThreads) {
    const int ly = i / kGH, lx = i - ly * kGH;
    int64_t g;
    tile[ly][lx] = halo_cell(rows, cols, periodic != 0, r0 + ly, c0 + lx, &g) ? x[g] : 0.0f;
  }
  __syncthreads();
  const int ty = tid / kGT, tx = tid - ty * kGT;
  const int64_t r = r0 + 1 + ty, c = c0 + 1 + tx;
  if (r >= rows || c >= cols) return;
  // bounded grids: staged zeros stand in for the missing neighbours (adding
  // +0.0 leaves a float64 sum unchanged, -0.0 aside; the result is rounded
  // from the same float64 value either way)
  double s = 0.0;
  s += (double)tile[ty][tx];
  s += (double)tile[ty][tx + 1];
  s += (double)tile[ty][tx + 2];
  s += (double)tile[ty + 1][tx];
  s += (double)tile[ty + 1][tx + 2];
  s += (double)tile[ty + 2][tx];
  s += (double)tile[ty + 2][tx + 1];
  s += (double)tile[ty + 2][tx + 2];
  out[r * cols + c] = (float)s;
}

__global__ void __launch_bounds__(kGThreads)
    k_moore_counts(int64_t rows, int64_t cols, int periodic, const int8_t *__restrict__ cells,
                   int32_t *__restrict__ same, int32_t *__restrict__ occupied) {
  __shared__ int8_t tile[kGH][kGH + 2];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kGT - 1, c0 = (int64_t)blockIdx.x * kGT - 1;
  for (int i = tid; i < kGH * kGH; i += kGThreads) {
    const int ly = i / kGH, lx = i - ly * kGH;
    int64_t g;
    tile[ly][lx] = halo_cell(rows, cols, periodic != 0, r0 + ly, c0 + lx, &g) ? cells[g] : 0;
  }
  __syncthreads();
  const int ty = tid / kGT, tx = tid - ty * kGT;
  const int64_t r = r0 + 1 + ty, c = c0 + 1 + tx;
  if (r >= rows || c >= cols) return;
  const int8_t me = tile[ty + 1][tx + 1];
  int32_t ns = 0, no = 0;
#pragma unroll
  for (int dy = 0; dy < 3; ++dy)
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
      if (dy == 1 && dx == 1) continue;
      const int8_t v = tile[ty + dy][tx + dx];
      no += v != 0;
      ns += (v != 0) & (v == me);
    }
  const int64_t g = r * cols + c;
  if (same) same[g] = ns;
  if (occupied) occupied[g] = no;
}

int grid_dims(int64_t rows, int64_t cols, dim3 *grid) {
  ABM_REQUIRE(rows >= 1 && cols >= 1 && rows * cols < (1ll << 40) &&
                  (rows + kGT - 1) / kGT < 65536 && (cols + kGT - 1) / kGT < (1ll << 31),
              "abm_grid_moore: bad grid shape");
  *grid = dim3((unsigned)((cols + kGT - 1) / kGT), (unsigned)((rows + kGT - 1) / kGT));
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_grid_moore_sum_f32(int64_t rows, int64_t cols, int32_t periodic, const float *x,
                           float *out, void *stream) {
  dim3 g;
  if (int rc = grid_dims(rows, cols, &g)) return rc;
  ABM_REQUIRE(x && out && x != out, "abm_grid_moore_sum_f32: null or aliased buffer");
  hipLaunchKernelGGL(k_moore_sum, g, dim3(kGThreads), 0, (hipStream_t)stream, rows, cols,
                     periodic ? 1 : 0, x, out);
  ABM_LAUNCH_CHECK("abm_grid_moore_sum_f32");
  return kOk;
}

int abm_grid_moore_counts_i8(int64_t rows, int64_t cols, int32_t periodic, const int8_t *cells,
                             int32_t *same, int32_t *occupied, void *stream) {
  dim3 g;
  if (int rc = grid_dims(rows, cols, &g)) return rc;
  ABM_REQUIRE(cells && (same || occupied), "abm_grid_moore_counts_i8: null buffer");
  hipLaunchKernelGGL(k_moore_counts, g, dim3(kGThreads), 0, (hipStream_t)stream, rows, cols,
                     periodic ? 1 : 0, cells, same, occupied);
  ABM_LAUNCH_CHECK("abm_grid_moore_counts_i8");
  return kOk;
}

}  // extern "C"
