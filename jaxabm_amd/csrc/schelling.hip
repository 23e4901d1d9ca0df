// schelling.hip -- Schelling segregation on a periodic Grid.
//
// Program (reference Grid stores only shape and the periodic flag,
// jaxabm/agentpy.py:465-527; step schedule jaxabm/model.py:146-216):
//  init   u = uniform(PRNGKey(seed), (H*W,)); cell type 1 if u < 0.45, 2 if
//         u < 0.9, else empty; agent k = k-th occupied cell in row-major order
//  step   empty list = empty cells of the pre-step grid in row-major order;
//         agent phase (agent keys = split(ck, n_agents)): Moore-8 on the torus,
//         happy iff 8 * same >= 3 * occupied (same/occupied >= 0.375, an
//         isolated agent is happy); an unhappy agent draws
//         j = randint(agent_key, (), 0, n_empty) and claims empty[j];
//         env phase: each claimed cell goes to its lowest-id claimant, winners
//         move (their old cell is not reusable until the next step)
//  record fraction happy (agent phase) and number of moves
//
// The empty cells are kept as a bitmap (bit c set when cell c is empty) and a
// count per 1024-cell block; the moves kernel keeps both current with atomics
// (a bit flip per moved agent, and a count update when a move crosses blocks;
// a cell changes state at most once per step), so the row-major empty list is
// never materialised: every agents block scans the block counts (1 KB at
// 512 x 512) into an LDS prefix, and empty[j] is then a search of that prefix
// and a popcount scan of one block's 32 bitmap words.  Two kernels per step
// (captured with the rest of the run):
//   k_schelling_agents  32x32 cell tiles + 1-cell torus halo in LDS;
//                       happiness, draw, empty[j], atomicMin claim
//   k_schelling_moves   winners move; bitmap, block counts and claims updated
// Layout (HBM): cells int8[H*W], id int32[H*W] (agent on a cell or -1),
// claim int32[H*W] (INT32_MAX when unclaimed), pos int32[n_agents],
// type int8[n_agents], target int32[n_agents], empty-cell bitmap
// uint32[ceil(H*W/32)] (in the grid's `empty` scratch), empty cells per
// 1024-cell block int32[8][ceil(H*W/1024)] (8 copies summed by readers; in
// the `block_counts` scratch).
#include <hip/hip_runtime.h>

#include <climits>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kTileDim = 32;
constexpr int kCellsPerBlock = 1024;
constexpr int kWordsPerBlock = kCellsPerBlock / 32;
// block prefix of the empty counts in LDS: with the agents kernel's 5.3 KB of
// static LDS, at most 64 KB in all
constexpr int64_t kMaxEmptyBlocks = 15000;
// copies of the per-block empty counts: the moves kernel's workgroup w updates
// copy w % kCountCopies, so the count atomics of a busy step spread over 16x
// the words (each word's device-scope atomics serialise); readers sum them
constexpr int kCountCopies = 16;

// number of cells with (cells[c] == 0) == want_empty per 1024-cell block
__global__ void __launch_bounds__(kCellsPerBlock)
    k_cells_count(const int8_t *__restrict__ cells, int64_t nc, int want_empty,
                  int32_t *__restrict__ block_counts) {
  const int64_t c = (int64_t)blockIdx.x * kCellsPerBlock + threadIdx.x;
  const bool hit = c < nc && ((cells[c] == 0) == (want_empty != 0));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ int32_t wc[16];
  const uint64_t m = __ballot(hit);
  if (lane == 0) wc[wave] = (int32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < 16; ++w) s += wc[w];
    block_counts[blockIdx.x] = s;
  }
}

// Writes the matching cells in row-major order: out[rank] = cell.  When
// agent arrays are given (init), also numbers the agents: pos, type, id.
__global__ void __launch_bounds__(kCellsPerBlock)
    k_cells_scatter(const int8_t *__restrict__ cells, int64_t nc, int want_empty,
                    const int32_t *__restrict__ block_counts, int32_t *__restrict__ out,
                    int8_t *__restrict__ type, int32_t *__restrict__ id) {
  __shared__ int32_t base;
  __shared__ int32_t wc[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // block offset = sum of the counts of earlier blocks
  int32_t part = 0;
  for (int32_t b = tid; b < (int32_t)blockIdx.x; b += kCellsPerBlock) part += block_counts[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if (lane == 0) wc[wave] = part;
  __syncthreads();
  if (tid == 0) {
    int32_t s = 0;
    for (int w = 0; w < 16; ++w) s += wc[w];
    base = s;
  }
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * kCellsPerBlock + tid;
  const bool hit = c < nc && ((cells[c] == 0) == (want_empty != 0));
  const uint64_t m = __ballot(hit);
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  __syncthreads();
  if (lane == 0) wc[wave] = (int32_t)__popcll(m);
  __syncthreads();
  int32_t before = base;
  for (int w = 0; w < wave; ++w) before += wc[w];
  if (hit) {
    const int32_t r = before + (int32_t)__popcll(m & lt);
    out[r] = (int32_t)c;
    if (type) type[r] = cells[c];
    if (id) id[c] = r;
  } else if (id && c < nc) {
    id[c] = -1;
  }
}

// Empty-cell bitmap of the current grid: one ballot per 64 cells.
__global__ void __launch_bounds__(kCellsPerBlock)
    k_empty_bits(const int8_t *__restrict__ cells, int64_t nc, int64_t nwords,
                 uint32_t *__restrict__ ebits) {
  const int64_t c = (int64_t)blockIdx.x * kCellsPerBlock + threadIdx.x;
  const bool e = c < nc && cells[c] == 0;
  const uint64_t m = __ballot(e);
  if ((threadIdx.x & 63) == 0) {
    const int64_t w = c >> 5;  // c is the wave's first cell (a multiple of 64)
    if (w < nwords) ebits[w] = (uint32_t)m;
    if (w + 1 < nwords) ebits[w + 1] = (uint32_t)(m >> 32);
  }
}

// Empty cells of block b < nb: the sum of its count copies.
__device__ __forceinline__ int32_t block_empty(const int32_t *__restrict__ ecount, int nb,
                                               int b) {
  int32_t v[kCountCopies];
#pragma unroll
  for (int r = 0; r < kCountCopies; ++r) v[r] = ecount[(int64_t)r * nb + b];
  int32_t s = 0;
#pragma unroll
  for (int r = 0; r < kCountCopies; ++r) s += v[r];
  return s;
}

// P[0] = 0, P[b + 1] = empty cells of blocks 0 .. b, for b < nb, in LDS, from
// the per-block counts (wave scans).  first: this thread's count for the
// first chunk of blockDim.x blocks, loaded by the caller ahead of its other
// loads.  Up to blockDim.x blocks (grids of up to 1M cells at 1024 threads)
// the scan takes two barriers; larger grids scan chunk by chunk.
__device__ void empty_prefix(const int32_t *__restrict__ ecount, int nb, int32_t first,
                             int32_t *P, int32_t *wsum, int32_t *carry) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nthreads = (int)blockDim.x;
  if (nb <= nthreads) {
    int32_t v = first;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(v, off);
      if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wave] = v;
    if (tid == 0) P[0] = 0;
    __syncthreads();
    if (tid < nb) {
      int32_t before = 0;
      for (int q = 0; q < wave; ++q) before += wsum[q];
      P[tid + 1] = before + v;
    }
    __syncthreads();
    return;
  }
  if (tid == 0) {
    P[0] = 0;
    *carry = 0;
  }
  __syncthreads();
  for (int base = 0; base < nb; base += nthreads) {
    const int b = base + tid;
    int32_t v = base == 0 ? first : (b < nb ? block_empty(ecount, nb, b) : 0);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(v, off);
      if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    int32_t before = *carry;
    for (int q = 0; q < wave; ++q) before += wsum[q];
    if (b < nb) P[b + 1] = before + v;
    __syncthreads();
    if (tid == nthreads - 1) *carry = before + v;
    __syncthreads();
  }
}

// Cell of the j-th empty cell in row-major order (j < P[nb]): the block b with
// P[b] <= j < P[b + 1], then the (j - P[b])-th set bit of its 32 bitmap words.
__device__ __forceinline__ int32_t jth_empty(uint32_t j, const int32_t *P, int nb,
                                             const uint32_t *__restrict__ ebits,
                                             int64_t nwords) {
  int lo = 0, hi = nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((uint32_t)P[mid] <= j)
      lo = mid;
    else
      hi = mid;
  }
  uint32_t r = j - (uint32_t)P[lo];
  const int64_t w0 = (int64_t)lo * kWordsPerBlock;
  uint32_t wv[kWordsPerBlock];
#pragma unroll
  for (int q = 0; q < kWordsPerBlock; ++q) wv[q] = w0 + q < nwords ? ebits[w0 + q] : 0u;
  int32_t cell = -1;
#pragma unroll
  for (int q = 0; q < kWordsPerBlock; ++q) {
    const uint32_t pc = (uint32_t)__popc(wv[q]);
    if (cell < 0) {
      if (r < pc) {
        uint32_t v = wv[q];
        for (uint32_t k = 0; k < r; ++k) v &= v - 1;
        cell = (int32_t)((w0 + q) * 32 + __builtin_ctz(v));
      } else {
        r -= pc;
      }
    }
  }
  return cell;
}

struct SchellingStep {
  int32_t H, W;
  int64_t n_agents, n_empty;
  const int8_t *cells;
  const int32_t *id;
  const uint32_t *ebits;
  const int32_t *ecount;
  int64_t nwords;
  int32_t nb;
  int32_t *claim, *target;
  const uint32_t *ck;
  int mode;
  int32_t *happy_count;
  // the previous step's happy count (null at a run's first step): when every
  // agent was happy then, nobody moved, the grid is unchanged and this step
  // is the identity
  const int32_t *prev_happy;
};

// True when the step is the identity: the previous step found every agent
// happy, so nothing moved since and every agent is happy again.
__device__ __forceinline__ bool schelling_identity(const int32_t *prev_happy, int64_t n_agents) {
  return prev_happy && (int64_t)*prev_happy == n_agents;
}

__global__ void __launch_bounds__(kTileDim * kTileDim) k_schelling_agents(SchellingStep a) {
  extern __shared__ int32_t empty_prefix_lds[];  // [nb + 1]
  __shared__ int8_t tile[kTileDim + 2][kTileDim + 2];
  __shared__ int32_t wc[16];
  __shared__ int32_t carry;
  __shared__ int32_t work[kTileDim * kTileDim];  // the tile's agents that draw a target
  __shared__ int32_t nwork;
  constexpr int kHalo = (kTileDim + 2) * (kTileDim + 2);  // 34 x 34 cells, <= 2 per thread
  const int tid = threadIdx.x;
  if (schelling_identity(a.prev_happy, a.n_agents)) {  // grid-uniform
    // every agent happy again; no target is drawn (the last step left each
    // agent's target at -1) and the moves kernel exits too
    if (blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) *a.happy_count = (int32_t)a.n_agents;
    return;
  }
  const int tx = tid % kTileDim, ty = tid / kTileDim;
  const int r0 = blockIdx.y * kTileDim, c0 = blockIdx.x * kTileDim;
  const int r = r0 + ty, c = c0 + tx;
  const bool inside = r < a.H && c < a.W;
  const int64_t cell = (int64_t)r * a.W + c;
  // one memory round trip: the tile and its torus halo, this thread's block
  // count for the prefix, the agent on this thread's cell
  int8_t tv[2] = {0, 0};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + i * (int)blockDim.x;
    if (q < kHalo) {
      const int lr = q / (kTileDim + 2), lc = q % (kTileDim + 2);
      int gr = r0 + lr - 1, gc = c0 + lc - 1;
      gr = (gr % a.H + a.H) % a.H;
      gc = (gc % a.W + a.W) % a.W;
      tv[i] = a.cells[(int64_t)gr * a.W + gc];
    }
  }
  const int32_t first = tid < a.nb ? block_empty(a.ecount, a.nb, tid) : 0;
  const int32_t k = inside ? a.id[cell] : -1;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + i * (int)blockDim.x;
    if (q < kHalo) tile[q / (kTileDim + 2)][q % (kTileDim + 2)] = tv[i];
  }
  if (tid == 0) nwork = 0;  // visible after the prefix's barriers
  empty_prefix(a.ecount, a.nb, first, empty_prefix_lds, wc, &carry);
  bool happy = false, draws = false;
  if (inside) {
    const int8_t me = tile[ty + 1][tx + 1];
    if (me) {
      int occ = 0, same = 0;
#pragma unroll
      for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc) {
          if (!dr && !dc) continue;
          const int8_t v = tile[ty + 1 + dr][tx + 1 + dc];
          occ += v != 0;
          same += v == me;
        }
      happy = 8 * same >= 3 * occ;
      draws = !happy && a.n_empty > 0;
      if (!draws) a.target[k] = -1;
    }
  }
  // The agents that draw go to the front of the workgroup, so that the draw
  // and the search run on full waves, not on every wave that holds one.
  const int lane = tid & 63;
  const uint64_t dm = __ballot(draws);
  if (dm) {  // wave-uniform
    int32_t base = 0;
    if (lane == 0) base = atomicAdd(&nwork, (int32_t)__popcll(dm));
    base = __shfl(base, 0);
    if (draws) work[base + (int32_t)__popcll(dm & ((1ull << lane) - 1ull))] = k;
  }
  __syncthreads();
  if (tid < nwork) {
    const int32_t kd = work[tid];
    Key k1, k2;
    split2(split_at(Key{a.ck[0], a.ck[1]}, a.n_agents, kd, a.mode), k1, k2, a.mode);
    const uint32_t j = randint_offset(bits_scalar(k1, a.mode), bits_scalar(k2, a.mode),
                                      (uint32_t)a.n_empty);
    const int32_t tg = jth_empty(j, empty_prefix_lds, a.nb, a.ebits, a.nwords);
    if (tg >= 0) atomicMin(&a.claim[tg], kd);
    a.target[kd] = tg;
  }
  const uint64_t m = __ballot(happy);
  if ((tid & 63) == 0) wc[tid >> 6] = (int32_t)__popcll(m);
  __syncthreads();
  if (tid == 0) {
    int32_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += wc[w];
    if (s) atomicAdd(a.happy_count, s);
  }
}

// one thread per agent; large workgroups keep the grid small, which is most
// of the cost of a step that only checks the identity condition
constexpr int kMovesThreads = 1024;

__global__ void __launch_bounds__(kMovesThreads) k_schelling_moves(int64_t n_agents, int8_t *__restrict__ cells,
                                  int32_t *__restrict__ id, int32_t *__restrict__ claim,
                                  int32_t *__restrict__ pos, const int8_t *__restrict__ type,
                                  const int32_t *__restrict__ target, uint32_t *__restrict__ ebits,
                                  int32_t *__restrict__ ecount, int32_t nb,
                                  int32_t *__restrict__ moves,
                                  const int32_t *__restrict__ prev_happy) {
  if (schelling_identity(prev_happy, n_agents)) return;  // grid-uniform: no move
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool moved = false;
  if (k < n_agents) {
    const int32_t tg = target[k];
    if (tg >= 0 && claim[tg] == (int32_t)k) {
      const int32_t old = pos[k];
      cells[old] = 0;
      id[old] = -1;
      cells[tg] = type[k];
      id[tg] = (int32_t)k;
      pos[k] = tg;
      claim[tg] = INT_MAX;
      // the old cell becomes empty, the target occupied (distinct cells: a
      // target was empty before the step, an old cell occupied)
      atomicOr(&ebits[old >> 5], 1u << (old & 31));
      atomicAnd(&ebits[tg >> 5], ~(1u << (tg & 31)));
      const int32_t bo = old / kCellsPerBlock, bt = tg / kCellsPerBlock;
      if (bo != bt) {
        int32_t *copy = ecount + (int64_t)(blockIdx.x % kCountCopies) * nb;
        atomicAdd(&copy[bo], 1);
        atomicSub(&copy[bt], 1);
      }
      moved = true;
    }
  }
  __shared__ int32_t wc[kMovesThreads / 64];
  const uint64_t m = __ballot(moved);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (int32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < kMovesThreads / 64; ++w) s += wc[w];
    if (s) atomicAdd(moves, s);
  }
}

// record[t] = (happy fraction, moves) from the integer counters
__global__ void k_schelling_record(const int32_t *__restrict__ counters, int32_t steps,
                                   int64_t n_agents, float *__restrict__ happy_frac,
                                   int32_t *__restrict__ moves) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= steps) return;
  happy_frac[t] = (float)counters[2 * t] / (float)n_agents;
  moves[t] = counters[2 * t + 1];
}

// settled steps: each is the identity when the step before them found every
// agent happy; otherwise the records are marked invalid (-1, NaN, -1)
__global__ void k_schelling_settled(const int32_t *__restrict__ prev_counters, int32_t steps,
                                    int64_t n_agents, int32_t *__restrict__ counters,
                                    float *__restrict__ happy_frac,
                                    int32_t *__restrict__ moves) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= steps) return;
  const bool ok = (int64_t)prev_counters[0] == n_agents;
  counters[2 * t] = ok ? (int32_t)n_agents : -1;
  counters[2 * t + 1] = ok ? 0 : -1;
  happy_frac[t] = ok ? (float)n_agents / (float)(n_agents > 0 ? n_agents : 1)
                     : __int_as_float(0x7fc00000);
  moves[t] = ok ? 0 : -1;
}

__global__ void k_schelling_init_cells(const uint32_t *__restrict__ key, int64_t nc,
                                       int8_t *__restrict__ cells, int32_t *__restrict__ claim,
                                       int mode) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const float u = unit_float(bits_at(Key{key[0], key[1]}, nc, c, mode));
  cells[c] = u < 0.45f ? 1 : (u < 0.9f ? 2 : 0);
  claim[c] = INT_MAX;
}

__global__ void k_fill_i32_s(int32_t *__restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int check_grid(const abm_schelling_grid *g, const char *who) {
  if (!g || g->height < 1 || g->width < 1 ||
      (int64_t)g->height * g->width >= (1ll << 31) || !g->cells || !g->id || !g->claim ||
      !g->pos || !g->type || !g->target || !g->empty || !g->block_counts ||
      g->n_agents < 0 || g->n_agents > (int64_t)g->height * g->width) {
    set_error(std::string(who) + ": bad grid descriptor");
    return kErrInvalid;
  }
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int64_t abm_schelling_block_count(int64_t n_cells) {
  return kCountCopies * ((n_cells + kCellsPerBlock - 1) / kCellsPerBlock);
}

int abm_schelling_init(abm_schelling_grid *g, const uint32_t *key, int mode,
                       int32_t *n_agents_out, void *stream) {
  if (!g || g->height < 1 || g->width < 1) {
    set_error("abm_schelling_init: bad grid");
    return kErrInvalid;
  }
  ABM_REQUIRE(key && n_agents_out && (mode == 0 || mode == 1),
              "abm_schelling_init: bad arguments");
  const int64_t nc = (int64_t)g->height * g->width;
  g->n_agents = 0;
  if (int rc = check_grid(g, "abm_schelling_init")) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_schelling_init_cells, dim3(nblk(nc, 256)), dim3(256), 0, s, key, nc,
                     g->cells, g->claim, mode);
  const unsigned nb = nblk(nc, kCellsPerBlock);
  hipLaunchKernelGGL(k_cells_count, dim3(nb), dim3(kCellsPerBlock), 0, s, g->cells, nc, 0,
                     g->block_counts);
  hipLaunchKernelGGL(k_cells_scatter, dim3(nb), dim3(kCellsPerBlock), 0, s, g->cells, nc, 0,
                     g->block_counts, g->pos, g->type, g->id);
  ABM_LAUNCH_CHECK("abm_schelling_init");
  // agent count: sum of the block counts (host read; set-up only)
  int32_t *host = new int32_t[nb];
  hipError_t e = hipMemcpyAsync(host, g->block_counts, nb * sizeof(int32_t),
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete[] host;
    return hip_fail(e, "abm_schelling_init");
  }
  int64_t na = 0;
  for (unsigned b = 0; b < nb; ++b) na += host[b];
  delete[] host;
  g->n_agents = na;
  *n_agents_out = (int32_t)na;
  return kOk;
}

}  // extern "C"

// the step loop of a run; prev0: the counters of the step before the first
// (nullptr: a fresh run, whose bitmap and block counts are built from the
// cells first)
static int schelling_steps(const abm_schelling_grid *g, const uint32_t *keys, int32_t steps,
                           int mode, const int32_t *prev0, int32_t *counters, float *happy_frac,
                           int32_t *moves, void *stream, const char *what) {
  if (int rc = check_grid(g, what)) return rc;
  ABM_REQUIRE(keys && counters && happy_frac && moves && steps >= 0 &&
                  (mode == 0 || mode == 1),
              "abm_schelling_run: bad arguments");
  if (steps == 0) return kOk;
  hipStream_t s = (hipStream_t)stream;
  const int64_t nc = (int64_t)g->height * g->width;
  const int64_t n_empty = nc - g->n_agents;
  const unsigned nb = nblk(nc, kCellsPerBlock);
  ABM_REQUIRE(nb <= kMaxEmptyBlocks,
              "abm_schelling_run: grids of more than 15,360,000 cells are not supported");
  const int64_t nwords = (nc + 31) / 32;
  uint32_t *ebits = reinterpret_cast<uint32_t *>(g->empty);  // nwords <= H*W words
  ABM_REQUIRE(((uintptr_t)ebits & 15u) == 0,
              "abm_schelling_run: the empty scratch must be 16-byte aligned");
  hipLaunchKernelGGL(k_fill_i32_s, dim3(nblk(2 * (int64_t)steps, 256)), dim3(256), 0, s,
                     counters, 2 * (int64_t)steps, 0);
  if (!prev0) {
    hipLaunchKernelGGL(k_empty_bits, dim3(nb), dim3(kCellsPerBlock), 0, s, g->cells, nc, nwords,
                       ebits);
    // copy 0 = the grid's empty counts, the other copies start at zero
    hipLaunchKernelGGL(k_cells_count, dim3(nb), dim3(kCellsPerBlock), 0, s, g->cells, nc, 1,
                       g->block_counts);
    const int64_t zero_words = (int64_t)(kCountCopies - 1) * nb;
    hipLaunchKernelGGL(k_fill_i32_s, dim3(nblk(zero_words, 256)), dim3(256), 0, s,
                       g->block_counts + nb, zero_words, 0);
  }
  const dim3 tiles((unsigned)((g->width + kTileDim - 1) / kTileDim),
                   (unsigned)((g->height + kTileDim - 1) / kTileDim));
  const size_t prefix_lds = (size_t)(nb + 1) * sizeof(int32_t);
  for (int32_t t = 0; t < steps; ++t) {
    SchellingStep a;
    a.H = g->height;
    a.W = g->width;
    a.n_agents = g->n_agents;
    a.n_empty = n_empty;
    a.cells = g->cells;
    a.id = g->id;
    a.ebits = ebits;
    a.ecount = g->block_counts;
    a.nwords = nwords;
    a.nb = (int32_t)nb;
    a.claim = g->claim;
    a.target = g->target;
    a.ck = keys + 2 * (int64_t)t;
    a.mode = mode;
    a.happy_count = counters + 2 * (int64_t)t;
    a.prev_happy = t ? counters + 2 * (int64_t)(t - 1) : prev0;
    hipLaunchKernelGGL(k_schelling_agents, tiles, dim3(kTileDim * kTileDim), prefix_lds, s, a);
    if (g->n_agents > 0)
      hipLaunchKernelGGL(k_schelling_moves, dim3(nblk(g->n_agents, kMovesThreads)),
                         dim3(kMovesThreads), 0, s,
                         g->n_agents, g->cells, g->id, g->claim, g->pos, g->type, g->target,
                         ebits, g->block_counts, (int32_t)nb, counters + 2 * (int64_t)t + 1,
                         a.prev_happy);
  }
  hipLaunchKernelGGL(k_schelling_record, dim3(nblk(steps, 256)), dim3(256), 0, s, counters,
                     steps, g->n_agents > 0 ? g->n_agents : 1, happy_frac, moves);
  ABM_LAUNCH_CHECK(what);
  return kOk;
}

extern "C" {

int abm_schelling_run(const abm_schelling_grid *g, const uint32_t *keys, int32_t steps,
                      int mode, int32_t *counters, float *happy_frac, int32_t *moves,
                      void *stream) {
  return schelling_steps(g, keys, steps, mode, nullptr, counters, happy_frac, moves, stream,
                         "abm_schelling_run");
}

int abm_schelling_run_from(const abm_schelling_grid *g, const uint32_t *keys, int32_t steps,
                           int mode, const int32_t *prev_counters, int32_t *counters,
                           float *happy_frac, int32_t *moves, void *stream) {
  ABM_REQUIRE(prev_counters, "abm_schelling_run_from: prev_counters is null");
  return schelling_steps(g, keys, steps, mode, prev_counters, counters, happy_frac, moves,
                         stream, "abm_schelling_run_from");
}

int abm_schelling_settled(const abm_schelling_grid *g, const int32_t *prev_counters, int32_t steps,
                          int32_t *counters, float *happy_frac, int32_t *moves, void *stream) {
  if (int rc = check_grid(g, "abm_schelling_settled")) return rc;
  ABM_REQUIRE(prev_counters && counters && happy_frac && moves && steps >= 0,
              "abm_schelling_settled: bad arguments");
  if (steps == 0) return kOk;
  hipLaunchKernelGGL(k_schelling_settled, dim3(nblk(steps, 256)), dim3(256), 0,
                     (hipStream_t)stream, prev_counters, steps, g->n_agents, counters, happy_frac,
                     moves);
  ABM_LAUNCH_CHECK("abm_schelling_settled");
  return kOk;
}

}  // extern "C"
