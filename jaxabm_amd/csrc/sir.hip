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
// Design for MI355X:
//  * Neighbour status comes from an infected-set bitmap (1 bit per agent,
//    125 KB at n = 1M, word count a multiple of 4) written by the previous
//    step.  Each of the 256 persistent 1024-thread blocks stages the whole
//    bitmap in LDS (160 KB per CU), so the random neighbour checks are LDS bit
//    tests instead of global gathers; graphs whose bitmap exceeds the LDS
//    budget read it from global memory (L2-resident) with the same kernel.
//    Staging costs a 125 KB read per block, so a block stages only when its
//    tile has many rows to scan (the epidemic peak) and otherwise tests bits
//    in the L2-resident global bitmap.
//  * A step whose previous step left no infected agent is the identity: the
//    kernel then only clears its bitmap spans and copies the counts.
//  * Work skipping with plain stores (no atomics): an agent u that is infected
//    after step s writes mark[w] = s + H (16-bit) into each neighbour w when
//    it was infected at s or (s + u) % H == 0.  At step t an S agent scans its
//    row only if (mark[v] - t) mod 2^16 <= H.  If v has a neighbour u infected
//    at the start of step t, u wrote a mark in some step s in [t-H, t-1] (its
//    infection step or its refresh phase) with value s + H in [t, t+H-1], and
//    later writes only store values in [t, t+H], so v is never skipped; stale
//    marks only cost a scan.  H = 16 with per-agent refresh phases cuts the
//    mark stores of long-lived infections by 16x without step-to-step spikes.
//  * Only its owner reads an agent's state byte, so state is updated in place
//    (an agent octet is written only when it changes); the bitmap is double-buffered (read
//    B[t&1], write B[(t+1)&1]) and each bitmap word is written whole by the
//    block owning its 64-agent span.
//  * The agent range is cut into 64-aligned chunks of equal (agents + degree
//    + heavy-row) cost and at most one tile, one per block.  A chunk is processed in tiles
//    of 8192 agents (eight per thread, loads issued together): S agents at risk go to an LDS
//    worklist; rows of degree <= 32 are scanned and marked by one lane, longer
//    rows by a whole wavefront (4 x 64 entries per round).
//  * S/I/R counts: wave reductions -> LDS -> one atomic per block into one of
//    64 slots, reduced after the run.
//
// Layout (HBM): state uint8[n] (low 2 bits S=0/I=1/R=2, bit 2 = row degree
// > 32), mark uint16[n], bitmap uint32[2][words], row_ptr int32[n+1],
// col int32[nnz] (rows ascending), chunk_ptr int32[n_chunks+1].
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kSirThreads = 1024;
constexpr int kPerThread = 8;
constexpr int kTile = kSirThreads * kPerThread;
constexpr int kSirBlocks = 256;
constexpr int kHeavyDeg = 32;
constexpr int kChunkAlign = 64;
constexpr int64_t kChunkNodeCost = 4;
// extra chunk cost of a heavy row: a wavefront pays about two serial memory
// round trips per row whatever its degree
constexpr int64_t kChunkHeavyCost = 128;
constexpr int kChunksPerBlock = 1;
constexpr int kSlots = 64;                 // count slots per step
constexpr int kStepSlots = kSlots + 1;     // + one flag slot
constexpr int kStride = kStepSlots * 3;    // ints per step in the slots array
// slots[t][kSlots][1] = 1 when step t left an infected agent (plain stores
// of the same value by any block that has one): the next step reads this one
// word to decide whether it is the identity
constexpr int kLiveFlag = kSlots * 3 + 1;
constexpr uint8_t kHeavyBit = 4;
constexpr uint32_t kMarkHorizon = 16;
// a tile stages the bitmap in LDS when it has at least this many rows to scan
constexpr int32_t kStageItems = 512;
// LDS bitmap budget: 32768 words (1,048,576 agents, 128 KB)
constexpr int64_t kLdsBitmapWords = 32768;

struct SirStep {
  int64_t n, nnz;
  const int32_t *row_ptr, *col, *chunk_ptr;
  int32_t n_chunks;
  uint8_t *state;
  uint16_t *mark;
  const uint32_t *bm_in;
  uint32_t *bm_out;
  int64_t words;
  uint32_t t;
  const int32_t *prev_slots;  // previous step's count slots, or null
  const int32_t *prev2_slots; // the step before that, or null
  const uint32_t *keys;
  uint32_t thr_beta, thr_gamma;
  int mode;
  int32_t *slots;
};

// Workgroup barrier that orders LDS only.  Every intra-block exchange in the
// step kernel goes through LDS (state, marks and bitmaps are never read back
// by another thread of the block in the same step), so the barrier need not
// wait for outstanding global stores the way __syncthreads() does.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ bool infected(const uint32_t *__restrict__ bm, int32_t u) {
  return (bm[u >> 5] >> (u & 31)) & 1u;
}

__device__ __forceinline__ bool at_risk(uint16_t m, uint32_t t) {
  return (uint16_t)(m - (uint16_t)t) <= kMarkHorizon;
}

// Rows of degree <= 32 are scanned by one lane, kScanGroup rows at a time:
// the group's row bounds load together, then each round loads the next four
// entries of every row still scanning together.  The S-I candidates of a
// round are drawn in order (one draw site); a row stops at its first
// infecting entry.  hitm bit g: row g of the group is infected.  One row per
// group is the fastest on MI355X (1M-agent BA run: 8.44 us per step against
// 9.07 with two rows and 10.15 with four): a busy step's scans are bound by
// the draws and the slowest lane, not by the lanes' memory round trips.
constexpr int kScanGroup = 1;

__device__ __forceinline__ uint32_t scan_light_group(const int32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ bm,
                                                     int32_t (&e)[kScanGroup],
                                                     const int32_t (&hi)[kScanGroup], Key kinf,
                                                     int64_t nnz, uint32_t thr, int mode) {
  uint32_t hitm = 0;
  for (;;) {
    uint32_t live = 0;
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g) live |= (uint32_t)(e[g] < hi[g]) << g;
    if (!live) break;
    int32_t c[kScanGroup][4];
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g)
#pragma unroll
      for (int k = 0; k < 4; ++k) c[g][k] = e[g] + k < hi[g] ? col[e[g] + k] : -1;
    uint32_t cm = 0;  // bit 4g + k: entry e[g] + k is an S-I candidate
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c[g][k] >= 0 && infected(bm, c[g][k])) cm |= 1u << (4 * g + k);
    while (cm) {
      const int bit = __builtin_ctz(cm);
      cm &= cm - 1;
      const int g = bit >> 2;
      int32_t base = e[0];
#pragma unroll
      for (int q = 1; q < kScanGroup; ++q) base = g == q ? e[q] : base;
      if ((bits_at(kinf, nnz, base + (bit & 3), mode) >> 9) < thr) {
        hitm |= 1u << g;
        cm &= ~(0xFu << (4 * g));
      }
    }
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g) e[g] = (hitm >> g) & 1u ? hi[g] : e[g] + 4;
  }
  return hitm;
}

// Scan of a long row by a whole wavefront (wave-uniform result): 8 x 64
// independent col loads per round.
__device__ __forceinline__ bool scan_heavy(const int32_t *__restrict__ col,
                                           const uint32_t *__restrict__ bm,
                                           int64_t lo, int64_t hi, Key kinf,
                                           int64_t nnz, uint32_t thr, int mode) {
  const int lane = threadIdx.x & 63;
  for (int64_t b = lo; b < hi; b += 8 * 64) {
    int32_t c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t e = b + k * 64 + lane;
      c[k] = e < hi ? col[e] : -1;
    }
    uint32_t cm = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c[k] >= 0 && infected(bm, c[k])) cm |= 1u << k;
    bool hit = false;
    while (cm && !hit) {
      const int k = __builtin_ctz(cm);
      cm &= cm - 1;
      hit = (bits_at(kinf, nnz, b + k * 64 + lane, mode) >> 9) < thr;
    }
    if (__ballot(hit)) return true;
  }
  return false;
}

// Marks of the rows of degree <= 32 of a group (as scan_light_group: the
// next four entries of every row per round, loaded together).
__device__ __forceinline__ void mark_light_group(const int32_t *__restrict__ col,
                                                 uint16_t *__restrict__ mark,
                                                 int32_t (&e)[kScanGroup],
                                                 const int32_t (&hi)[kScanGroup], uint16_t m) {
  for (;;) {
    uint32_t live = 0;
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g) live |= (uint32_t)(e[g] < hi[g]) << g;
    if (!live) break;
    int32_t c[kScanGroup][4];
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g)
#pragma unroll
      for (int k = 0; k < 4; ++k) c[g][k] = e[g] + k < hi[g] ? col[e[g] + k] : -1;
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c[g][k] >= 0) mark[c[g][k]] = m;
      e[g] += 4;
    }
  }
}

// Row bounds of this lane's group of light items items[k0 + g * kSirThreads]
// (g < kScanGroup, k < nl), loaded together; rows past nl are empty.
__device__ __forceinline__ void light_group_rows(const SirStep &a, int64_t sub, int32_t k0,
                                                 int32_t nl, const uint16_t *items,
                                                 int (&loc)[kScanGroup],
                                                 int32_t (&e)[kScanGroup],
                                                 int32_t (&hi)[kScanGroup]) {
#pragma unroll
  for (int g = 0; g < kScanGroup; ++g) {
    const int32_t k = k0 + g * kSirThreads;
    loc[g] = k < nl ? (int)items[k] : -1;
  }
#pragma unroll
  for (int g = 0; g < kScanGroup; ++g) {
    e[g] = 0;
    hi[g] = 0;
    if (loc[g] >= 0) {
      const int64_t u = sub + loc[g];
      e[g] = a.row_ptr[u];
      hi[g] = a.row_ptr[u + 1];
    }
  }
}

__device__ __forceinline__ void mark_heavy(const int32_t *__restrict__ col,
                                           uint16_t *__restrict__ mark,
                                           int64_t lo, int64_t hi, uint16_t m) {
  const int lane = threadIdx.x & 63;
  for (int64_t b = lo; b < hi; b += 8 * 64) {
    int32_t c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t e = b + k * 64 + lane;
      c[k] = e < hi ? col[e] : -1;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c[k] >= 0) mark[c[k]] = m;
  }
}

// Push this thread's agents whose bit is set in lmask (light rows, front of
// the list) / hmask (heavy rows, back of the list); agent j of a thread has
// tile index tid * kPerThread + j.  One ballot per (list, j) and one LDS
// atomic per wave and list.
__device__ __forceinline__ void push_items(uint32_t lmask, uint32_t hmask, uint16_t *items,
                                           int32_t *nl, int32_t *nh) {
  const int tid = threadIdx.x, lane = tid & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint64_t bl[kPerThread], bh[kPerThread];
  int32_t tl = 0, th = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    bl[j] = __ballot((lmask >> j) & 1u);
    bh[j] = __ballot((hmask >> j) & 1u);
    tl += (int32_t)__popcll(bl[j]);
    th += (int32_t)__popcll(bh[j]);
  }
  int32_t fl = 0, fh = 0;
  if (lane == 0) {
    if (tl) fl = atomicAdd(nl, tl);
    if (th) fh = atomicAdd(nh, th);
  }
  fl = __shfl(fl, 0);
  fh = __shfl(fh, 0);
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const uint16_t loc = (uint16_t)(tid * kPerThread + j);
    if ((lmask >> j) & 1u) items[fl + __popcll(bl[j] & lt)] = loc;
    if ((hmask >> j) & 1u) items[kTile - 1 - (fh + __popcll(bh[j] & lt))] = loc;
    fl += (int32_t)__popcll(bl[j]);
    fh += (int32_t)__popcll(bh[j]);
  }
}

// Phase-2 scans of a tile's worklist: heavy rows (back of the list) by
// wavefronts, light rows (front) by lanes; a hit marks the agent infected.
template <int kMode>
__device__ __forceinline__ void scan_items(const SirStep &a, const uint32_t *__restrict__ bm,
                                           int64_t sub, int32_t nl, int32_t nh,
                                           const uint16_t *items, uint8_t *outb, Key kinf,
                                           int32_t *next_heavy) {
  const int tid = threadIdx.x, lane = tid & 63;
  for (;;) {
    int32_t k = 0;
    if (lane == 0) k = atomicAdd(next_heavy, 1);
    k = __shfl(k, 0);
    if (k >= nh) break;
    const int loc = items[kTile - 1 - k];
    const int64_t u = sub + loc;
    const bool hit = scan_heavy(a.col, bm, a.row_ptr[u], a.row_ptr[u + 1], kinf, a.nnz,
                                a.thr_beta, kMode);
    if (lane == 0 && hit) outb[loc] = 1;
  }
  for (int32_t k0 = tid; k0 < nl; k0 += kSirThreads * kScanGroup) {
    int loc[kScanGroup];
    int32_t e[kScanGroup], hi[kScanGroup];
    light_group_rows(a, sub, k0, nl, items, loc, e, hi);
    const uint32_t hitm = scan_light_group(a.col, bm, e, hi, kinf, a.nnz, a.thr_beta, kMode);
#pragma unroll
    for (int g = 0; g < kScanGroup; ++g)
      if ((hitm >> g) & 1u) outb[loc[g]] = 1;
  }
}

template <bool kLds, int kMode>
__global__ void __launch_bounds__(kSirThreads) k_sir_step(SirStep a) {
  extern __shared__ uint4 lds_bm4[];  // kLds: the bitmap words
  __shared__ __align__(16) uint8_t outb[kTile];
  __shared__ __align__(16) uint16_t items[kTile];
  __shared__ int32_t n_light, n_heavy, m_light, m_heavy, next_heavy, next_mark;
  __shared__ int32_t cnt[3];
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid < 3) cnt[tid] = 0;
  // the first chunk's bounds and the previous step's live flag (one
  // uniform word: no block-wide reduction or barrier) load together
  int64_t c0 = a.chunk_ptr[blockIdx.x], c1 = a.chunk_ptr[blockIdx.x + 1];
  const bool any_inf = a.prev_slots ? a.prev_slots[kLiveFlag] != 0 : true;
  if (!any_inf) {
    // identity step.  Its slots stay zero (the count reduction carries the
    // previous step's counts forward).  The next bitmap B[(t+1)&1] is the
    // infected set after step t-2: empty when step t-2 left no infected agent,
    // so only the first identity step clears its spans.
    if (!(a.prev2_slots && a.prev2_slots[kLiveFlag] == 0)) {
      for (int32_t c = blockIdx.x; c < a.n_chunks; c += gridDim.x) {
        const int64_t w0 = (int64_t)a.chunk_ptr[c] >> 5;
        const int64_t w1 = c + 1 == a.n_chunks ? a.words : ((int64_t)a.chunk_ptr[c + 1] >> 5);
        for (int64_t w = w0 + tid; w < w1; w += kSirThreads) a.bm_out[w] = 0u;
      }
    }
    return;
  }
  bool staged = false;
  const Key kinf{a.keys[0], a.keys[1]};
  const Key krec{a.keys[2], a.keys[3]};
  const uint16_t new_mark = (uint16_t)(a.t + kMarkHorizon);
  uint32_t c_s = 0, c_i = 0, c_all = 0;  // wave-uniform counts
  uint8_t *bm_out_bytes = reinterpret_cast<uint8_t *>(a.bm_out);
  for (int32_t c = blockIdx.x; c < a.n_chunks; c += gridDim.x) {
    if (c != (int32_t)blockIdx.x) {
      c0 = a.chunk_ptr[c];
      c1 = a.chunk_ptr[c + 1];
    }
    for (int64_t sub = c0; sub < c1; sub += kTile) {
      if (tid == 0) {
        n_light = n_heavy = m_light = m_heavy = next_heavy = next_mark = 0;
      }
      // phase 1: this thread's 8 consecutive agents v0 .. v0+7 (one 8-byte
      // state load, one 16-byte mark load; chunk starts are 64-aligned)
      const int64_t v0 = sub + (int64_t)tid * kPerThread;
      const int nv = v0 >= c1 ? 0 : (int)(c1 - v0 < kPerThread ? c1 - v0 : kPerThread);
      uint8_t b[kPerThread];
      uint16_t mk[kPerThread];
      if (nv == kPerThread) {
        const uint2 sw = *reinterpret_cast<const uint2 *>(a.state + v0);
        const uint4 mw = *reinterpret_cast<const uint4 *>(a.mark + v0);
        const uint32_t sws[2] = {sw.x, sw.y};
        const uint32_t mws[4] = {mw.x, mw.y, mw.z, mw.w};
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
          b[j] = (uint8_t)(sws[j >> 2] >> (8 * (j & 3)));
          mk[j] = (uint16_t)(mws[j >> 1] >> (16 * (j & 1)));
        }
      } else {
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
          b[j] = j < nv ? a.state[v0 + j] : (uint8_t)(kHeavyBit | 2);
          mk[j] = j < nv ? a.mark[v0 + j] : (uint16_t)(a.t + 0x8000u);
        }
      }
      uint32_t lmask = 0, hmask = 0, imask = 0;
      uint32_t ob[2] = {0u, 0u};
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) {
        const uint8_t s = b[j] & 3;
        ob[j >> 2] |= (uint32_t)s << (8 * (j & 3));
        if (s == 1) imask |= 1u << j;
        if (s == 0 && at_risk(mk[j], a.t)) {
          if (b[j] & kHeavyBit) hmask |= 1u << j;
          else lmask |= 1u << j;
        }
      }
      for (uint32_t m = imask; m; m &= m - 1) {
        const int j = __builtin_ctz(m);
        if ((bits_at(krec, a.n, v0 + j, kMode) >> 9) < a.thr_gamma)
          ob[j >> 2] ^= 3u << (8 * (j & 3));  // I (1) -> R (2)
      }
      *reinterpret_cast<uint2 *>(outb + tid * kPerThread) = make_uint2(ob[0], ob[1]);
      lds_barrier();  // counters reset before any push
      push_items(lmask, hmask, items, &n_light, &n_heavy);
      lds_barrier();
      // phase 2: scans (heavy rows by waves, light rows by lanes) against the
      // LDS copy of the bitmap when the tile has enough rows, else the global one
      // (block-uniform: a tile with no row to scan writes no other thread's
      // outb byte, so it needs no barrier before phase 3)
      {
        const int32_t nl = n_light, nh = n_heavy;
        if (nl + nh > 0) {
          if constexpr (kLds) {
            if (!staged && nl + nh >= kStageItems) {
              const uint4 *src = reinterpret_cast<const uint4 *>(a.bm_in);
              for (int64_t i = tid; i < (a.words >> 2); i += kSirThreads) lds_bm4[i] = src[i];
              staged = true;
              lds_barrier();
            }
          }
          if (kLds && staged)
            scan_items<kMode>(a, reinterpret_cast<const uint32_t *>(lds_bm4), sub, nl, nh,
                              items, outb, kinf, &next_heavy);
          else
            scan_items<kMode>(a, a.bm_in, sub, nl, nh, items, outb, kinf, &next_heavy);
          lds_barrier();
        }
      }
      // phase 3: outputs (one 8-byte state store when anything changed), the
      // next bitmap (one byte = these 8 agents), counts, the mark worklist
      const uint2 rw = *reinterpret_cast<const uint2 *>(outb + tid * kPerThread);
      const uint32_t rws[2] = {rw.x, rw.y};
      lmask = 0;
      hmask = 0;
      uint32_t ibits = 0, sbits = 0, changed = 0;
      uint32_t nw[2] = {0u, 0u};
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) {
        const uint8_t s = b[j] & 3;
        const uint8_t r = (uint8_t)(rws[j >> 2] >> (8 * (j & 3)));
        nw[j >> 2] |= (uint32_t)(r | (b[j] & kHeavyBit)) << (8 * (j & 3));
        if (j < nv) {
          changed |= r != s;
          ibits |= (uint32_t)(r == 1) << j;
          sbits |= (uint32_t)(r == 0) << j;
          const bool refresh = ((a.t + (uint32_t)(v0 + j)) % kMarkHorizon) == 0;
          if (r == 1 && (s == 0 || refresh)) {
            if (b[j] & kHeavyBit) hmask |= 1u << j;
            else lmask |= 1u << j;
          }
        }
      }
      if (changed) {
        if (nv == kPerThread) {
          *reinterpret_cast<uint2 *>(a.state + v0) = make_uint2(nw[0], nw[1]);
        } else {
#pragma unroll
          for (int j = 0; j < kPerThread; ++j)
            if (j < nv) a.state[v0 + j] = (uint8_t)(nw[j >> 2] >> (8 * (j & 3)));
        }
      }
      if (nv > 0) bm_out_bytes[v0 >> 3] = (uint8_t)ibits;
      c_s += (uint32_t)__popc(sbits);
      c_i += (uint32_t)__popc(ibits);
      c_all += (uint32_t)nv;
      // the scan worklist is dead after the barrier above; reuse it for marks
      push_items(lmask, hmask, items, &m_light, &m_heavy);
      lds_barrier();
      // phase 4: marks of newly infected agents and periodic refreshes
      // (block-uniform: with none, no thread reads the worklist or its
      // counters, and the next tile's reset only rewrites zeros)
      {
        const int32_t nl = m_light, nh = m_heavy;
        if (nl + nh > 0) {
          for (;;) {
            int32_t k = 0;
            if (lane == 0) k = atomicAdd(&next_mark, 1);
            k = __shfl(k, 0);
            if (k >= nh) break;
            const int64_t u = sub + items[kTile - 1 - k];
            mark_heavy(a.col, a.mark, a.row_ptr[u], a.row_ptr[u + 1], new_mark);
          }
          for (int32_t k0 = tid; k0 < nl; k0 += kSirThreads * kScanGroup) {
            int loc[kScanGroup];
            int32_t e[kScanGroup], hi[kScanGroup];
            light_group_rows(a, sub, k0, nl, items, loc, e, hi);
            mark_light_group(a.col, a.mark, e, hi, new_mark);
          }
          lds_barrier();
        }
      }
    }
  }
  // block counts: per-thread counts -> wave sums -> LDS -> one atomic per class
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    c_s += __shfl_xor(c_s, off);
    c_i += __shfl_xor(c_i, off);
    c_all += __shfl_xor(c_all, off);
  }
  if (lane == 0) {
    atomicAdd(&cnt[0], (int32_t)c_s);
    atomicAdd(&cnt[1], (int32_t)c_i);
    atomicAdd(&cnt[2], (int32_t)(c_all - c_s - c_i));
  }
  lds_barrier();
  if (tid < 3 && cnt[tid]) atomicAdd(&a.slots[(blockIdx.x % kSlots) * 3 + tid], cnt[tid]);
  if (tid == 1 && cnt[1]) a.slots[kLiveFlag] = 1;
}

// counts[t][k] = sum over the count slots of slots[t][s][k]; an identity
// step leaves its slots zero (S + I + R = n > 0 for any computed step), which
// this pass records in the flag slot's spare word for the carry pass
constexpr int kComputedMark = kSlots * 3;

__global__ void k_sir_reduce_slots(int32_t *__restrict__ slots, int64_t steps,
                                   int32_t *__restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= steps * 3) return;
  const int64_t t = i / 3, k = i - t * 3;
  const int32_t *p = slots + t * kStride + k;
  int32_t s = 0;
  for (int j = 0; j < kSlots; ++j) s += p[j * 3];
  counts[i] = s;
  if (k == 0) {
    int32_t tot = 0;
    for (int j = 0; j < kSlots * 3; ++j) tot |= slots[t * kStride + j];
    slots[t * kStride + kComputedMark] = tot != 0;
  }
}

// Identity steps form a suffix of a run (no infected agent, no new
// infection): counts[t] = counts[t_a - 1] for t >= t_a, the first identity
// step, found by binary search over the computed marks.
__global__ void k_sir_carry_counts(const int32_t *__restrict__ slots, int64_t steps,
                                   int32_t *__restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= steps * 3) return;
  const int64_t t = i / 3, k = i - t * 3;
  if (slots[t * kStride + kComputedMark]) return;
  int64_t lo = 0, hi = t;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (slots[mid * kStride + kComputedMark])
      lo = mid + 1;
    else
      hi = mid;
  }
  counts[i] = lo > 0 ? counts[(lo - 1) * 3 + k] : 0;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// state[v] = S with the heavy-row flag.
__global__ void k_sir_init_state(int64_t n, const int32_t *__restrict__ row_ptr,
                                 uint8_t *__restrict__ state) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) state[v] = (row_ptr[v + 1] - row_ptr[v] > kHeavyDeg) ? kHeavyBit : 0;
}

__global__ void k_sir_seed(const int32_t *__restrict__ perm, int64_t n0,
                           uint8_t *__restrict__ state) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) state[perm[i]] = (state[perm[i]] & kHeavyBit) | 1;
}

__global__ void k_fill_u16(uint16_t *__restrict__ p, int64_t n, uint16_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Initial marks: every infected agent marks its neighbours with `m`.
__global__ void k_sir_mark_infected(int64_t n, const int32_t *__restrict__ row_ptr,
                                    const int32_t *__restrict__ col,
                                    const uint8_t *__restrict__ state,
                                    uint16_t *__restrict__ mark, uint16_t m) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || (state[v] & 3) != 1) return;
  for (int64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) mark[col[e]] = m;
}

// bitmap word w = infected bits of agents 32w .. 32w+31
__global__ void k_sir_bitmap(int64_t n, const uint8_t *__restrict__ state,
                             int64_t words, uint32_t *__restrict__ bm) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  uint32_t x = 0;
  for (int k = 0; k < 32; ++k) {
    const int64_t v = w * 32 + k;
    if (v < n && (state[v] & 3) == 1) x |= 1u << k;
  }
  bm[w] = x;
}

__global__ void k_mask_state(const uint8_t *__restrict__ in, int64_t n,
                             uint8_t *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] & 3;
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

inline int64_t bitmap_words(int64_t n) { return ((n + 127) / 128) * 4; }

int check_graph(const abm_sir_graph *g, const char *who) {
  if (!g || g->n <= 0 || g->n >= (1ll << 31) - 64 || g->nnz < 0 ||
      g->nnz >= (1ll << 31) - 64 || g->n_chunks < 1 || g->n_chunks > g->n ||
      !g->row_ptr || (g->nnz && !g->col) || !g->chunk_ptr) {
    set_error(std::string(who) + ": bad graph descriptor");
    return kErrInvalid;
  }
  return kOk;
}

// The chunk table the step kernel assumes: chunk_ptr[0] = 0, chunk_ptr
// [n_chunks] = n, ascending, every inner boundary a multiple of kChunkAlign
// (each bitmap byte and 8-agent state word then has one owning block) and no
// chunk longer than one tile.  Read back once per instance set-up.
int check_chunks(const abm_sir_graph *g, hipStream_t s) {
  std::vector<int32_t> cp((size_t)g->n_chunks + 1);
  ABM_CHECK_HIP(hipMemcpyAsync(cp.data(), g->chunk_ptr, cp.size() * sizeof(int32_t),
                               hipMemcpyDeviceToHost, s));
  ABM_CHECK_HIP(hipStreamSynchronize(s));
  bool ok = cp.front() == 0 && cp.back() == g->n;
  for (size_t i = 1; ok && i < cp.size(); ++i) {
    ok = cp[i] > cp[i - 1] && cp[i] - cp[i - 1] <= kTile;
    if (i + 1 < cp.size()) ok = ok && cp[i] % kChunkAlign == 0;
  }
  if (!ok) {
    set_error("abm_sir_seed: chunk_ptr must run from 0 to n, ascending, with inner boundaries "
              "multiples of 64 and chunks of at most 8192 agents (abm_node_chunks with "
              "abm_sir_chunk_params)");
    return kErrInvalid;
  }
  return kOk;
}

// The bitmap-staging kernels need more than 64 KB of dynamic LDS: the limit
// is raised once per device (under a lock; each device keeps its own
// function attributes).
constexpr int kMaxDevices = 64;
std::mutex g_lds_attr_mu;
bool g_lds_attr_set[kMaxDevices];

int sir_lds_attr() {
  int dev = -1;
  ABM_CHECK_HIP(hipGetDevice(&dev));
  ABM_REQUIRE(dev >= 0 && dev < kMaxDevices, "sir: device index out of range");
  std::lock_guard<std::mutex> lock(g_lds_attr_mu);
  if (!g_lds_attr_set[dev]) {
    for (const void *k : {(const void *)k_sir_step<true, kModePartitionable>,
                          (const void *)k_sir_step<true, kModeOriginal>})
      ABM_CHECK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(kLdsBitmapWords * 4)));
    g_lds_attr_set[dev] = true;
  }
  return kOk;
}

int check_aligned(const void *p, uintptr_t bytes, const char *who, const char *what) {
  if (((uintptr_t)p & (bytes - 1)) != 0) {
    set_error(std::string(who) + ": " + what + " must be " + std::to_string(bytes) +
              "-byte aligned");
    return kErrInvalid;
  }
  return kOk;
}

int check_buffers(const uint8_t *state, const uint16_t *mark, const uint32_t *b0,
                  const uint32_t *b1, const char *who) {
  if (int rc = check_aligned(state, 8, who, "state")) return rc;
  if (int rc = check_aligned(mark, 16, who, "mark")) return rc;
  if (int rc = check_aligned(b0, 16, who, "bitmap")) return rc;
  return check_aligned(b1, 16, who, "bitmap");
}

int launch_step(const abm_sir_graph *g, uint8_t *state, uint16_t *mark,
                const uint32_t *bm_in, uint32_t *bm_out, uint32_t t,
                const uint32_t *keys4, uint32_t tb, uint32_t tg, int mode,
                const int32_t *prev_slots, const int32_t *prev2_slots, int32_t *slots,
                hipStream_t s) {
  SirStep a;
  a.n = g->n;
  a.nnz = g->nnz;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.chunk_ptr = g->chunk_ptr;
  a.n_chunks = (int32_t)g->n_chunks;
  a.state = state;
  a.mark = mark;
  a.bm_in = bm_in;
  a.bm_out = bm_out;
  a.words = bitmap_words(g->n);
  a.t = t;
  a.prev_slots = prev_slots;
  a.prev2_slots = prev2_slots;
  a.keys = keys4;
  a.thr_beta = tb;
  a.thr_gamma = tg;
  a.mode = mode;
  a.slots = slots;
  const unsigned blocks = (unsigned)(g->n_chunks < kSirBlocks ? g->n_chunks : kSirBlocks);
  // one instantiation per stream layout: each carries one copy of the draw
  const bool part = a.mode == kModePartitionable;
  if (a.words <= kLdsBitmapWords) {
    auto *kp = k_sir_step<true, kModePartitionable>;
    auto *ko = k_sir_step<true, kModeOriginal>;
    if (int rc = sir_lds_attr()) return rc;
    hipLaunchKernelGGL(part ? kp : ko, dim3(blocks), dim3(kSirThreads), (size_t)a.words * 4,
                       s, a);
  } else {
    auto *kg = part ? k_sir_step<false, kModePartitionable> : k_sir_step<false, kModeOriginal>;
    hipLaunchKernelGGL(kg, dim3(blocks), dim3(kSirThreads), 0, s, a);
  }
  return kOk;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream) {
  ABM_REQUIRE(n >= 0, "abm_fill_i32: n must be >= 0");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, p, n, v);
  ABM_LAUNCH_CHECK("abm_fill_i32");
  return kOk;
}

uint32_t abm_unit_threshold(float p) { return unit_threshold(p); }

int abm_sir_slots(void) { return kStepSlots; }

int64_t abm_sir_bitmap_words(int64_t n) { return n > 0 ? bitmap_words(n) : 0; }

int abm_sir_chunk_params(int32_t *align, int64_t *node_cost, int32_t *heavy_deg,
                         int64_t *heavy_cost, int32_t *chunks_per_block, int32_t *blocks,
                         int32_t *max_nodes) {
  if (align) *align = kChunkAlign;
  if (heavy_deg) *heavy_deg = kHeavyDeg;
  if (heavy_cost) *heavy_cost = kChunkHeavyCost;
  if (max_nodes) *max_nodes = kTile;
  if (node_cost) *node_cost = kChunkNodeCost;
  if (chunks_per_block) *chunks_per_block = kChunksPerBlock;
  if (blocks) *blocks = kSirBlocks;
  return kOk;
}

int abm_sir_seed(const abm_sir_graph *g, const int32_t *perm, int64_t n0,
                 uint8_t *state, uint16_t *mark, uint32_t *bitmap, uint32_t t0,
                 void *stream) {
  if (int rc = check_graph(g, "abm_sir_seed")) return rc;
  ABM_REQUIRE(state && mark && bitmap && n0 >= 0 && n0 <= g->n && (n0 == 0 || perm),
              "abm_sir_seed: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (int rc = check_chunks(g, s)) return rc;
  // the per-device launch set-up, here rather than inside a captured run
  if (int rc = sir_lds_attr()) return rc;
  const int64_t n = g->n, words = bitmap_words(n);
  hipLaunchKernelGGL(k_sir_init_state, dim3(blocks_for(n, 256)), dim3(256), 0,
                     s, n, g->row_ptr, state);
  if (n0 > 0)
    hipLaunchKernelGGL(k_sir_seed, dim3(blocks_for(n0, 256)), dim3(256), 0, s,
                       perm, n0, state);
  hipLaunchKernelGGL(k_fill_u16, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     mark, n, (uint16_t)(t0 + 0x8000u));
  hipLaunchKernelGGL(k_sir_mark_infected, dim3(blocks_for(n, 256)), dim3(256), 0,
                     s, n, g->row_ptr, g->col, state, mark,
                     (uint16_t)(t0 + kMarkHorizon));
  hipLaunchKernelGGL(k_sir_bitmap, dim3(blocks_for(words, 256)), dim3(256), 0,
                     s, n, state, words, bitmap);
  ABM_LAUNCH_CHECK("abm_sir_seed");
  return kOk;
}

int abm_sir_run(const abm_sir_graph *g, uint8_t *state, uint16_t *mark,
                uint32_t *bitmaps, uint32_t t0, const uint32_t *keys,
                int32_t steps, float beta, float gamma, int mode, int32_t *slots,
                int32_t *counts, void *stream) {
  if (int rc = check_graph(g, "abm_sir_run")) return rc;
  ABM_REQUIRE(state && mark && bitmaps && keys && slots && counts,
              "abm_sir_run: null buffer");
  ABM_REQUIRE(steps >= 0 && (mode == 0 || mode == 1),
              "abm_sir_run: bad steps or mode");
  if (int rc = check_buffers(state, mark, bitmaps, bitmaps + bitmap_words(g->n),
                             "abm_sir_run"))
    return rc;
  if (steps == 0) return kOk;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t tb = unit_threshold(beta), tg = unit_threshold(gamma);
  const int64_t words = bitmap_words(g->n);
  const int64_t ns = (int64_t)steps * kStride;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(ns, 256)), dim3(256), 0, s,
                     slots, ns, 0);
  uint32_t *bm[2] = {bitmaps, bitmaps + words};
  for (int32_t t = 0; t < steps; ++t)
    if (int rc = launch_step(g, state, mark, bm[t & 1], bm[(t + 1) & 1],
                             t0 + (uint32_t)t, keys + 4 * (int64_t)t, tb, tg, mode,
                             t ? slots + (int64_t)(t - 1) * kStride : nullptr,
                             t > 1 ? slots + (int64_t)(t - 2) * kStride : nullptr,
                             slots + (int64_t)t * kStride, s))
      return rc;
  hipLaunchKernelGGL(k_sir_reduce_slots, dim3(blocks_for(3 * (int64_t)steps, 256)),
                     dim3(256), 0, s, slots, (int64_t)steps, counts);
  hipLaunchKernelGGL(k_sir_carry_counts, dim3(blocks_for(3 * (int64_t)steps, 256)),
                     dim3(256), 0, s, (const int32_t *)slots, (int64_t)steps, counts);
  if (steps & 1)
    ABM_CHECK_HIP(hipMemcpyAsync(bm[0], bm[1], (size_t)words * 4,
                                 hipMemcpyDeviceToDevice, s));
  ABM_LAUNCH_CHECK("abm_sir_run");
  return kOk;
}

int abm_sir_step(const abm_sir_graph *g, uint8_t *state, uint16_t *mark,
                 const uint32_t *bm_in, uint32_t *bm_out, uint32_t t,
                 const uint32_t *keys4, float beta, float gamma, int mode,
                 const int32_t *prev_slots, int32_t *slots, void *stream) {
  if (int rc = check_graph(g, "abm_sir_step")) return rc;
  ABM_REQUIRE(state && mark && bm_in && bm_out && keys4 && slots && bm_in != bm_out &&
                  (mode == 0 || mode == 1),
              "abm_sir_step: bad argument");
  if (int rc = check_buffers(state, mark, bm_in, bm_out, "abm_sir_step")) return rc;
  if (int rc = launch_step(g, state, mark, bm_in, bm_out, t, keys4, unit_threshold(beta),
                           unit_threshold(gamma), mode, prev_slots, nullptr, slots,
                           (hipStream_t)stream))
    return rc;
  ABM_LAUNCH_CHECK("abm_sir_step");
  return kOk;
}

int abm_sir_reduce_counts(int32_t *slots, int32_t steps, int32_t *counts, void *stream) {
  ABM_REQUIRE(slots && counts && steps >= 0, "abm_sir_reduce_counts: bad args");
  if (steps == 0) return kOk;
  hipLaunchKernelGGL(k_sir_reduce_slots, dim3(blocks_for(3 * (int64_t)steps, 256)),
                     dim3(256), 0, (hipStream_t)stream, slots, (int64_t)steps,
                     counts);
  hipLaunchKernelGGL(k_sir_carry_counts, dim3(blocks_for(3 * (int64_t)steps, 256)),
                     dim3(256), 0, (hipStream_t)stream, (const int32_t *)slots,
                     (int64_t)steps, counts);
  ABM_LAUNCH_CHECK("abm_sir_reduce_counts");
  return kOk;
}

int abm_sir_states(const uint8_t *state, int64_t n, uint8_t *out, void *stream) {
  ABM_REQUIRE(state && out && n >= 0, "abm_sir_states: bad args");
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_mask_state, dim3(blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, state, n, out);
  ABM_LAUNCH_CHECK("abm_sir_states");
  return kOk;
}

}  // extern "C"
