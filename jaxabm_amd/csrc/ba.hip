// ba.hip -- Barabasi-Albert edge lists built on the device.
//
// The same graph as the host builder (abm_ba_edges): a star on nodes 0..m
// centred on 0, then each node s = m+1 .. n-1 attaches to m distinct nodes
// drawn from the degree-repeated node list, draw a of node s taking slot
// floor(r * len / 2^32) with r = threefry(key, (s, a)).y0.  The reference
// grows its edge list one add_edge at a time (jaxabm/agentpy.py:559-582).
//
// Preferential attachment looks sequential, but the repeated-node list has a
// closed form: slots 0..m-1 hold 0, slots m..2m-1 hold 1..m, and node s >= m+1
// owns slots [2m(s-m), 2m(s-m+1)): its m picks in order, then m copies of s.
// Its length when node s draws is 2m(s-m), and every draw is a pure function
// of (s, a).  So a draw resolves to a known node (a copy slot) or to pick q of
// an earlier node s' -- the q-th distinct node s' drew.  Rounds of one thread
// per node resolve them: a node walks its draws from a = 0, writes each new
// distinct pick into its edge slot as soon as it is found, and stops at the
// first draw whose slot holds a pick not yet written (-1).  A written pick
// never changes, so any non-negative value a thread reads is final; the
// lowest unresolved node always completes, so every round makes progress
// (in practice a few dozen rounds at n = 1M).
//
// Layout: src/dst int32[E], E = m (n - m): entries 0..m-1 the star (0, i),
// then node s's picks at m + (s-m-1) m + q, q < m -- the host builder's order.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace abm {

constexpr int kBaThreads = 256;
constexpr int kBaRoundsPerCheck = 8;

__device__ __forceinline__ int64_t ba_pick_slot(int64_t s, int32_t m, int32_t q) {
  return (int64_t)m + (s - m - 1) * (int64_t)m + q;
}

// Node in slot p of the repeated-node list, or -1 while that slot is a pick
// not yet written.
__device__ __forceinline__ int32_t ba_rep(int64_t p, int32_t m, const int32_t *dst) {
  if (p < m) return 0;
  if (p < 2 * (int64_t)m) return (int32_t)(p - m + 1);
  const int64_t q = p - 2 * (int64_t)m;
  const int64_t s = m + 1 + q / (2 * (int64_t)m);
  const int64_t off = q - (s - m - 1) * 2 * (int64_t)m;
  if (off >= m) return (int32_t)s;
  return __hip_atomic_load(dst + ba_pick_slot(s, m, (int32_t)off), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ void __launch_bounds__(kBaThreads)
    k_ba_init(int64_t n, int32_t m, int64_t E, int32_t *__restrict__ src,
              int32_t *__restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * kBaThreads + threadIdx.x;
  if (e >= E) return;
  if (e < m || n <= m) {
    src[e] = 0;
    dst[e] = (int32_t)(e + 1);
  } else {
    src[e] = (int32_t)(m + 1 + (e - m) / m);
    dst[e] = -1;
  }
}

__global__ void __launch_bounds__(kBaThreads)
    k_ba_round(int64_t n, int32_t m, uint32_t k0, uint32_t k1, int32_t *dst,
               uint32_t *unresolved) {
  const int64_t s = m + 1 + (int64_t)blockIdx.x * kBaThreads + threadIdx.x;
  bool open = false;
  // this node is done once its last pick is written (by itself, in an
  // earlier round); other nodes' pick slots are read with relaxed atomics: a
  // value read while its node is still drawing is -1 (retried next round) or
  // final
  if (s < n && dst[ba_pick_slot(s, m, m - 1)] < 0) {
    const uint64_t len = 2ull * (uint64_t)m * (uint64_t)(s - m);
    int32_t have = 0;
    for (uint32_t a = 0; have < m; ++a) {
      uint32_t y0, y1;
      threefry(k0, k1, (uint32_t)s, a, y0, y1);
      const int32_t t = ba_rep((int64_t)(((uint64_t)y0 * len) >> 32), m, dst);
      if (t < 0) {
        open = true;
        break;
      }
      bool seen = false;
      for (int32_t q = 0; q < have; ++q) seen |= dst[ba_pick_slot(s, m, q)] == t;
      if (!seen) {
        __hip_atomic_store(dst + ba_pick_slot(s, m, have), t, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
        ++have;
      }
    }
  }
  const uint64_t b = __ballot(open);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(unresolved, (uint32_t)__popcll(b));
}

__global__ void k_ba_zero(uint32_t *p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_ba_edges_device(int64_t n, int32_t m, const uint32_t key[2], int32_t *src,
                        int32_t *dst, uint32_t *unresolved, void *stream) {
  ABM_REQUIRE(n >= 0 && n < (1ll << 31) && key, "abm_ba_edges_device: bad arguments");
  if (n <= 1) return kOk;
  ABM_REQUIRE(m >= 1, "abm_ba_edges_device: m must be >= 1");
  const int64_t E = abm_ba_num_edges(n, m);
  ABM_REQUIRE(E < (1ll << 31), "abm_ba_edges_device: too many edges");
  ABM_REQUIRE(src && dst && unresolved, "abm_ba_edges_device: null buffer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ba_init, dim3((unsigned)((E + kBaThreads - 1) / kBaThreads)),
                     dim3(kBaThreads), 0, s, n, m, E, src, dst);
  ABM_LAUNCH_CHECK("abm_ba_edges_device");
  if (n <= m + 1) return kOk;
  const int64_t nodes = n - m - 1;
  const unsigned blocks = (unsigned)((nodes + kBaThreads - 1) / kBaThreads);
  // every round completes at least the lowest open node: at most `nodes`
  // rounds, checked every kBaRoundsPerCheck rounds
  for (int64_t done = 0; done < nodes + kBaRoundsPerCheck; done += kBaRoundsPerCheck) {
    for (int r = 0; r < kBaRoundsPerCheck; ++r) {
      if (r == kBaRoundsPerCheck - 1)
        hipLaunchKernelGGL(k_ba_zero, dim3(1), dim3(64), 0, s, unresolved);
      hipLaunchKernelGGL(k_ba_round, dim3(blocks), dim3(kBaThreads), 0, s, n, m, key[0], key[1],
                         dst, unresolved);
    }
    ABM_LAUNCH_CHECK("abm_ba_edges_device");
    uint32_t open = 0;
    ABM_CHECK_HIP(hipMemcpyAsync(&open, unresolved, sizeof(open), hipMemcpyDeviceToHost, s));
    ABM_CHECK_HIP(hipStreamSynchronize(s));
    if (open == 0) return kOk;
  }
  set_error("abm_ba_edges_device: picks left unresolved");
  return kErrState;
}

}  // extern "C"
