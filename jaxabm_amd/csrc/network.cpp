// network.cpp -- host-side Network construction: Barabasi-Albert edge lists
// and undirected CSR.
//
// Replaces the reference's edge-at-a-time Network (jaxabm/agentpy.py:559-582:
// one jnp.concatenate per add_edge, O(E^2) to build) with a bulk builder whose
// observable neighbour order matches Network.get_neighbors
// (jaxabm/agentpy.py:584-615: ascending unique ids for undirected networks).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "jaxabm_amd.h"

using namespace abm;

extern "C" {

int64_t abm_ba_num_edges(int64_t n, int32_t m) {
  if (n <= m) return n > 0 ? n - 1 : 0;
  return (int64_t)m * (n - m);
}

// Barabasi-Albert (n, m): star on nodes 0..m centred on 0, then each node
// s = m+1 .. n-1 attaches to m distinct nodes drawn from the degree-repeated
// node list; draw a of node s picks index floor(r * len / 2^32) with
// r = threefry(key, (s, a)).y0.
int64_t abm_ba_edges(int64_t n, int32_t m, const uint32_t key[2], int32_t *src,
                     int32_t *dst) {
  if (n <= 1 || m < 1 || !key || !src || !dst) {
    if (n > 1 && m < 1) set_error("abm_ba_edges: m must be >= 1");
    return n <= 1 ? 0 : -1;
  }
  int64_t e = 0;
  if (n <= m) {
    for (int64_t i = 1; i < n; ++i, ++e) {
      src[e] = 0;
      dst[e] = (int32_t)i;
    }
    return e;
  }
  std::vector<int32_t> rep;
  rep.reserve((size_t)(2 * abm_ba_num_edges(n, m) + 2));
  for (int32_t i = 1; i <= m; ++i, ++e) {
    src[e] = 0;
    dst[e] = i;
  }
  rep.insert(rep.end(), (size_t)m, 0);
  for (int32_t i = 1; i <= m; ++i) rep.push_back(i);
  std::vector<int32_t> pick((size_t)m);
  // draws are a pure function of (s, attempt): compute them four at a time
  // (independent chains), consume them in attempt order
  constexpr int kBatch = 4;
  uint32_t draw[kBatch];
  for (int64_t s = m + 1; s < n; ++s) {
    int32_t have = 0;
    const uint64_t len = rep.size();
    for (uint32_t a = 0; have < m; a += kBatch) {
      for (int j = 0; j < kBatch; ++j) {
        uint32_t y0, y1;
        threefry(key[0], key[1], (uint32_t)s, a + (uint32_t)j, y0, y1);
        draw[j] = y0;
      }
      for (int j = 0; j < kBatch && have < m; ++j) {
        const int32_t t = rep[(size_t)(((uint64_t)draw[j] * len) >> 32)];
        bool seen = false;
        for (int32_t q = 0; q < have; ++q) seen |= (pick[q] == t);
        if (!seen) pick[have++] = t;
      }
    }
    for (int32_t q = 0; q < m; ++q, ++e) {
      src[e] = (int32_t)s;
      dst[e] = pick[q];
      rep.push_back(pick[q]);
    }
    rep.insert(rep.end(), (size_t)m, (int32_t)s);
  }
  return e;
}

// Undirected CSR with ascending neighbour ids per row (duplicates kept):
// degree count, scatter of both directions of every edge, then a sort of each
// row (rows are short; insertion sort below 32 entries).
int abm_csr_undirected(int64_t n, int64_t n_edges, const int32_t *src,
                       const int32_t *dst, int32_t *row_ptr, int32_t *col) {
  ABM_REQUIRE(n >= 0 && n_edges >= 0, "abm_csr_undirected: bad sizes");
  ABM_REQUIRE(2 * n_edges < (1ll << 31), "abm_csr_undirected: too many edges");
  ABM_REQUIRE(row_ptr && (n_edges == 0 || (src && dst && col)),
              "abm_csr_undirected: null pointer");
  for (int64_t e = 0; e < n_edges; ++e) {
    if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n) {
      set_error("abm_csr_undirected: edge endpoint out of range");
      return kErrInvalid;
    }
  }
  std::fill(row_ptr, row_ptr + n + 1, 0);
  for (int64_t e = 0; e < n_edges; ++e) {
    row_ptr[src[e] + 1]++;
    row_ptr[dst[e] + 1]++;
  }
  for (int64_t v = 0; v < n; ++v) row_ptr[v + 1] += row_ptr[v];
  std::vector<int32_t> pos(row_ptr, row_ptr + n);
  for (int64_t e = 0; e < n_edges; ++e) {
    col[pos[(size_t)src[e]]++] = dst[e];
    col[pos[(size_t)dst[e]]++] = src[e];
  }
  for (int64_t v = 0; v < n; ++v) {
    int32_t *a = col + row_ptr[v];
    const int32_t len = row_ptr[v + 1] - row_ptr[v];
    if (len > 32) {
      std::sort(a, a + len);
      continue;
    }
    for (int32_t i = 1; i < len; ++i) {
      const int32_t x = a[i];
      int32_t j = i - 1;
      while (j >= 0 && a[j] > x) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = x;
    }
  }
  return kOk;
}

// Contiguous node chunks of roughly equal cost (cost(v) = node_cost +
// degree(v), plus heavy_cost when degree(v) > heavy_deg) and at most
// max_nodes nodes, whose boundaries are multiples of `align` (except the last;
// max_nodes must be a multiple of align): chunk c covers nodes chunk_ptr[c] ..
// chunk_ptr[c+1].  Returns the number of chunks; chunk_ptr may be NULL to
// count.
int64_t abm_node_chunks(int64_t n, const int32_t *row_ptr, int32_t align,
                        int64_t node_cost, int32_t heavy_deg, int64_t heavy_cost,
                        int64_t target, int32_t max_nodes, int32_t *chunk_ptr) {
  if (!row_ptr || n < 0 || align < 1 || target < 1 || max_nodes < align ||
      max_nodes % align)
    return -1;
  int64_t c = 0, acc = 0, start = 0;
  if (chunk_ptr) chunk_ptr[0] = 0;
  for (int64_t v = 0; v < n; ++v) {
    const int64_t d = row_ptr[v + 1] - row_ptr[v];
    acc += node_cost + d + (d > heavy_deg ? heavy_cost : 0);
    const bool full = v + 1 - start >= max_nodes;
    if ((acc >= target || full) && (v + 1) % align == 0 && v + 1 < n) {
      ++c;
      if (chunk_ptr) chunk_ptr[c] = (int32_t)(v + 1);
      acc = 0;
      start = v + 1;
    }
  }
  if (n > 0) {
    ++c;
    if (chunk_ptr) chunk_ptr[c] = (int32_t)n;
  }
  return c;
}

}  // extern "C"
