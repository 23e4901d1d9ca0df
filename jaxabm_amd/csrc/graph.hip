// graph.hip -- capture of a device-resident step loop as a hipGraph, and
// stream-ordered event timing.
//
// Replaces the host loop of JaxModel.run (jaxabm/model.py:240-241: one
// Python iteration and a dozen un-jitted dispatches per step) by one graph
// launch for the whole run.  Captured code records kernels and device copies
// only: buffers that must start at zero are cleared by fill kernels, never by
// memset nodes.  abm_graph_sanitize enforces the same on a graph captured
// from other code (a generic model step of torch ops): memset nodes become
// fill-kernel nodes, and graphs with host, event or allocation nodes are
// refused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "jaxabm_amd.h"

struct abm_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace abm {

// Stands in for a captured memset node: `height` rows of `width` elements of
// `elem` bytes (1, 2 or 4) at dst + row * pitch, each set to `value`.
__global__ void __launch_bounds__(256)
    k_graph_fill(char *dst, size_t pitch, size_t width, size_t height, unsigned elem,
                 unsigned value) {
  const size_t total = width * height;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const size_t r = i / width, c = i - r * width;
    char *p = dst + r * pitch + c * elem;
    if (elem == 4)
      *(uint32_t *)p = value;
    else if (elem == 2)
      *(uint16_t *)p = (uint16_t)value;
    else
      *(uint8_t *)p = (uint8_t)value;
  }
}

const char *node_type_name(hipGraphNodeType t) {
  switch (t) {
    case hipGraphNodeTypeHost: return "host";
    case hipGraphNodeTypeGraph: return "child graph";
    case hipGraphNodeTypeWaitEvent: return "event wait";
    case hipGraphNodeTypeEventRecord: return "event record";
    case hipGraphNodeTypeMemAlloc: return "memory alloc";
    case hipGraphNodeTypeMemFree: return "memory free";
    case hipGraphNodeTypeMemcpyFromSymbol: return "memcpy from symbol";
    case hipGraphNodeTypeMemcpyToSymbol: return "memcpy to symbol";
    default: return "unsupported";
  }
}

// Replaces one memset node by a k_graph_fill kernel node with the same
// dependencies and dependents.
int replace_memset(hipGraph_t g, hipGraphNode_t node) {
  hipMemsetParams mp{};
  ABM_CHECK_HIP(hipGraphMemsetNodeGetParams(node, &mp));
  ABM_REQUIRE(mp.dst && (mp.elementSize == 1 || mp.elementSize == 2 || mp.elementSize == 4),
              "abm_graph_sanitize: memset node with unsupported parameters");
  const size_t height = mp.height ? mp.height : 1;
  ABM_REQUIRE(height == 1 || mp.pitch >= mp.width * mp.elementSize,
              "abm_graph_sanitize: memset node with pitch below its row width");
  size_t ndep = 0, nout = 0;
  ABM_CHECK_HIP(hipGraphNodeGetDependencies(node, nullptr, &ndep));
  ABM_CHECK_HIP(hipGraphNodeGetDependentNodes(node, nullptr, &nout));
  std::vector<hipGraphNode_t> deps(ndep), outs(nout);
  if (ndep) ABM_CHECK_HIP(hipGraphNodeGetDependencies(node, deps.data(), &ndep));
  if (nout) ABM_CHECK_HIP(hipGraphNodeGetDependentNodes(node, outs.data(), &nout));
  char *dst = (char *)mp.dst;
  size_t pitch = mp.pitch, width = mp.width, h = height;
  unsigned elem = mp.elementSize, value = mp.value;
  void *args[] = {&dst, &pitch, &width, &h, &elem, &value};
  const size_t total = width * h;
  size_t blocks = (total + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  hipKernelNodeParams kp{};
  kp.blockDim = dim3(256);
  kp.gridDim = dim3((unsigned)blocks);
  kp.func = (void *)k_graph_fill;
  kp.kernelParams = args;
  kp.sharedMemBytes = 0;
  kp.extra = nullptr;
  hipGraphNode_t fill = nullptr;
  ABM_CHECK_HIP(hipGraphAddKernelNode(&fill, g, ndep ? deps.data() : nullptr, ndep, &kp));
  for (hipGraphNode_t o : outs) ABM_CHECK_HIP(hipGraphAddDependencies(g, &fill, &o, 1));
  ABM_CHECK_HIP(hipGraphDestroyNode(node));
  return kOk;
}

}  // namespace abm

namespace abm {

// A step's result copies in one launch: item y copies items.it[y].bytes
// from src to dst + row * row_stride (row read from device memory when
// row_stride != 0; rows at or past row_limit are not written).  The list
// travels by value in the kernel arguments, so a captured launch keeps it.
struct CopyItems {
  abm_copy_item it[ABM_MAX_COPY_ITEMS];
};

__global__ void __launch_bounds__(256)
    k_copy_items(const CopyItems items, const int64_t *__restrict__ row) {
  const abm_copy_item &c = items.it[blockIdx.y];
  int64_t r = 0;
  if (c.row_stride != 0) {
    r = row[0];
    if (r < 0 || r >= c.row_limit) return;
  }
  char *dst = (char *)c.dst + r * c.row_stride;
  const char *src = (const char *)c.src;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((((uintptr_t)dst | (uintptr_t)src | (uintptr_t)c.bytes) & 15u) == 0) {
    const int64_t n16 = c.bytes >> 4;
    for (int64_t i = t0; i < n16; i += step)
      ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
  } else {
    for (int64_t i = t0; i < c.bytes; i += step) dst[i] = src[i];
  }
}

// k_copy_items plus the step's tail, done by the last block to finish (the
// one whose arrival makes `done` reach the grid size; every block reads row
// before it arrives): row and ts advance, `done` returns to 0, and row r + 1
// of the key table is copied to key_dst for the next step (if r + 1 <
// key_rows).
__global__ void __launch_bounds__(256)
    k_commit_step(const CopyItems items, int32_t n_items, int64_t *__restrict__ row,
                  int32_t *__restrict__ ts, uint32_t *__restrict__ done,
                  const uint32_t *__restrict__ key_table, int64_t key_rows, int32_t key_words,
                  uint32_t *__restrict__ key_dst) {
  if ((int32_t)blockIdx.y < n_items) {
    const abm_copy_item &c = items.it[blockIdx.y];
    int64_t r = 0;
    bool write = true;
    if (c.row_stride != 0) {
      r = row[0];
      write = r >= 0 && r < c.row_limit;
    }
    if (write) {
      char *dst = (char *)c.dst + r * c.row_stride;
      const char *src = (const char *)c.src;
      const int64_t step = (int64_t)gridDim.x * blockDim.x;
      const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
      if ((((uintptr_t)dst | (uintptr_t)src | (uintptr_t)c.bytes) & 15u) == 0) {
        const int64_t n16 = c.bytes >> 4;
        for (int64_t i = t0; i < n16; i += step) ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
      } else {
        for (int64_t i = t0; i < c.bytes; i += step) dst[i] = src[i];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned total = gridDim.x * gridDim.y;
    if (atomicAdd(done, 1u) == total - 1) {
      __threadfence();
      const int64_t r = row[0] + 1;
      row[0] = r;
      ts[0] += 1;
      *done = 0u;
      if (key_dst && r < key_rows)
        for (int32_t w = 0; w < key_words; ++w) key_dst[w] = key_table[r * key_words + w];
      __threadfence();
    }
  }
}

__global__ void k_step_counters(int64_t *__restrict__ row, int32_t *__restrict__ ts) {
  if (threadIdx.x == 0) {
    row[0] += 1;
    ts[0] += 1;
  }
}

}  // namespace abm

using namespace abm;

extern "C" {

int abm_copy_items(const abm_copy_item *items, int32_t n_items, const int64_t *row,
                   void *stream) {
  ABM_REQUIRE(items && n_items >= 0 && n_items <= ABM_MAX_COPY_ITEMS,
              "abm_copy_items: bad item list");
  if (n_items == 0) return kOk;
  CopyItems list{};
  int64_t max_bytes = 0;
  for (int32_t i = 0; i < n_items; ++i) {
    const abm_copy_item &c = items[i];
    ABM_REQUIRE(c.src && c.dst && c.bytes >= 0 && (c.row_stride == 0 || (row && c.row_limit > 0)),
                "abm_copy_items: bad item");
    list.it[i] = c;
    if (c.bytes > max_bytes) max_bytes = c.bytes;
  }
  const int64_t per_block = 256 * 16;
  const unsigned bx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (max_bytes + per_block - 1) / per_block), 1024);
  hipLaunchKernelGGL(k_copy_items, dim3(bx, (unsigned)n_items), dim3(256), 0, (hipStream_t)stream,
                     list, row);
  ABM_LAUNCH_CHECK("abm_copy_items");
  return kOk;
}

int abm_commit_step(const abm_copy_item *items, int32_t n_items, int64_t *row, int32_t *ts,
                    uint32_t *done, const uint32_t *key_table, int64_t key_rows,
                    int32_t key_words, uint32_t *key_dst, void *stream) {
  ABM_REQUIRE((items || n_items == 0) && n_items >= 0 && n_items <= ABM_MAX_COPY_ITEMS && row &&
                  ts && done && key_rows >= 0 && key_words >= 0 &&
                  (!key_dst || (key_table && key_words > 0)),
              "abm_commit_step: bad arguments");
  CopyItems list{};
  int64_t max_bytes = 0;
  for (int32_t i = 0; i < n_items; ++i) {
    const abm_copy_item &c = items[i];
    ABM_REQUIRE(c.src && c.dst && c.bytes >= 0 && (c.row_stride == 0 || c.row_limit > 0),
                "abm_commit_step: bad item");
    list.it[i] = c;
    if (c.bytes > max_bytes) max_bytes = c.bytes;
  }
  const int64_t per_block = 256 * 16;
  const unsigned bx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (max_bytes + per_block - 1) / per_block), 1024);
  hipLaunchKernelGGL(k_commit_step, dim3(bx, (unsigned)std::max<int32_t>(n_items, 1)), dim3(256), 0,
                     (hipStream_t)stream, list, n_items, row, ts, done, key_table, key_rows,
                     key_words, key_dst);
  ABM_LAUNCH_CHECK("abm_commit_step");
  return kOk;
}

int abm_step_counters(int64_t *row, int32_t *ts, void *stream) {
  ABM_REQUIRE(row && ts, "abm_step_counters: null pointer");
  hipLaunchKernelGGL(k_step_counters, dim3(1), dim3(64), 0, (hipStream_t)stream, row, ts);
  ABM_LAUNCH_CHECK("abm_step_counters");
  return kOk;
}

int abm_graph_begin(void *stream) {
  ABM_CHECK_HIP(hipStreamBeginCapture((hipStream_t)stream,
                                      hipStreamCaptureModeThreadLocal));
  return kOk;
}

int abm_graph_end(void *stream, abm_graph **out) {
  ABM_REQUIRE(out, "abm_graph_end: null out");
  hipGraph_t g = nullptr;
  ABM_CHECK_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);  // the instantiation error is the one reported
    return hip_fail(e, "hipGraphInstantiate");
  }
  abm_graph *h = new abm_graph{g, ex};
  *out = h;
  return kOk;
}

int abm_graph_launch(abm_graph *g, void *stream) {
  ABM_REQUIRE(g, "abm_graph_launch: null graph");
  ABM_CHECK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return kOk;
}

int abm_graph_destroy(abm_graph *g) {
  if (!g) return kOk;
  hipError_t e = hipSuccess;
  if (g->exec) e = hipGraphExecDestroy(g->exec);
  if (g->graph) {
    const hipError_t e2 = hipGraphDestroy(g->graph);
    if (e == hipSuccess) e = e2;
  }
  delete g;
  if (e != hipSuccess) return hip_fail(e, "abm_graph_destroy");
  return kOk;
}

int abm_graph_sanitize(void *graph, int32_t *counts) {
  ABM_REQUIRE(graph && counts, "abm_graph_sanitize: null argument");
  hipGraph_t g = (hipGraph_t)graph;
  size_t n = 0;
  ABM_CHECK_HIP(hipGraphGetNodes(g, nullptr, &n));
  std::vector<hipGraphNode_t> nodes(n);
  if (n) ABM_CHECK_HIP(hipGraphGetNodes(g, nodes.data(), &n));
  int32_t c[4] = {0, 0, 0, 0};
  std::vector<hipGraphNode_t> memsets;
  for (hipGraphNode_t node : nodes) {
    hipGraphNodeType t;
    ABM_CHECK_HIP(hipGraphNodeGetType(node, &t));
    if (t == hipGraphNodeTypeKernel) {
      ++c[1];
    } else if (t == hipGraphNodeTypeMemcpy) {
      ++c[2];
    } else if (t == hipGraphNodeTypeMemset) {
      memsets.push_back(node);
    } else if (t == hipGraphNodeTypeEmpty) {
      ++c[3];
    } else {
      set_error(std::string("abm_graph_sanitize: the graph holds a ") + node_type_name(t) +
                " node; only kernels, device copies and memsets are replayed");
      return kErrInvalid;
    }
  }
  for (hipGraphNode_t node : memsets) {
    if (int rc = replace_memset(g, node)) return rc;
    ++c[0];
  }
  for (int i = 0; i < 4; ++i) counts[i] = c[i];
  return kOk;
}

int abm_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
  ABM_REQUIRE(dst && src, "abm_memcpy_d2d: null pointer");
  if (bytes == 0) return kOk;
  ABM_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
  return kOk;
}

// phase 0: record start; phase 1: record stop, wait, return elapsed ms.
int abm_event_timer(void *stream, int phase, float *ms_out) {
  static thread_local hipEvent_t ev[2] = {nullptr, nullptr};
  if (!ev[0]) {
    ABM_CHECK_HIP(hipEventCreate(&ev[0]));
    ABM_CHECK_HIP(hipEventCreate(&ev[1]));
  }
  ABM_REQUIRE(phase == 0 || phase == 1, "abm_event_timer: phase must be 0/1");
  ABM_CHECK_HIP(hipEventRecord(ev[phase], (hipStream_t)stream));
  if (phase == 1) {
    ABM_CHECK_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    ABM_CHECK_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (ms_out) *ms_out = ms;
  }
  return kOk;
}

}  // extern "C"
