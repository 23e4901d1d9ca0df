// graph.hip -- capture of a device-resident step loop as a hipGraph, and
// stream-ordered event timing.
//
// Replaces the host loop of JaxModel.run (jaxabm/model.py:240-241: one
// Python iteration and a dozen un-jitted dispatches per step) by one graph
// launch for the whole run.  Captured code records kernels and device copies
// only: buffers that must start at zero are cleared by fill kernels, never by
// memset nodes.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "jaxabm_amd.h"

struct abm_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

using namespace abm;

extern "C" {

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
