// rtc.cpp -- run-time compilation and launch of generated agent-update
// kernels.
//
// The reference maps a per-agent Python function over the agents with
// jax.vmap inside AgentCollection.update (jaxabm/agent.py:156-177), which XLA
// fuses.  jaxabm_amd/fuse.py traces such a function into the text of one HIP
// kernel; this file compiles that text for the GPU with hiprtc, loads it, and
// launches it on the caller's stream.  A generated kernel takes one argument:
// a struct of 64-bit words (pointers, sizes and scalar parameters) that the
// caller packs.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "common.hpp"
#include "jaxabm_amd.h"

namespace {

struct RtcKernel {
  hipModule_t module;
  hipFunction_t function;
};

}  // namespace

using namespace abm;

extern "C" {

int abm_rtc_compile(const char *src, const char *arch, void **code, size_t *size) {
  ABM_REQUIRE(src && arch && code && size, "abm_rtc_compile: null argument");
  *code = nullptr;
  *size = 0;
  hiprtcProgram prog = nullptr;
  if (hiprtcCreateProgram(&prog, src, "abm_fused.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    set_error("abm_rtc_compile: hiprtcCreateProgram failed");
    return kErrInvalid;
  }
  const std::string arch_opt = std::string("--offload-arch=") + arch;
  // the library's own flags: no contraction into fused multiply-adds, no
  // packed FP32 instructions
  const char *opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      log.resize(n);
      hiprtcGetProgramLog(prog, &log[0]);
    }
    hiprtcDestroyProgram(&prog);
    set_error("abm_rtc_compile: " + std::string(hiprtcGetErrorString(rc)) + "\n" + log);
    return kErrInvalid;
  }
  size_t n = 0;
  if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
    hiprtcDestroyProgram(&prog);
    set_error("abm_rtc_compile: no code object");
    return kErrInvalid;
  }
  void *buf = malloc(n);
  if (!buf) {
    hiprtcDestroyProgram(&prog);
    set_error("abm_rtc_compile: out of host memory");
    return kErrNoMem;
  }
  if (hiprtcGetCode(prog, (char *)buf) != HIPRTC_SUCCESS) {
    free(buf);
    hiprtcDestroyProgram(&prog);
    set_error("abm_rtc_compile: hiprtcGetCode failed");
    return kErrInvalid;
  }
  hiprtcDestroyProgram(&prog);
  *code = buf;
  *size = n;
  return kOk;
}

int abm_rtc_code_free(void *code) {
  free(code);
  return kOk;
}

int abm_rtc_load(const void *code, const char *kernel, void **handle) {
  ABM_REQUIRE(code && kernel && handle, "abm_rtc_load: null argument");
  *handle = nullptr;
  hipModule_t mod = nullptr;
  ABM_CHECK_HIP(hipModuleLoadData(&mod, code));
  hipFunction_t fn = nullptr;
  const hipError_t e = hipModuleGetFunction(&fn, mod, kernel);
  if (e != hipSuccess) {
    (void)hipModuleUnload(mod);
    return hip_fail(e, "hipModuleGetFunction");
  }
  *handle = new RtcKernel{mod, fn};
  return kOk;
}

int abm_rtc_unload(void *handle) {
  if (!handle) return kOk;
  RtcKernel *k = (RtcKernel *)handle;
  const hipError_t e = hipModuleUnload(k->module);
  delete k;
  if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
  return kOk;
}

int abm_rtc_launch(void *handle, uint32_t grid, uint32_t block, const void *args,
                   size_t args_bytes, void *stream) {
  ABM_REQUIRE(handle && args && args_bytes > 0 && args_bytes % 8 == 0 && grid >= 1 &&
                  block >= 1 && block <= 1024,
              "abm_rtc_launch: bad arguments");
  RtcKernel *k = (RtcKernel *)handle;
  size_t bytes = args_bytes;
  void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args),
                    HIP_LAUNCH_PARAM_BUFFER_SIZE, &bytes, HIP_LAUNCH_PARAM_END};
  ABM_CHECK_HIP(hipModuleLaunchKernel(k->function, grid, 1, 1, block, 1, 1, 0,
                                      (hipStream_t)stream, nullptr, config));
  return kOk;
}

}  // extern "C"
