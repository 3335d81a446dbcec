// Runtime-compiled element kernels and the fixed kernels of the program
// interpreter.
//
// The reference JIT-compiles every Simit function (LLVM IR, and NVVM IR for
// its GPU backend, loaded with cuModuleLoadData in gpu_function.cpp).  Here
// the front-end generates HIP C++ for each map's element function; hiprtc
// compiles it for the device's architecture (gfx950) and the module is
// loaded and launched through these entry points.  Everything else a
// program's exported function does runs through fixed kernels: the staged
// `reduce +` below, the vector statements (smh_vec_op), SpMV and dots.
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>

#include "smh_internal.h"

using namespace smh;

struct smh_module {
  hipModule_t mod = nullptr;
  std::map<std::string, hipFunction_t> fns;
};

namespace {

void copy_log(const std::string &s, char *log, size_t log_len) {
  if (!log || !log_len) return;
  const size_t n = std::min(s.size(), log_len - 1);
  std::memcpy(log, s.data(), n);
  log[n] = '\0';
}

// Entry i (bs components) of a map's output from its staged writes
// src[ptr[i]..ptr[i+1]), listed in the order the reference performs them
// (elements ascending, then the element function's execution order).
// `reduce +` (taken == NULL): out = ((0 + w0) + w1) + ..., the reference's
// compound writes into the zeroed global tensor (a write a branch skipped is
// staged as 0 and adds nothing).  No reduction: plain writes, so out takes the
// last write taken, and an entry no element wrote keeps its value.
__global__ void __launch_bounds__(256) k_stage_reduce(uint64_t nout, uint32_t bs,
                                                      const uint64_t *__restrict__ ptr,
                                                      const uint32_t *__restrict__ src,
                                                      const double *__restrict__ stage,
                                                      const uint8_t *__restrict__ taken,
                                                      double *__restrict__ out) {
#pragma clang fp contract(off)
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout * bs) return;
  const uint64_t i = t / bs;
  const uint32_t c = (uint32_t)(t - i * bs);
  if (!taken) {
    double acc = 0.0;
    for (uint64_t k = ptr[i]; k < ptr[i + 1]; ++k) acc += stage[(uint64_t)src[k] * bs + c];
    out[t] = acc;
    return;
  }
  for (uint64_t k = ptr[i + 1]; k > ptr[i]; --k)
    if (taken[src[k - 1]]) {
      out[t] = stage[(uint64_t)src[k - 1] * bs + c];
      return;
    }
}

// one rounding per operation, no contraction: the elementwise statements of a
// Simit function evaluate exactly as the reference's scalar loops do
__global__ void __launch_bounds__(256) k_vec_op(int op, uint64_t n, double a, const double *x,
                                                double b, const double *y, double *out) {
#pragma clang fp contract(off)
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double r;
    switch (op) {
      case SMH_VEC_LINCOMB: {
        const double ax = a * x[i];
        r = y ? ax + b * y[i] : ax;
        break;
      }
      case SMH_VEC_DIVS: r = x[i] / a; break;
      case SMH_VEC_EMUL: r = x[i] * y[i]; break;
      case SMH_VEC_EDIV: r = x[i] / y[i]; break;
      case SMH_VEC_FILL: r = a; break;
      case SMH_VEC_ADDS: r = x[i] + a; break;
      case SMH_VEC_SDIV: r = a / x[i]; break;
      default: r = 0.0; break;
    }
    out[i] = r;
  }
}

}  // namespace

extern "C" int smh_hiprtc_compile(const char *source, const char *name, const char *arch,
                                  void **image, size_t *bytes, char *log, size_t log_len) {
  if (!source || !image || !bytes) {
    copy_log("smh_hiprtc_compile: null argument", log, log_len);
    return SMH_ERR_INVALID;
  }
  *image = nullptr;
  *bytes = 0;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, name ? name : "simit_map.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    copy_log("hiprtcCreateProgram failed", log, log_len);
    return SMH_ERR_HIP;
  }
  const std::string off = std::string("--offload-arch=") + (arch && *arch ? arch : "gfx950");
  // the library's own flags: no FMA contraction (the element arithmetic of
  // the reference rounds every operation) and no packed FP32 instructions
  const char *opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, sizeof(opts) / sizeof(opts[0]), opts);
  size_t ls = 0;
  std::string text;
  if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
    text.resize(ls);
    hiprtcGetProgramLog(prog, &text[0]);
  }
  if (rc != HIPRTC_SUCCESS) {
    copy_log("hiprtc: " + std::string(hiprtcGetErrorString(rc)) + "\n" + text, log, log_len);
    hiprtcDestroyProgram(&prog);
    return SMH_ERR_INVALID;
  }
  size_t cs = 0;
  if (hiprtcGetCodeSize(prog, &cs) != HIPRTC_SUCCESS || cs == 0) {
    copy_log("hiprtc produced no code", log, log_len);
    hiprtcDestroyProgram(&prog);
    return SMH_ERR_HIP;
  }
  char *buf = (char *)std::malloc(cs);
  if (!buf || hiprtcGetCode(prog, buf) != HIPRTC_SUCCESS) {
    std::free(buf);
    copy_log("hiprtcGetCode failed", log, log_len);
    hiprtcDestroyProgram(&prog);
    return SMH_ERR_HIP;
  }
  hiprtcDestroyProgram(&prog);
  copy_log(text, log, log_len);
  *image = buf;
  *bytes = cs;
  return SMH_OK;
}

extern "C" void smh_image_free(void *image) { std::free(image); }

extern "C" int smh_module_load(smh_ctx *ctx, const void *image, size_t bytes,
                               smh_module **out) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, image && bytes && out, SMH_ERR_INVALID, "smh_module_load: null argument");
  hipDeviceProp_t prop{};
  SMH_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
  SMH_CHECK(ctx, std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, SMH_ERR_UNSUPPORTED,
            "smh_module_load: device %s is not gfx950", prop.gcnArchName);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  auto m = std::make_unique<smh_module>();
  SMH_HIP(ctx, hipModuleLoadData(&m->mod, image));
  *out = m.release();
  return SMH_OK;
}

extern "C" int smh_module_launch(smh_ctx *ctx, smh_module *m, const char *kernel,
                                 uint64_t nthreads, const void *args, size_t args_bytes) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, m && kernel && (args || !args_bytes), SMH_ERR_INVALID,
            "smh_module_launch: null argument");
  if (!nthreads) return SMH_OK;
  SMH_CHECK(ctx, nthreads <= (uint64_t)0x7fffffff * 256, SMH_ERR_UNSUPPORTED,
            "smh_module_launch: %llu threads", (unsigned long long)nthreads);
  auto it = m->fns.find(kernel);
  if (it == m->fns.end()) {
    hipFunction_t f = nullptr;
    SMH_HIP(ctx, hipModuleGetFunction(&f, m->mod, kernel));
    it = m->fns.emplace(kernel, f).first;
  }
  size_t sz = args_bytes;
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args),
                 HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
  const unsigned grid = (unsigned)((nthreads + 255) / 256);
  SMH_HIP(ctx, hipModuleLaunchKernel(it->second, grid, 1, 1, 256, 1, 1, 0, ctx->stream, nullptr,
                                     cfg));
  return SMH_OK;
}

extern "C" int smh_module_destroy(smh_ctx *ctx, smh_module *m) {
  if (!m) return SMH_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  const hipError_t e = hipModuleUnload(m->mod);
  delete m;
  if (e != hipSuccess && ctx)
    return fail(ctx, SMH_ERR_HIP, "hipModuleUnload failed: %s", hipGetErrorString(e));
  return SMH_OK;
}

extern "C" int smh_stage_reduce(smh_ctx *ctx, uint64_t nout, uint32_t bs, const uint64_t *ptr,
                                const uint32_t *src, const double *stage, const uint8_t *taken,
                                double *out) {
  if (!ctx) return SMH_ERR_INVALID;
  if (!nout || !bs) return SMH_OK;
  SMH_CHECK(ctx, ptr && out, SMH_ERR_INVALID, "smh_stage_reduce: null argument");
  const uint64_t nt = nout * bs;
  hipLaunchKernelGGL(k_stage_reduce, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, nout,
                     bs, ptr, src, stage, taken, out);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" int smh_vec_op(smh_ctx *ctx, int32_t op, uint64_t n, double a, const double *x,
                          double b, const double *y, double *out) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, op >= SMH_VEC_LINCOMB && op <= SMH_VEC_SDIV, SMH_ERR_INVALID,
            "smh_vec_op: unknown operation %d", op);
  if (!n) return SMH_OK;
  const bool needs_x = op != SMH_VEC_FILL;
  const bool needs_y = op == SMH_VEC_EMUL || op == SMH_VEC_EDIV;
  SMH_CHECK(ctx, out && (!needs_x || x) && (!needs_y || y), SMH_ERR_INVALID,
            "smh_vec_op: null argument");
  hipLaunchKernelGGL(k_vec_op, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, op, n, a,
                     x, b, op == SMH_VEC_LINCOMB || needs_y ? y : nullptr, out);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}
