// Context, device memory, scans and reductions of libsimit_hip.
#include "smh_internal.h"

#include <cstdlib>
#include <cstring>

namespace smh {

int fail(smh_ctx *ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->last_error = buf;
  return code;
}

int dmalloc(smh_ctx *ctx, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ctx, SMH_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                hipGetErrorString(e));
  }
  return SMH_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// ---- exclusive scan --------------------------------------------------------
// Three-phase scan (per-block local scan, scan of block totals, add-back),
// 1024 items per 256-thread block.  Used for index construction only.
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256)
k_scan_local(const Tin *in, Tout *out, uint64_t *bsum, uint64_t n) {
  __shared__ uint64_t s[256];
  const uint64_t base = (uint64_t)blockIdx.x * 1024 + threadIdx.x * 4;
  uint64_t v[4], t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? (uint64_t)in[base + k] : 0;
    t += v[k];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint64_t add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t run = s[threadIdx.x] - t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = (Tout)run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

template <typename Tout>
__global__ void __launch_bounds__(256)
k_scan_add(Tout *out, const uint64_t *boff, uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * 1024;
  const uint64_t add = boff[blockIdx.x];
  for (int k = threadIdx.x; k < 1024; k += 256)
    if (base + k < n) out[base + k] = (Tout)((uint64_t)out[base + k] + add);
}

// exclusive scan of n values of Tin into out (uint64), writes out[n] = total
static int scan_u64(smh_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n) {
  const uint64_t nb = (n + 1023) / 1024;
  uint64_t *bsum = nullptr, *boff = nullptr;
  int rc = dmalloc(ctx, (void **)&bsum, sizeof(uint64_t) * (nb + 1));
  if (rc) return rc;
  rc = dmalloc(ctx, (void **)&boff, sizeof(uint64_t) * (nb + 1));
  if (rc) { dfree(bsum); return rc; }
  hipLaunchKernelGGL((k_scan_local<uint64_t, uint64_t>), dim3((unsigned)nb), dim3(256), 0,
                     ctx->stream, in, out, bsum, n);
  if (nb > 1) {
    rc = scan_u64(ctx, bsum, boff, nb);
    if (rc) { dfree(bsum); dfree(boff); return rc; }
    hipLaunchKernelGGL((k_scan_add<uint64_t>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                       out, boff, n);
    hipMemcpyAsync(out + n, boff + nb, sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream);
  } else {
    hipMemcpyAsync(out + n, bsum, sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream);
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  dfree(bsum);
  dfree(boff);
  if (e != hipSuccess) return fail(ctx, SMH_ERR_HIP, "scan: %s", hipGetErrorString(e));
  return SMH_OK;
}

int scan_u32(smh_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n,
             uint64_t *total) {
  const uint64_t nb = ((uint64_t)n + 1023) / 1024;
  uint64_t *bsum = nullptr, *boff = nullptr;
  int rc = dmalloc(ctx, (void **)&bsum, sizeof(uint64_t) * (nb + 1));
  if (rc) return rc;
  rc = dmalloc(ctx, (void **)&boff, sizeof(uint64_t) * (nb + 1));
  if (rc) { dfree(bsum); return rc; }
  hipLaunchKernelGGL((k_scan_local<uint32_t, uint32_t>), dim3((unsigned)(nb ? nb : 1)),
                     dim3(256), 0, ctx->stream, in, out, bsum, (uint64_t)n);
  uint64_t tot = 0;
  if (nb > 1) {
    rc = scan_u64(ctx, bsum, boff, nb);
    if (rc) { dfree(bsum); dfree(boff); return rc; }
    hipLaunchKernelGGL((k_scan_add<uint32_t>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                       out, boff, (uint64_t)n);
    SMH_HIP(ctx, hipMemcpyAsync(&tot, boff + nb, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                ctx->stream));
  } else {
    SMH_HIP(ctx, hipMemcpyAsync(&tot, bsum, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                ctx->stream));
  }
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dfree(bsum);
  dfree(boff);
  SMH_CHECK(ctx, tot <= 0xffffffffull, SMH_ERR_UNSUPPORTED,
            "scan total %llu exceeds the uint32 index range",
            (unsigned long long)tot);
  uint32_t t32 = (uint32_t)tot;
  SMH_HIP(ctx, hipMemcpyAsync(out + n, &t32, sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *total = tot;
  return SMH_OK;
}

// ---- deterministic dot --------------------------------------------------------
// Fixed grid (a function of n only), FP64 accumulation, fixed-order block
// and final reductions: the result is identical run to run.
template <typename T>
__global__ void __launch_bounds__(256)
k_dot_partial(uint64_t n, const T *__restrict__ x, const T *__restrict__ y,
              double *__restrict__ part) {
  __shared__ double s[4];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    acc += (double)x[i] * (double)y[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

// hi = fl32(v), lo = fl32(v - hi): the assemblies' SMH_F32X2 rounding
__global__ void __launch_bounds__(256)
k_split_hilo(uint64_t n, const double *__restrict__ v, float *__restrict__ hi,
             float *__restrict__ lo) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    const double x = v[i];
    const float h = (float)x;
    hi[i] = h;
    lo[i] = (float)(x - (double)h);
  }
}

__global__ void __launch_bounds__(256)
k_axpby(uint64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    y[i] = a * x[i] + b * y[i];
}

__global__ void __launch_bounds__(256)
k_sum_partials(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double s[4];
  double acc = thread_sum(part, np);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = ((s[0] + s[1]) + s[2]) + s[3];
}

}  // namespace smh

using namespace smh;

// ---- C ABI: context / memory -------------------------------------------------
extern "C" int smh_ctx_create(int device, smh_ctx **out) {
  if (!out) return SMH_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return SMH_ERR_HIP;
  }
  if (device < 0 || device >= ndev) return SMH_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return SMH_ERR_HIP;
  smh_ctx *c = new smh_ctx();
  c->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      ncu > 0)
    c->num_cu = ncu;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SMH_ERR_HIP;
  }
  c->stream = c->own_stream;
  // the allocation-free solver's scratch array: an error here is reported
  // by that entry point, the context is usable without it
  (void)scratch_setup(c);
  c->last_error.clear();
  if (hipMalloc((void **)&c->red, sizeof(double) * (kRedSlots + 64)) != hipSuccess ||
      hipMalloc((void **)&c->red_count, sizeof(unsigned) * kRedCounters) != hipSuccess ||
      hipHostMalloc((void **)&c->host_pinned, sizeof(double) * 64, hipHostMallocDefault) !=
          hipSuccess) {
    if (c->red) hipFree(c->red);
    if (c->red_count) hipFree(c->red_count);
    hipStreamDestroy(c->own_stream);
    delete c;
    return SMH_ERR_NOMEM;
  }
  void *sig = nullptr, *sig_dev = nullptr;
  if (hipHostMalloc(&sig, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer(&sig_dev, sig, 0) != hipSuccess) {
    (void)hipGetLastError();
    if (sig) hipHostFree(sig);
    hipFree(c->red);
    hipFree(c->red_count);
    hipHostFree(c->host_pinned);
    hipStreamDestroy(c->own_stream);
    delete c;
    return SMH_ERR_NOMEM;
  }
  std::memset(sig, 0, 64);
  c->signal_host = (volatile unsigned long long *)sig;
  c->signal_dev = (unsigned long long *)sig_dev;
  if (hipMemset(c->red_count, 0, sizeof(unsigned) * kRedCounters) != hipSuccess) {
    delete c;
    return SMH_ERR_HIP;
  }
  *out = c;
  return SMH_OK;
}

extern "C" int smh_ctx_destroy(smh_ctx *ctx) {
  if (!ctx) return SMH_ERR_INVALID;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream) hipStreamSynchronize(ctx->own_stream);
  ctx->cached.clear();  // cached solver workspaces and graphs
  for (auto &p : ctx->plans) dfree(p.tile_row);
  if (ctx->red) hipFree(ctx->red);
  if (ctx->red_count) hipFree(ctx->red_count);
  if (ctx->host_pinned) hipHostFree(ctx->host_pinned);
  if (ctx->signal_host) hipHostFree((void *)ctx->signal_host);
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return SMH_OK;
}

extern "C" const char *smh_ctx_last_error(const smh_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" int smh_ctx_set_stream(smh_ctx *ctx, void *s) {
  if (!ctx) return SMH_ERR_INVALID;
  ctx->stream = (hipStream_t)s;
  return SMH_OK;
}

extern "C" void *smh_ctx_get_stream(const smh_ctx *ctx) {
  return ctx ? (void *)ctx->stream : nullptr;
}

extern "C" int smh_ctx_synchronize(smh_ctx *ctx) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_ctx_describe(const smh_ctx *ctx, char *buf, size_t len) {
  if (!ctx || !buf || !len) return SMH_ERR_INVALID;
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    return SMH_ERR_HIP;
  }
  snprintf(buf, len, "device %d: %s (%s), %d compute units, %zu GiB", ctx->device, prop.name,
           prop.gcnArchName, prop.multiProcessorCount, (size_t)(prop.totalGlobalMem >> 30));
  return SMH_OK;
}

extern "C" int smh_malloc(smh_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return SMH_ERR_INVALID;
  return dmalloc(ctx, dptr, bytes);
}

extern "C" int smh_free(smh_ctx *ctx, void *dptr) {
  if (!ctx) return SMH_ERR_INVALID;
  if (dptr) SMH_HIP(ctx, hipFree(dptr));
  return SMH_OK;
}

extern "C" int smh_memcpy_htod(smh_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return SMH_ERR_INVALID;
  if (!bytes) return SMH_OK;
  SMH_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_memcpy_dtoh(smh_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return SMH_ERR_INVALID;
  if (!bytes) return SMH_OK;
  SMH_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_memset(smh_ctx *ctx, void *dst, int value, size_t bytes) {
  if (!ctx || (bytes && !dst)) return SMH_ERR_INVALID;
  if (!bytes) return SMH_OK;
  SMH_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
  return SMH_OK;
}

extern "C" int smh_dot(smh_ctx *ctx, int32_t precision, uint64_t n, const void *x,
                       const void *y, double *result) {
  if (!ctx || !result || (n && (!x || !y))) return ctx ? fail(ctx, SMH_ERR_INVALID, "smh_dot: null argument") : SMH_ERR_INVALID;
  SMH_CHECK(ctx, precision == SMH_F64 || precision == SMH_F32, SMH_ERR_INVALID,
            "smh_dot: precision must be SMH_F64 or SMH_F32");
  const unsigned g = grid_for(n, 256, 1024);
  if (precision == SMH_F64)
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(g), dim3(256), 0, ctx->stream, n,
                       (const double *)x, (const double *)y, ctx->red);
  else
    hipLaunchKernelGGL(k_dot_partial<float>, dim3(g), dim3(256), 0, ctx->stream, n,
                       (const float *)x, (const float *)y, ctx->red);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->red, (int)g,
                     ctx->red + kRedSlots);
  SMH_LAUNCH_CHECK(ctx);
  SMH_HIP(ctx, hipMemcpyAsync(ctx->host_pinned, ctx->red + kRedSlots, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *result = ctx->host_pinned[0];
  return SMH_OK;
}

extern "C" int smh_split_hilo(smh_ctx *ctx, uint64_t n, const double *v, float *hi, float *lo) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, n == 0 || (v && hi && lo), SMH_ERR_INVALID, "smh_split_hilo: null argument");
  if (!n) return SMH_OK;
  hipLaunchKernelGGL(k_split_hilo, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, n, v,
                     hi, lo);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" int smh_axpby(smh_ctx *ctx, uint64_t n, double a, const double *x, double b,
                         double *y) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, n == 0 || (x && y), SMH_ERR_INVALID, "smh_axpby: null argument");
  if (!n) return SMH_OK;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, n, a, x, b,
                     y);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}
