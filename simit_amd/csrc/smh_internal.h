// Internal definitions shared by the HIP translation units of libsimit_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/simit_hip.h"

// Row-tile partition of one blocked-CSR pattern (see smh_linalg.h), cached
// per context and keyed by the pattern's rowptr buffer, row and block count.
struct smh_tile_plan {
  const uint32_t *rowptr = nullptr;
  uint32_t nrows = 0;
  uint64_t nnz = 0;
  uint32_t maxrow = 0;
  uint32_t tile_blocks = 0;
  uint32_t tprime = 0;
  uint32_t ntiles = 0;
  uint32_t *tile_row = nullptr;  // device [ntiles+1]
};

// Something a solver builds once per pattern and keeps in the context
// (device workspace, captured graphs); dropped by smh_bsr_forget(rowptr) and
// with the context.
struct smh_cached {
  const uint32_t *rowptr = nullptr;
  virtual ~smh_cached() = default;
};

struct smh_ctx {
  int device = 0;
  int num_cu = 256;  // compute units of the device (persistent grids)
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string last_error;
  std::vector<smh_tile_plan> plans;
  std::vector<std::unique_ptr<smh_cached>> cached;
  struct Tuned {
    const uint32_t *rowptr;
    uint32_t nrows, bdim;
    uint64_t nnz;
    unsigned grid;
    int shape;  // batched SpMV tile shape
    int prec;   // the inner precision it was tuned for
  };
  std::vector<Tuned> tuned;  // autotuned batched SpMV shapes and grids by pattern
  // small device scratch for reductions: [0..kRedSlots) partials + counters
  double *red = nullptr;
  unsigned *red_count = nullptr;
  double *host_pinned = nullptr;  // pinned host mirror for scalar readbacks
  // pinned host words the batched solver's replays write their progress to
  // (host pointer and its device mapping)
  volatile unsigned long long *signal_host = nullptr;
  unsigned long long *signal_dev = nullptr;
  // this device's copy of the library's static scratch array (the
  // allocation-free solver's workspace), looked up once per device
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
};

struct smh_index {
  uint32_t nrows = 0;
  uint64_t nnz = 0;
  uint32_t *rowptr = nullptr;
  uint32_t *colidx = nullptr;
  uint32_t *wave_row = nullptr;     // row holding block 64*w (assembly waves)
  uint32_t nsets = 0;
  std::vector<uint32_t *> ve_ptr;   // per set, [nrows+1]
  std::vector<uint32_t *> ve_elem;  // per set
  std::vector<uint64_t> n_inc;      // per set incidences
  std::vector<uint64_t> ne;         // per set element count
  std::vector<uint32_t> card;       // per set cardinality
  std::vector<const int32_t *> ep;  // per set endpoints (borrowed, device)
  // Per set, made by the first assembly over the set and kept with the pattern:
  // each incidence's endpoint slot, each block's contributing incidences, and
  // which 64-block waves the fast assembly path takes.
  struct AsmPlan {
    uint8_t *slot = nullptr;   // [n_inc] slot of the row in the incidence's element
    void *mask = nullptr;      // [nnz] bit k: the row's k-th incidence writes the block
    uint32_t *rank = nullptr;  // [nnz] per contribution (mask order, 2 bits): the rank of
                               // the column's slot among the element's other slots
    uint8_t *wnd = nullptr;    // [waves] diagonal blocks in the wave (255: generic path)
    uint8_t *wnc = nullptr;    // [waves] LDS chunks the wave's pairs are split into
    int bits = 0;              // 32 or 64 mask bits (0: not built yet)
    uint64_t nslow[3] = {0, 0, 0};  // generic waves for 1x1, 2x2 and 3x3 blocks
  };
  mutable std::vector<AsmPlan> asm_plan;
};

namespace smh {

constexpr int kRedSlots = 4096;  // max partial sums per reduction
constexpr int kRedCounters = 16;

int fail(smh_ctx *ctx, int code, const char *fmt, ...);

#define SMH_HIP(ctx, call)                                                   \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return smh::fail((ctx), SMH_ERR_HIP, "%s failed: %s (%s:%d)", #call,  \
                       hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

#define SMH_CHECK(ctx, cond, code, ...)                           \
  do {                                                            \
    if (!(cond)) return smh::fail((ctx), (code), __VA_ARGS__);    \
  } while (0)

#define SMH_LAUNCH_CHECK(ctx)                                               \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess)                                                   \
      return smh::fail((ctx), SMH_ERR_HIP, "kernel launch failed: %s (%s:%d)", \
                       hipGetErrorString(e_), __FILE__, __LINE__);          \
  } while (0)

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 30) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// exclusive scan of uint32 counts into out[0..n] (out[n] = total); returns
// SMH status; total (uint64) written to *total on the host (synchronises).
int scan_u32(smh_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n,
             uint64_t *total);

// device allocation helpers that report through ctx
int dmalloc(smh_ctx *ctx, void **p, size_t bytes);
void dfree(void *p);

// the row-tile plan of a pattern (built on first use, synchronises then)
int tile_plan(smh_ctx *ctx, const uint32_t *rowptr, uint32_t nrows, uint64_t nnz,
              uint32_t tile_blocks, smh_tile_plan *out);
// drop cached plans of a rowptr buffer (before it is freed or reused)
void forget_plans(smh_ctx *ctx, const uint32_t *rowptr);
// Builds the index's assembly plan for the set with these endpoints (before a
// stream capture that records the assembly).
int prepare_assembly(smh_ctx *ctx, const smh_index *idx, const int32_t *ep, uint32_t card);
// device address of the static scratch array on ctx's (current) device
int scratch_setup(smh_ctx *ctx);

}  // namespace smh

// ---- fixed-order partial sums ----------------------------------------------
// This thread's share of a sum over part[i * st]: i = threadIdx.x,
// threadIdx.x + nt, ... added in that order, the loads issued 16 at a time
// (a load-add loop waits one L2 round trip per value).
__device__ __forceinline__ double thread_sum(const double *part, int np, int st = 1,
                                             int nt = 256) {
  constexpr int U = 16;
  double acc = 0.0;
  int i = threadIdx.x;
  for (; i + (U - 1) * nt < np; i += U * nt) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = part[(size_t)(i + u * nt) * st];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
  }
  for (; i < np; i += nt) acc += part[(size_t)i * st];
  return acc;
}

// ---- wave-level helpers (64-wide wavefronts) ------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sumf(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
