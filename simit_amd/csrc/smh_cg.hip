// Blocked SpMV entry point and the device-resident CG / PCG solver.
//
// The CG loop is what Simit programs write inline (apps/springs/isprings.sim:
// 77-93, apps/fem/fem_linear.sim:211-230, test/input/program/cg.sim:25-53,
// apps/thermal/thermal.sim:75-130).  One iteration here is four launches on
// one stream (five with the learned correction's pass), with every scalar
// kept on the device:
//   K1  q = A p (row tiles) + FP64 partial of p'q per workgroup;
//   K2  every workgroup sums the p'q partials in one fixed order (so all
//       hold the same alpha = r'z / p'q); x += alpha p; r -= alpha q;
//       z = M r; partials of z'r and r'r;
//   R2  one workgroup: z'r, r'r, beta, the stopping test;
//   K3  p = z + beta p.
// Cross-workgroup results only pass through kernel boundaries (no
// device-wide fences inside a kernel), and every reduction has a fixed
// shape, so a solve is bitwise repeatable.  The host reads the state every
// check_every iterations; iterations enqueued past convergence exit at once.
//
// Mixed precision: the inner iteration streams the FP32 matrix (hi part)
// with FP32 vectors and FP64 dots; x is accumulated in FP64.  Whenever the
// iterated residual fell by reliable_delta since the last update, or claims
// convergence, the iteration pauses right after its x/r update and a
// reliable update recomputes r = b - A x in FP64 from hi + lo (~2^-48
// relative to the FP64 matrix); beta and p then continue from the replaced
// residual.  Convergence is always certified by that FP64 residual.
#include "smh_linalg.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>

namespace smh {

// ---- row-tile plans ---------------------------------------------------------------
__global__ void k_rowmax(const uint32_t *__restrict__ rowptr, uint32_t nrows, unsigned *out) {
  unsigned m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, rowptr[i + 1] - rowptr[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_tile_rows(const uint32_t *__restrict__ rowptr, uint32_t nrows, uint32_t tprime,
                            uint32_t ntiles, uint32_t *__restrict__ tile_row) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ntiles) return;
  if (k == ntiles) {
    tile_row[k] = nrows;
    return;
  }
  const uint64_t key = (uint64_t)k * tprime;
  uint32_t lo = 0, hi = nrows;  // first r in [0, nrows] with rowptr[r] >= key
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)rowptr[mid] < key) lo = mid + 1; else hi = mid;
  }
  tile_row[k] = k == 0 ? 0 : lo;
}

int tile_plan(smh_ctx *ctx, const uint32_t *rowptr, uint32_t nrows, uint64_t nnz,
              uint32_t tile_blocks, smh_tile_plan *out) {
  for (auto &p : ctx->plans)
    if (p.rowptr == rowptr && p.nrows == nrows && p.nnz == nnz && p.tile_blocks == tile_blocks) {
      *out = p;
      return SMH_OK;
    }
  unsigned *dmax = nullptr;
  int rc = dmalloc(ctx, (void **)&dmax, sizeof(unsigned));
  if (rc) return rc;
  SMH_HIP(ctx, hipMemsetAsync(dmax, 0, sizeof(unsigned), ctx->stream));
  if (nrows)
    hipLaunchKernelGGL(k_rowmax, dim3(grid_for(nrows, 256, 2048)), dim3(256), 0, ctx->stream,
                       rowptr, nrows, dmax);
  unsigned maxrow = 0;
  SMH_HIP(ctx, hipMemcpyAsync(&maxrow, dmax, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dfree(dmax);
  smh_tile_plan p;
  p.rowptr = rowptr;
  p.nrows = nrows;
  p.nnz = nnz;
  p.maxrow = maxrow;
  p.tile_blocks = tile_blocks;
  if (maxrow <= tile_blocks / 2) {
    p.tprime = tile_blocks - (maxrow ? maxrow : 1) + 1;
    uint64_t nt = (nnz + p.tprime - 1) / p.tprime;
    if (nt < 1) nt = 1;
    SMH_CHECK(ctx, nt < 0xffffffffull, SMH_ERR_UNSUPPORTED, "tile plan: too many tiles");
    p.ntiles = (uint32_t)nt;
    rc = dmalloc(ctx, (void **)&p.tile_row, sizeof(uint32_t) * (p.ntiles + 1));
    if (rc) return rc;
    hipLaunchKernelGGL(k_tile_rows, dim3(grid_for(p.ntiles + 1, 256)), dim3(256), 0, ctx->stream,
                       rowptr, nrows, p.tprime, p.ntiles, p.tile_row);
    SMH_LAUNCH_CHECK(ctx);
    // a plan is complete when it is returned: its users may launch on
    // another stream (the batched solver runs on the context's own stream)
    SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  ctx->plans.push_back(p);
  *out = p;
  return SMH_OK;
}

// ---- the allocation-free solver's workspace -------------------------------------
// One static array per device, held until the process ends.  Host code uses
// the array only through the device address hipGetSymbolAddress returns for
// the current device (the name itself is a host-side stand-in).
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int scratch_setup(smh_ctx *ctx) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *ptrs[kMaxDevices] = {};
  SMH_CHECK(ctx, ctx->device >= 0 && ctx->device < kMaxDevices, SMH_ERR_UNSUPPORTED,
            "scratch: device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ctx, SMH_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                  hipGetErrorString(e));
    }
    ptrs[ctx->device] = p;
  }
  ctx->scratch = ptrs[ctx->device];
  ctx->scratch_bytes = kScratchBytes;
  return SMH_OK;
}

void forget_plans(smh_ctx *ctx, const uint32_t *rowptr) {
  for (size_t i = 0; i < ctx->cached.size();) {
    if (ctx->cached[i]->rowptr == rowptr) {
      hipStreamSynchronize(ctx->own_stream);
      hipStreamSynchronize(ctx->stream);
      ctx->cached.erase(ctx->cached.begin() + i);
    } else {
      ++i;
    }
  }
  for (size_t i = 0; i < ctx->tuned.size();) {
    if (ctx->tuned[i].rowptr == rowptr) ctx->tuned.erase(ctx->tuned.begin() + i);
    else ++i;
  }
  for (size_t i = 0; i < ctx->plans.size();) {
    if (ctx->plans[i].rowptr == rowptr) {
      hipStreamSynchronize(ctx->stream);
      dfree(ctx->plans[i].tile_row);
      ctx->plans.erase(ctx->plans.begin() + i);
    } else {
      ++i;
    }
  }
}

// The lane-planar copy of a pattern's FP32 3x3 values that the CG solver's
// inner SpMV streams: one device buffer per pattern, kept in the context
// (dropped with the pattern's other plans) and refilled by each solve.
struct PlanarBuf : smh_cached {
  uint64_t nnz = 0;
  float *vals = nullptr;
  ~PlanarBuf() override { dfree(vals); }
};

int planar_buffer(smh_ctx *ctx, const smh_bsr *A, float **out) {
  for (auto &c : ctx->cached) {
    auto *pb = dynamic_cast<PlanarBuf *>(c.get());
    if (pb && pb->rowptr == A->rowptr && pb->nnz == A->nnz) {
      *out = pb->vals;
      return SMH_OK;
    }
  }
  std::unique_ptr<PlanarBuf> pb(new PlanarBuf);
  pb->rowptr = A->rowptr;
  pb->nnz = A->nnz;
  const int rc = dmalloc(ctx, (void **)&pb->vals, sizeof(float) * planar_floats(A->nnz));
  if (rc) return rc;
  *out = pb->vals;
  ctx->cached.push_back(std::move(pb));
  return SMH_OK;
}

}  // namespace smh

namespace {
using namespace smh;

// ---- plain SpMV ---------------------------------------------------------------
#pragma clang fp contract(off)
// thread per scalar row, ascending ij then bj, separate multiply and add:
// bit-identical to the CPU loop nest.  Also the path for rows longer than
// half a tile.
__global__ void k_bsrmv_exact(uint32_t nrows, int bd, const uint32_t *__restrict__ rowptr,
                              const uint32_t *__restrict__ colidx, const double *__restrict__ vals,
                              const double *__restrict__ x, double *__restrict__ y) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)nrows * bd) return;
  const uint32_t i = (uint32_t)(g / bd);
  const int bi = (int)(g % bd);
  const int bb = bd * bd;
  double tmp = 0.0;
  for (uint32_t ij = rowptr[i]; ij < rowptr[i + 1]; ++ij) {
    const uint32_t j = colidx[ij];
    for (int bj = 0; bj < bd; ++bj) {
      const double p = vals[(uint64_t)ij * bb + bi * bd + bj] * x[(uint64_t)j * bd + bj];
      tmp = tmp + p;
    }
  }
  y[g] = tmp;
}
#pragma clang fp contract(fast)


template <typename T, int BD>
struct EpiStore {
  T *y;
  __device__ void operator()(uint32_t r, const T *s) {
    for (int a = 0; a < BD; ++a) y[(uint64_t)r * BD + a] = s[a];
  }
};

template <int PREC, int BD>
__global__ void __launch_bounds__(TileShape<BD>::kThreads)
k_bsrmv_tiles(TileDev A, const typename Val<PREC>::T *__restrict__ x,
              typename Val<PREC>::T *__restrict__ y) {
  using T = typename Val<PREC>::T;
  __shared__ TileSmem<PREC, BD> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XPlain<T> xg{x, BD};
  EpiStore<T, BD> epi{y};
  spmv_tiles<PREC, BD>(A, k0, k1, sm, xg, epi);
}

// y = A x over lane-planar FP32 3x3 values (SMH_F32P), float vectors
__global__ void __launch_bounds__(TileShapePl::kThreads)
k_bsrmv_planar(TileDev A, const float *__restrict__ x, float *__restrict__ y) {
  __shared__ TileSmem<SMH_F32, 3, TileShapePl> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XPlain<float> xg{x, 3};
  EpiStore<float, 3> epi{y};
  spmv_tiles<SMH_F32, 3>(A, k0, k1, sm, xg, epi);
}

// row-major FP32 3x3 values -> lane-planar (one thread per block; the
// blocks past nnz in the last group are written as zeros)
__global__ void __launch_bounds__(256)
k_to_planar(uint64_t nnz, const float *__restrict__ v, float *__restrict__ out) {
  const uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= ((nnz + 63) & ~63ull)) return;
  float e[9];
  if (b < nnz) {
    for (int k = 0; k < 9; ++k) e[k] = v[b * 9 + k];
  } else {
    for (int k = 0; k < 9; ++k) e[k] = 0.0f;
  }
  typedef float v4 __attribute__((ext_vector_type(4)));
  float *g = out + (b >> 6) * kPlanarGroup;
  const uint64_t l = b & 63u;
  *(v4 *)(g + 4 * l) = v4{e[0], e[1], e[2], e[3]};
  *(v4 *)(g + 256 + 4 * l) = v4{e[4], e[5], e[6], e[7]};
  g[512 + l] = e[8];
}

// ---- CG device state ------------------------------------------------------------
struct CGState {
  double rz, pq, alpha, beta, rr, rr_true, thr, rr_ref, bb, delta2;
  int iter, done, status, want_ru, claimed, maxiters, precond_stop, mixed, restart, replace_every;
  int precond;  // a preconditioner is applied: z'r must stay positive
};

__device__ __forceinline__ bool finite_pos(double v) { return v > 0.0 && v < INFINITY; }

// fixed-order sum of np values (stride st) by one 256-thread workgroup (the
// update kernels sum the SpMV's 8,192 partials this way in every workgroup:
// with one load in flight per thread that held each workgroup ~4 us before
// its first streaming load)
__device__ double wg_sum(const double *part, int np, int st, double *s4) {
  return block_sum256(thread_sum(part, np, st), s4);
}

// entry (a, b), b <= a, of a lower-triangular 3x3 block stored as its six
// entries row by row
__host__ __device__ constexpr int tri_idx(int a, int b) { return a * (a + 1) / 2 + b; }

// Linv [n][9] (row-major, zero above the diagonal) -> [n][6] (tri_idx)
__global__ void k_pack_linv(uint32_t n, const float *__restrict__ L9, float *__restrict__ L6) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * 6) return;
  const uint64_t i = t / 6;
  const int k = (int)(t - i * 6), a = k >= 3 ? 2 : (k >= 1 ? 1 : 0), b = k - tri_idx(a, 0);
  L6[t] = L9[i * 9 + 3 * a + b];
}

template <int BD>
__device__ __forceinline__ void apply_precond(int pc, const double *r, double *z,
                                              const void *dinv, bool f32, uint64_t i) {
  if (pc == SMH_PRECOND_NONE) {
    for (int a = 0; a < BD; ++a) z[a] = r[a];
  } else if (BD == 1 && pc == SMH_PRECOND_LEARNED) {
    z[0] = (double)((const float *)dinv)[i] * r[0];  // u = Linv r, Linv = 1/sqrt(a_ii)
  } else if (BD == 3 && pc == SMH_PRECOND_LEARNED) {
    // first half of the learned preconditioner: u = Linv r (the rest,
    // z = Linv^T (u + C u), needs the neighbours' u: k_cg_lapply); Linv
    // packed as its six lower-triangular entries (tri_idx)
    const float *L = (const float *)dinv + i * 6;
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b <= a; ++b) s += (double)L[tri_idx(a, b)] * r[b];
      z[a] = s;
    }
  } else if (BD == 1 || pc == SMH_PRECOND_JACOBI) {
    for (int a = 0; a < BD; ++a)
      z[a] = r[a] * (f32 ? (double)((const float *)dinv)[i * BD + a]
                         : ((const double *)dinv)[i * BD + a]);
  } else {
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b)
        s += (f32 ? (double)((const float *)dinv)[i * 9 + 3 * a + b]
                  : ((const double *)dinv)[i * 9 + 3 * a + b]) * r[b];
      z[a] = s;
    }
  }
}

// K1: q = A p and the workgroup partial of p'q.  With dlo (mixed precision),
// the inner operator is hi + blockdiag(sum_j lo_ij): the FP32 values plus
// each row's rounding residuals moved onto its diagonal block, so its block
// row sums equal the exact matrix's and the difference to the exact matrix
// annihilates constant (rigid translation) vectors.
template <typename VT>
struct EpiPq {
  VT *q;
  const VT *p;
  const float *dlo;
  double d;
  int bd;
  // AT: the row sums' type (double for FP32 vectors: q is rounded once, after
  // the low parts' term)
  template <typename AT>
  __device__ void operator()(uint32_t r, const AT *s) {
    AT sv[3];
    for (int a = 0; a < bd; ++a) sv[a] = s[a];
    if (dlo) {
      const float *D = dlo + (uint64_t)r * bd * bd;
      for (int a = 0; a < bd; ++a)
        for (int b = 0; b < bd; ++b) sv[a] += (AT)D[bd * a + b] * (AT)p[(uint64_t)r * bd + b];
    }
    for (int a = 0; a < bd; ++a) {
      const uint64_t o = (uint64_t)r * bd + a;
      const VT qv = (VT)sv[a];
      q[o] = qv;
      d += (double)qv * (double)p[o];
    }
  }
};

// The CG inner SpMV's accumulation precision: FP32 values times FP32 vectors
// are summed in FP64 (kF32D's loop: every FP32 x FP32 product is exact in
// FP64, so a row's sum carries one rounding, that of q).  FP32 sums lose the
// low modes' small A p to the cancellation of large terms: on the 32^3 x 5
// and 48^3 x 5 tet systems they cost 27% and 41% more PCG iterations than
// FP64 sums, which match FP64 vectors' counts.
template <int IP>
constexpr int cg_acc_prec() {
  return IP == SMH_F32 ? kF32D : IP;
}
// Gather of FP32/FP64 vector entries widened to the loop's type AT.
template <typename VT, typename AT>
struct XWide {
  const VT *x;
  int bd;
  __device__ AT operator()(uint32_t j, int b) const { return (AT)x[(uint64_t)j * bd + b]; }
  template <int BD>
  __device__ void row(uint32_t j, AT *out) const {
    VT t[BD];
    load_row<VT, BD, false>(x + (uint64_t)j * BD, t);
    for (int b = 0; b < BD; ++b) out[b] = (AT)t[b];
  }
};

// dlo_i = sum over the blocks of row i of the low parts of an SMH_F32X2
// matrix; one thread per (row, block entry), so a wave reads whole blocks of
// consecutive rows (ascending block order, FP64 accumulation)
__global__ void k_lo_rowsum(uint32_t n, int bb, const uint32_t *__restrict__ rp,
                            const float *__restrict__ lo, float *__restrict__ dlo) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * bb) return;
  const uint32_t i = (uint32_t)(t / bb);
  const int q = (int)(t - (uint64_t)i * bb);
  double s = 0.0;
  for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) s += (double)lo[(uint64_t)k * bb + q];
  dlo[t] = (float)s;
}

template <int IP, int BD>
__global__ void __launch_bounds__(TileShape<BD>::kThreads)
k_cg_spmv(TileDev A, const typename Val<IP>::T *__restrict__ p, typename Val<IP>::T *__restrict__ q,
          const CGState *S, const float *dlo, double *part) {
  using VT = typename Val<IP>::T;
  constexpr int AP = cg_acc_prec<IP>();
  __shared__ TileSmem<AP, BD, TileShape<BD>> sm;
  if (S->done) return;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XWide<VT, typename Val<AP>::T> xg{p, BD};
  EpiPq<VT> epi{q, p, dlo, 0.0, BD};
  spmv_tiles<AP, BD>(A, k0, k1, sm, xg, epi);
  const double t = block_sum<TileShape<BD>::kWaves>(epi.d, sm.red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// K1 over lane-planar FP32 3x3 values (the mixed solver's inner SpMV)
template <int IP>
__global__ void __launch_bounds__(TileShapePlS::kThreads)
k_cg_spmv_pl(TileDev A, const typename Val<IP>::T *__restrict__ p,
             typename Val<IP>::T *__restrict__ q, const CGState *S, const float *dlo,
             double *part) {
  using VT = typename Val<IP>::T;
  constexpr int AP = cg_acc_prec<IP>();
  __shared__ TileSmem<AP, 3, TileShapePlS> sm;
  if (S->done) return;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XWide<VT, typename Val<AP>::T> xg{p, 3};
  EpiPq<VT> epi{q, p, dlo, 0.0, 3};
  spmv_tiles<AP, 3>(A, k0, k1, sm, xg, epi);
  const double t = block_sum<TileShapePlS::kWaves>(epi.d, sm.red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// K2: alpha = z'r / p'q (R1 folded in: every workgroup sums K1's np
// partials of p'q in the same fixed order, so all hold the same alpha, and
// workgroup 0 records it); x += alpha p; r -= alpha q; z = M r; partials of
// z'r and r'r.  z32: z is stored as float (the learned correction's input
// u = Linv r when the correction pass gathers FP32 u beside FP64 vectors)
template <typename VT, int BD>
__global__ void __launch_bounds__(256)
k_cg_update(uint32_t nrows, int pc, const void *dinv, double *__restrict__ x, VT *__restrict__ r,
            void *__restrict__ z, int z32, const VT *__restrict__ p, const VT *__restrict__ q,
            CGState *S, const double *pq_part, int np, double *part) {
  __shared__ double s4[4];
  if (S->done) return;
  const double pq = wg_sum(pq_part, np, 1, s4);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!finite_pos(pq)) {  // the same verdict in every workgroup
    if (lead) {
      S->pq = pq;
      S->status = SMH_ERR_BREAKDOWN;
      S->done = 1;
    }
    return;
  }
  const double alpha = S->rz / pq;
  if (lead) {
    S->pq = pq;
    S->alpha = alpha;
  }
  double drz = 0.0, drr = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nrows;
       i += (uint64_t)gridDim.x * 256) {
    double rv[BD], zv[BD];
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = i * BD + a;
      x[o] = x[o] + alpha * (double)p[o];
      const VT rn = (VT)((double)r[o] - alpha * (double)q[o]);
      r[o] = rn;
      rv[a] = (double)rn;
    }
    apply_precond<BD>(pc, rv, zv, dinv, sizeof(VT) == 4, i);
    for (int a = 0; a < BD; ++a) {
      const VT zz = z32 ? (VT)(float)zv[a] : (VT)zv[a];
      if (z32) ((float *)z)[i * BD + a] = (float)zv[a];
      else ((VT *)z)[i * BD + a] = zz;
      drz += (double)zz * rv[a];
      drr += rv[a] * rv[a];
    }
  }
  const double t1 = block_sum256(drz, s4);
  const double t2 = block_sum256(drr, s4);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t1;
    part[2 * blockIdx.x + 1] = t2;
  }
}

// K2b (learned preconditioner): z_i = Linv_i^T (u_i + sum_k C_k u_j) over
// the pattern of A (C zero on the diagonal; FP8 for 3x3 blocks, FP32 for
// scalar systems), and the workgroup partials of z'r and r'r.  check_done = 0
// on the residual-replacement path, which runs after the stopping test
// already set done.
template <typename VT, int BD, typename UT = VT>
struct EpiLearned {
  const UT *u;
  const VT *r;
  VT *z;
  const float *linv;
  double drz, drr;
  template <typename AT>  // the row loop's accumulator type
  __device__ void operator()(uint32_t row, const AT *s) {
    if constexpr (BD == 1) {
      const double t = (double)linv[row] * ((double)u[row] + (double)s[0]);
      const VT zz = (VT)t;
      z[row] = zz;
      const double rv = (double)r[row];
      drz += (double)zz * rv;
      drr += rv * rv;
      return;
    }
    double y[3];
    for (int a = 0; a < 3; ++a) y[a] = (double)u[(uint64_t)row * 3 + a] + (double)s[a];
    const float *L = linv + (uint64_t)row * 6;  // packed (tri_idx)
    for (int a = 0; a < 3; ++a) {
      double t = 0.0;
      for (int b = a; b < 3; ++b) t += (double)L[tri_idx(b, a)] * y[b];
      const uint64_t o = (uint64_t)row * 3 + a;
      const VT zz = (VT)t;
      z[o] = zz;
      const double rv = (double)r[o];
      drz += (double)zz * rv;
      drr += rv * rv;
    }
  }
};

template <int LP, int BD>
__global__ void __launch_bounds__(TileShape<BD>::kThreads)
k_cg_lapply(TileDev C, const typename Val<LP>::T *__restrict__ u,
            const typename Val<LP>::T *__restrict__ r, typename Val<LP>::T *__restrict__ z,
            const float *__restrict__ linv, const CGState *S, int check_done, double *part) {
  using VT = typename Val<LP>::T;
  __shared__ TileSmem<LP, BD> sm;
  if (check_done && S->done) return;
  uint32_t k0, k1;
  wg_tiles(C, k0, k1);
  XPlain<VT> xg{u, BD};
  EpiLearned<VT, BD> epi{u, r, z, linv, 0.0, 0.0};
  spmv_tiles<LP, BD>(C, k0, k1, sm, xg, epi);
  const double t1 = block_sum<TileShape<BD>::kWaves>(epi.drz, sm.red);
  const double t2 = block_sum<TileShape<BD>::kWaves>(epi.drr, sm.red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t1;
    part[2 * blockIdx.x + 1] = t2;
  }
}

// stopping rule of the reference loops: r'r > thr (and z'r > thr for PCG)
__device__ __forceinline__ bool stop_test(const CGState *S, double rr, double rz) {
  return !(rr > S->thr) || (S->precond_stop && !(rz > S->thr));
}

// R2: z'r, r'r, beta and the stopping test
__global__ void __launch_bounds__(256) k_cg_r2(CGState *S, const double *part, int np) {
  __shared__ double s4[4];
  if (S->done) return;
  const double rz = wg_sum(part, np, 2, s4);
  const double rr = wg_sum(part + 1, np, 2, s4);
  if (threadIdx.x == 0) {
    S->iter += 1;
    S->rr = rr;
    if (!(rr < INFINITY) || !(rz == rz)) {
      S->status = SMH_ERR_BREAKDOWN;
      S->done = 1;
    } else if (stop_test(S, rr, rz)) {
      S->rz = rz;
      S->claimed = 1;  // verified against the FP64 residual by the host
      S->done = 1;
    } else if (S->precond && !(rz > 0.0)) {
      // an indefinite preconditioner (z'r <= 0): the recurrence has no
      // meaning from here on
      S->status = SMH_ERR_BREAKDOWN;
      S->done = 1;
    } else if (S->iter >= S->maxiters) {
      S->done = 1;
    } else if ((S->mixed && rr < S->delta2 * S->rr_ref) ||
               (S->replace_every > 0 && S->iter % S->replace_every == 0)) {
      // pause before beta and p: the host replaces r (and z) by the true
      // residual, then k_cg_rres forms beta from the replaced z'r
      S->want_ru = 1;
      S->done = 1;
    } else {
      S->beta = rz / S->rz;
      S->rz = rz;
    }
    if (rr > S->rr_ref) S->rr_ref = rr;
  }
}

// K3: p = z + beta p
template <typename VT>
__global__ void __launch_bounds__(256)
k_cg_pupd(uint64_t n, VT *__restrict__ p, const VT *__restrict__ z, const CGState *S) {
  if (S->done) return;
  const double beta = S->beta;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    p[i] = (VT)((double)z[i] + beta * (double)p[i]);
}

// Residual replacement: r = b - A x in FP64 (A in precision RP), stored as
// VT, z = M r, and (init) p = z; workgroup partials of z'r and ||r||^2.
template <typename VT, int BD>
struct EpiResidual {
  const double *b;
  VT *r;
  void *z;  // VT, or float when z32 (as k_cg_update)
  VT *p;
  const void *dinv;
  int pc, init, z32;
  double drz, drr;
  __device__ void operator()(uint32_t row, const double *s) {
    double rv[BD], zv[BD];
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = (uint64_t)row * BD + a;
      const double v = b[o] - s[a];
      drr += v * v;
      r[o] = (VT)v;
      rv[a] = (double)(VT)v;
    }
    apply_precond<BD>(pc, rv, zv, dinv, sizeof(VT) == 4, row);
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = (uint64_t)row * BD + a;
      const VT zz = z32 ? (VT)(float)zv[a] : (VT)zv[a];
      if (z32) ((float *)z)[o] = (float)zv[a];
      else ((VT *)z)[o] = zz;
      if (init) p[o] = zz;
      drz += (double)zz * rv[a];
    }
  }
};

template <int RP, int BD, typename VT>
__global__ void __launch_bounds__(TileShape<BD>::kThreads)
k_residual(TileDev A, const double *__restrict__ b, const double *__restrict__ x,
           EpiResidual<VT, BD> epi, double *part) {
  __shared__ TileSmem<RP, BD> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XPlain<double> xg{x, BD};
  epi.drz = 0.0;
  epi.drr = 0.0;
  spmv_tiles<RP, BD>(A, k0, k1, sm, xg, epi);
  const double t1 = block_sum<TileShape<BD>::kWaves>(epi.drz, sm.red);
  const double t2 = block_sum<TileShape<BD>::kWaves>(epi.drr, sm.red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t1;
    part[2 * blockIdx.x + 1] = t2;
  }
}

// after a replacement: z'r, the true ||r||^2, the stopping test; the search
// direction is kept (reliable update) or set to z (start of the solve)
__global__ void __launch_bounds__(256) k_cg_rres(CGState *S, const double *part, int np,
                                                 int certify_only) {
  __shared__ double s4[4];
  const double rz = wg_sum(part, np, 2, s4);
  const double rr = wg_sum(part + 1, np, 2, s4);
  if (threadIdx.x) return;
  S->rr_true = rr;
  if (certify_only) return;
  // The replacement sits between the x/r update and beta (as in the
  // reference loop's periodic r = b - A x), so the search direction
  // continues as p = z + beta p with beta = (z'r)_true / (z'r)_old.  After a
  // false convergence claim, or when the true residual is more than 10x the
  // iterated one, restart from p = z instead: the old direction belongs to a
  // different residual.
  if (S->claimed || !(rr <= 100.0 * S->rr)) {
    S->restart = 1;
  } else if (S->want_ru) {
    S->restart = 2;
    S->beta = rz / S->rz;
  } else {
    S->restart = 0;
  }
  S->rz = rz;
  S->rr = rr;
  S->rr_ref = rr;
  S->want_ru = 0;
  S->claimed = 0;
  if (!(rr < INFINITY) || !(rz == rz)) {
    S->status = SMH_ERR_BREAKDOWN;
    S->done = 1;
  } else if (stop_test(S, rr, rz)) {
    S->done = 1;
  } else if (S->precond && !(rz > 0.0)) {
    S->status = SMH_ERR_BREAKDOWN;
    S->done = 1;
  } else if (S->status == SMH_OK) {
    S->done = S->iter >= S->maxiters ? 1 : 0;
  }
}

// after a residual replacement: p = z (restart 1) or p = z + beta p (restart 2)
template <typename VT>
__global__ void __launch_bounds__(256)
k_cg_restart(uint64_t n, VT *__restrict__ p, const VT *__restrict__ z, const CGState *S) {
  const int mode = S->restart;
  if (!mode) return;
  const double beta = S->beta;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    p[i] = mode == 1 ? z[i] : (VT)((double)z[i] + beta * (double)p[i]);
}

// Jacobi / block-Jacobi inverse of the diagonal (thermal.sim:229-233 and
// inv3 of runtime.cpp:135-160), from the matrix in its best precision.
template <int RP, typename OT>
__global__ void k_dinv(uint32_t nrows, const uint32_t *__restrict__ rowptr,
                       const uint32_t *__restrict__ colidx, const void *vals, const void *vals_lo,
                       int bd, int pc, OT *out) {
  using S = typename Val<RP>::S;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int bb = bd * bd;
  double d[9];
  for (int q = 0; q < bb; ++q) d[q] = 0.0;
  for (uint32_t ij = rowptr[i]; ij < rowptr[i + 1]; ++ij)
    if (colidx[ij] == i) {
      for (int q = 0; q < bb; ++q) {
        const uint64_t o = (uint64_t)ij * bb + q;
        d[q] = (double)((const S *)vals)[o];
        if (Val<RP>::kLo) d[q] += (double)((const float *)vals_lo)[o];
      }
      break;
    }
  if (bd == 1 || pc == SMH_PRECOND_JACOBI) {
    for (int a = 0; a < bd; ++a) out[(uint64_t)i * bd + a] = (OT)(1.0 / d[a * bd + a]);
    return;
  }
  const double *a = d;
  double inv[9];
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = -a[3] * a[8] + a[5] * a[6],
               c02 = a[3] * a[7] - a[4] * a[6];
  const double c10 = -a[1] * a[8] + a[2] * a[7], c11 = a[0] * a[8] - a[2] * a[6],
               c12 = -a[0] * a[7] + a[1] * a[6];
  const double c20 = a[1] * a[5] - a[2] * a[4], c21 = -a[0] * a[5] + a[2] * a[3],
               c22 = a[0] * a[4] - a[1] * a[3];
  const double det = 1.0 / (a[0] * c00 + a[1] * c01 + a[2] * c02);
  inv[0] = c00 * det; inv[1] = c10 * det; inv[2] = c20 * det;
  inv[3] = c01 * det; inv[4] = c11 * det; inv[5] = c21 * det;
  inv[6] = c02 * det; inv[7] = c12 * det; inv[8] = c22 * det;
  for (int q = 0; q < 9; ++q) out[(uint64_t)i * 9 + q] = (OT)inv[q];
}

__global__ void k_cg_init_state(CGState *S, double tol, int rel, int norm_form, int maxiters,
                                int precond_stop, double delta, int mixed, int replace_every,
                                int precond, const double *bbp) {
  if (threadIdx.x || blockIdx.x) return;
  const double bb = *bbp;
  memset(S, 0, sizeof(CGState));
  S->bb = bb;
  double thr = rel ? tol * tol * bb : tol;
  if (norm_form && !rel) thr = tol * tol;
  S->thr = thr;
  S->maxiters = maxiters;
  S->precond_stop = precond_stop;
  S->delta2 = delta * delta;
  S->mixed = mixed;
  S->replace_every = replace_every;
  S->precond = precond;
}

__global__ void __launch_bounds__(256) k_dot_wg(uint64_t n, const double *__restrict__ x,
                                                double *part) {
  __shared__ double s4[4];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    acc += x[i] * x[i];
  const double t = block_sum256(acc, s4);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) k_sum_to(const double *part, int np, double *out) {
  __shared__ double s4[4];
  const double t = wg_sum(part, np, 1, s4);
  if (threadIdx.x == 0) *out = t;
}

// the CG's per-workgroup partial sums: two per workgroup of a tile kernel
constexpr int kPartSlots = 2 * kMaxTileGrid;

// ---- host driver ---------------------------------------------------------------
// Bump allocation from a caller-owned device buffer (the scratch array of the
// allocation-free solver); 256-byte aligned pieces.
struct Arena {
  unsigned char *base = nullptr;
  size_t cap = 0, used = 0;
  void *take(size_t bytes) {
    const size_t off = (used + 255) & ~size_t(255);
    if (bytes > cap || off > cap - bytes) return nullptr;
    used = off + bytes;
    return base + off;
  }
};

struct Work {
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr, *u = nullptr,
       *dlo = nullptr, *linv = nullptr;
  CGState *S = nullptr;
  double *part = nullptr;
  void release() {
    dfree(r); dfree(z); dfree(p); dfree(q); dfree(dinv); dfree(u); dfree(dlo); dfree(linv);
    dfree(S); dfree(part);
  }
};


// ---- scalar 27-point stencil storage ----------------------------------------------------
// Row i couples with the columns i + off[s], s = 0..26 in ascending order
// (s = 13 is the diagonal); values are [27][n] (slot-major, so consecutive
// rows read consecutive addresses) and zero where there is no coupling.
struct StDev {
  uint32_t n;
  int32_t off[27];
  const void *vals;
  const void *vals_lo;
  uint32_t n0, n1, n2;  // the grid of rows (periodic neighbours)
};

// row j of slot s of row i on a periodic n0 x n1 x n2 grid (offsets of slot s:
// s % 3 - 1, s / 3 % 3 - 1, s / 9 - 1 along the three axes)
__device__ __forceinline__ uint32_t periodic_nbr(const StDev &A, uint32_t i, int s) {
  auto wrap = [](uint32_t c, int d, uint32_t n) -> uint32_t {
    return d < 0 ? (c == 0 ? n - 1 : c - 1) : d > 0 ? (c + 1 == n ? 0u : c + 1) : c;
  };
  const uint32_t c0 = i % A.n0, c1 = (i / A.n0) % A.n1, c2 = i / (A.n0 * A.n1);
  return wrap(c0, s % 3 - 1, A.n0) +
         A.n0 * (wrap(c1, s / 3 % 3 - 1, A.n1) + A.n1 * wrap(c2, s / 9 - 1, A.n2));
}

// One thread per row: the 27 products in ascending column order, each
// rounded before it is added -- the order and rounding of the tile loop, so
// the result is the blocked-CSR product bit for bit (zero slots add exact
// zeros; out-of-range columns read a clamped x times a zero value).  The
// periodic storage runs this loop; the non-periodic one runs stencil_rows_n
// below, which groups adjacent rows with the same per-row arithmetic.
//
// A plain grid-stride loop over the rows with plain loads.  Measured on the
// 192^3 hex system with one row per thread and slower than it: giving each
// XCD a contiguous eighth of the rows (226 us against 212 us), staging the
// nine x pencils of a 256-row block in LDS (9 global loads per row instead
// of 27 gathers; no faster, the solver iteration slower) and non-temporal
// value loads (239 us).
template <int PREC, class XG, class EPI, bool PER = false>
__device__ __forceinline__ void stencil_rows(const StDev &A, const XG &xg, EPI &epi) {
#pragma clang fp contract(off)
  using T = typename Val<PREC>::T;
  using S = typename Val<PREC>::S;
  const uint64_t n = A.n;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    // all 27 values, then all 27 gathers, then the products: the loads go
    // out together (one slot at a time leaves two loads in flight)
    S vs[27];
    float vl[Val<PREC>::kLo ? 27 : 1];
    T xs[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      vs[s] = ((const S *)A.vals)[(uint64_t)s * n + i];
      if constexpr (Val<PREC>::kLo) vl[s] = ((const float *)A.vals_lo)[(uint64_t)s * n + i];
    }
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      if constexpr (PER) {
        xs[s] = xg(periodic_nbr(A, (uint32_t)i, s), 0);
      } else {
        const int64_t j = (int64_t)i + A.off[s];
        const uint32_t jc = j < 0 ? 0u : (j >= (int64_t)n ? (uint32_t)(n - 1) : (uint32_t)j);
        xs[s] = xg(jc, 0);
      }
    }
    T acc = 0;
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      T v = val_of<PREC>(vs[s]);
      if constexpr (Val<PREC>::kLo) v += (T)vl[s];
      const T prod = v * xs[s];
      acc = acc + prod;
    }
    epi((uint32_t)i, &acc);
  }
}

// SMH_ST_ROWS adjacent rows per thread (non-periodic storage): per slot the
// rows' values are one load, and per pencil (d1, d2) the R + 2 entries
// x[i + o - 1 .. i + o + R] the rows gather (d0 = -1, 0, +1) are one load,
// so a row costs 27 / R value loads and 9 (R + 2) / R x loads instead of 27
// and 27.  Each row's 27 products are added in slot order with the one-row
// loop's rounding, so the results are the same bits (windows that leave
// [0, n) load entry by entry, clamped; a ragged last group runs row by row).
// 192^3 hex system, FP32 values and FP64 vectors, 4,096 workgroups: one row
// 211 us, two rows 179 us (the solver's SpMV 231 -> 203 us), four rows
// 171 us but the solver's SpMV 210 us (its epilogue pushes the kernel to 236
// VGPRs); each XCD walking a contiguous eighth of the rows was slower again
// (189 us).  At 8,192 workgroups: two rows 174 us (the solver's 197 us), four
// rows 163 us (the solver's 206 us).  So the plain product (smh_stencil27_mv)
// takes four rows and the solver's kernels two.
#ifndef SMH_ST_ROWS
#define SMH_ST_ROWS 2
#endif
#ifndef SMH_ST_ROWS_MV
#define SMH_ST_ROWS_MV 4
#endif
template <typename S, int R>
__device__ __forceinline__ void load_run(const S *p, S *o) {
  typedef S vR __attribute__((ext_vector_type(R), aligned(sizeof(S))));
  const vR r = *(const vR *)p;
#pragma unroll
  for (int k = 0; k < R; ++k) o[k] = r[k];
}
template <typename T, int W>
__device__ __forceinline__ void load_window(const T *x, int64_t j0, uint64_t n, T *w) {
  static_assert(W % 2 == 0, "windows load in pairs");
  if (j0 >= 0 && (uint64_t)j0 + W <= n) {
#pragma unroll
    for (int k = 0; k < W; k += 2) load_run<T, 2>(x + j0 + k, w + k);
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int64_t j = j0 + k;
      w[k] = x[j < 0 ? 0 : (j >= (int64_t)n ? (int64_t)n - 1 : j)];
    }
  }
}
// AT: the row sums' type (the CG inner SpMV sums FP32 products in FP64); R:
// rows per thread
template <int PREC, class EPI, typename AT = typename Val<PREC>::T, int R = SMH_ST_ROWS>
__device__ __forceinline__ void stencil_rows_n(const StDev &A, const typename Val<PREC>::T *x,
                                               EPI &epi) {
#pragma clang fp contract(off)
  using T = typename Val<PREC>::T;
  using S = typename Val<PREC>::S;
  const uint64_t n = A.n, ng = (n + R - 1) / R;
  for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < ng;
       gi += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = R * gi;
    if (i + R > n) {  // the ragged last group: row by row, as stencil_rows
      for (uint64_t r = i; r < n; ++r) {
        AT acc = 0;
#pragma unroll
        for (int s = 0; s < 27; ++s) {
          T v = val_of<PREC>(((const S *)A.vals)[(uint64_t)s * n + r]);
          if constexpr (Val<PREC>::kLo) v += (T)((const float *)A.vals_lo)[(uint64_t)s * n + r];
          const int64_t j = (int64_t)r + A.off[s];
          const uint64_t jc = j < 0 ? 0 : (j >= (int64_t)n ? n - 1 : (uint64_t)j);
          const AT prod = (AT)v * (AT)x[jc];
          acc = acc + prod;
        }
        epi((uint32_t)r, &acc);
      }
      continue;
    }
    S v[27][R];
    float l[Val<PREC>::kLo ? 27 : 1][R];
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      load_run<S, R>((const S *)A.vals + (uint64_t)s * n + i, v[s]);
      if constexpr (Val<PREC>::kLo) load_run<float, R>((const float *)A.vals_lo + (uint64_t)s * n + i, l[s]);
    }
    T w[9][R + 2];
#pragma unroll
    for (int pc = 0; pc < 9; ++pc)
      load_window<T, R + 2>(x, (int64_t)i + A.off[3 * pc + 1] - 1, n, w[pc]);
    AT acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0;
#pragma unroll
    for (int pc = 0; pc < 9; ++pc)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          T vv = val_of<PREC>(v[3 * pc + k][r]);
          if constexpr (Val<PREC>::kLo) vv += (T)l[3 * pc + k][r];
          const AT prod = (AT)vv * (AT)w[pc][k + r];
          acc[r] = acc[r] + prod;
        }
#pragma unroll
    for (int r = 0; r < R; ++r) epi((uint32_t)(i + r), &acc[r]);
  }
}

template <int PREC, bool PER = false>
__global__ void __launch_bounds__(256) k_stmv(StDev A, const typename Val<PREC>::T *__restrict__ x,
                                             typename Val<PREC>::T *__restrict__ y) {
  using T = typename Val<PREC>::T;
  XPlain<T> xg{x, 1};
  EpiStore<T, 1> epi{y};
  if constexpr (PER)
    stencil_rows<PREC, XPlain<T>, EpiStore<T, 1>, PER>(A, xg, epi);
  else
    stencil_rows_n<PREC, EpiStore<T, 1>, T, SMH_ST_ROWS_MV>(A, x, epi);
}

template <int IP>
__global__ void __launch_bounds__(256)
k_cg_spmv_st(StDev A, const typename Val<IP>::T *__restrict__ p,
             typename Val<IP>::T *__restrict__ q, const CGState *S, const float *dlo,
             double *part) {
  using VT = typename Val<IP>::T;
  __shared__ double red[4];
  if (S->done) return;
  EpiPq<VT> epi{q, p, dlo, 0.0, 1};
  stencil_rows_n<IP, EpiPq<VT>, typename Val<cg_acc_prec<IP>()>::T>(A, p, epi);
  const double t = block_sum256(epi.d, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <int RP, typename VT>
__global__ void __launch_bounds__(256)
k_residual_st(StDev A, const double *__restrict__ b, const double *__restrict__ x,
              EpiResidual<VT, 1> epi, double *part) {
  __shared__ double red[4];
  epi.drz = 0.0;
  epi.drr = 0.0;
  stencil_rows_n<RP>(A, x, epi);
  const double t1 = block_sum256(epi.drz, red);
  const double t2 = block_sum256(epi.drr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t1;
    part[2 * blockIdx.x + 1] = t2;
  }
}

// u in the correction's vector type (Val<LP>::T: float for kF16F, also under
// FP64 r and z, which keeps the 27 gathered values in 27 registers)
template <int LP, typename VT>
__global__ void __launch_bounds__(256)
k_cg_lapply_st(StDev C, const typename Val<LP>::T *__restrict__ u, const VT *__restrict__ r,
               VT *__restrict__ z, const float *__restrict__ linv, const CGState *S,
               int check_done, double *part) {
  using UT = typename Val<LP>::T;
  __shared__ double red[4];
  if (check_done && S->done) return;
  EpiLearned<VT, 1, UT> epi{u, r, z, linv, 0.0, 0.0};
  stencil_rows_n<LP>(C, u, epi);
  const double t1 = block_sum256(epi.drz, red);
  const double t2 = block_sum256(epi.drr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t1;
    part[2 * blockIdx.x + 1] = t2;
  }
}

// dlo_i = sum of the 27 low parts of row i (ascending, FP64: the CSR sum)
__global__ void k_lo_rowsum_st(uint32_t n, const float *__restrict__ lo,
                               float *__restrict__ dlo) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int q = 0; q < 27; ++q) s += (double)lo[(uint64_t)q * n + i];
  dlo[i] = (float)s;
}

// Jacobi: 1 / the diagonal slot, in the matrix's best precision
template <int RP, typename OT>
__global__ void k_dinv_st(uint32_t n, const void *vals, const void *vals_lo, OT *out) {
  using S = typename Val<RP>::S;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = (double)((const S *)vals)[13ull * n + i];
  if (Val<RP>::kLo) d += (double)((const float *)vals_lo)[13ull * n + i];
  out[i] = (OT)(1.0 / d);
}

// CSR -> stencil through an LDS table: a workgroup takes 64 consecutive
// rows, drops each of their entries into a zeroed [27][64] slot table (row
// by binary search over the rows' offsets, slot by the first offset equal
// to column - row), then writes the table out slot by slot, so both the
// reads of the entries and the writes of each slot are coalesced.  An entry
// outside its row's 27 offsets raises *err.
template <typename OT, typename VT>
__device__ __forceinline__ OT st_out(VT v) {
  if constexpr (sizeof(OT) == 2)  // IEEE half, round to nearest even
    return __builtin_bit_cast(uint16_t, (_Float16)v);
  else
    return (OT)v;
}
template <typename VT, typename OT>
__global__ void __launch_bounds__(256)
k_st_convert(StDev D, const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ colidx,
             const VT *__restrict__ vals, const float *__restrict__ vals_lo, OT *__restrict__ out,
             float *__restrict__ out_lo, int *__restrict__ err) {
  constexpr int kRows = 64;
  __shared__ VT tv[27 * kRows];
  __shared__ float tl[27 * kRows];
  __shared__ uint32_t rp[kRows + 1];
  const uint64_t r0 = (uint64_t)blockIdx.x * kRows;
  if (r0 >= D.n) return;
  const int nr = (int)(D.n - r0 < (uint64_t)kRows ? D.n - r0 : (uint64_t)kRows);
  const bool lo = out_lo != nullptr;
  for (int t = threadIdx.x; t < 27 * kRows; t += 256) {
    tv[t] = (VT)0;
    tl[t] = 0.0f;
  }
  for (int t = threadIdx.x; t <= nr; t += 256) rp[t] = rowptr[r0 + t];
  __syncthreads();
  const uint32_t base = rp[0], cnt = rp[nr] - base;
  for (uint32_t k = threadIdx.x; k < cnt; k += 256) {
    const uint32_t g = base + k;
    int a = 0, b = nr - 1;  // the last row starting at or before entry g
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (rp[mid] <= g) a = mid;
      else b = mid - 1;
    }
    const int64_t o = (int64_t)colidx[g] - (int64_t)(r0 + a);
    int sl = -1;
    for (int q = 0; q < 27; ++q)
      if (D.off[q] == o) {
        sl = q;
        break;
      }
    if (sl < 0) {
      atomicMax(err, 1);
      continue;
    }
    tv[sl * kRows + a] = vals[g];
    if (lo) tl[sl * kRows + a] = vals_lo[g];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 27 * kRows; t += 256) {
    const int sl = t / kRows, r = t - sl * kRows;
    if (r >= nr) continue;
    const uint64_t dst = (uint64_t)sl * D.n + r0 + r;
    out[dst] = st_out<OT>(tv[t]);
    if (lo) out_lo[dst] = tl[t];
  }
}

// The matrix passes of one CG solve over a 27-point stencil matrix.
template <int IP, int RP, int LP>
struct StencilOps {
  using VT = typename Val<IP>::T;
  StDev A{}, C{};
  StDev R{};  // the residual's (and the Jacobi inverse's) values, precision RP
  unsigned g = 0;
  unsigned parts_spmv() const { return g; }
  unsigned parts_res() const { return g; }
  unsigned parts_corr() const { return g; }
  void spmv(hipStream_t s, const VT *p, VT *q, const CGState *S, const float *dlo,
            double *part) const {
    hipLaunchKernelGGL((k_cg_spmv_st<IP>), dim3(g), dim3(256), 0, s, A, p, q, S, dlo, part);
  }
  void residual(hipStream_t s, const double *b, const double *x, const EpiResidual<VT, 1> &epi,
                double *part) const {
    hipLaunchKernelGGL((k_residual_st<RP, VT>), dim3(g), dim3(256), 0, s, R, b, x, epi, part);
  }
  // the correction pass reads u in the correction's vector type
  static constexpr bool kU32 = sizeof(typename Val<LP>::T) == 4 && sizeof(VT) == 8;
  void corr(hipStream_t s, const void *u, const VT *r, VT *z, const float *linv, const CGState *S,
            int check_done, double *part) const {
    hipLaunchKernelGGL((k_cg_lapply_st<LP, VT>), dim3(g), dim3(256), 0, s, C,
                       (const typename Val<LP>::T *)u, r, z, linv, S, check_done, part);
  }
  void lo_rowsum(hipStream_t s, float *dlo) const {
    hipLaunchKernelGGL(k_lo_rowsum_st, dim3(grid_for(A.n, 256)), dim3(256), 0, s, A.n,
                       (const float *)A.vals_lo, dlo);
  }
  void dinv(hipStream_t s, int, VT *out) const {
    hipLaunchKernelGGL((k_dinv_st<RP, VT>), dim3(grid_for(R.n, 256)), dim3(256), 0, s, R.n,
                       R.vals, R.vals_lo, out);
  }
};

// the device view of a stencil matrix (offsets of the lexicographic
// (d2, d1, d0) order, i.e. ascending)
StDev st_dev(const smh_stencil27 *S, const void *vals, const void *vals_lo) {
  StDev d{};
  d.n = S->n0 * S->n1 * S->n2;
  int q = 0;
  for (int d2 = -1; d2 <= 1; ++d2)
    for (int d1 = -1; d1 <= 1; ++d1)
      for (int d0 = -1; d0 <= 1; ++d0)
        d.off[q++] = d0 + (int32_t)S->n0 * d1 + (int32_t)(S->n0 * S->n1) * d2;
  d.vals = vals;
  d.vals_lo = vals_lo;
  d.n0 = S->n0;
  d.n1 = S->n1;
  d.n2 = S->n2;
  return d;
}

// The matrix passes of one CG solve over blocked-CSR row tiles: the inner
// SpMV (precision IP), the true residual (RP), the learned correction pass
// (LP) and the two per-row set-ups (row sums of the low parts, the Jacobi
// diagonal).  cg_run drives the same state machine over any such policy.
template <int IP, int RP, int LP, int BD>
struct TileOps {
  using VT = typename Val<IP>::T;
  // the inner SpMV of FP32 3x3 values may run over a lane-planar copy
  static constexpr bool kPlanar = BD == 3 && (IP == SMH_F32 || IP == kF32D);
  const smh_bsr *A = nullptr;
  TileDev T{}, TR{}, TC{};
  TileDev TP{};  // lane-planar tiles (TP.vals set: the inner SpMV uses them)
  unsigned gt = 0, grt = 0, gc = 0, gp = 0;
  bool planar() const { return kPlanar && TP.vals != nullptr; }
  unsigned parts_spmv() const { return planar() ? gp : gt; }
  unsigned parts_res() const { return grt; }
  unsigned parts_corr() const { return gc; }
  void spmv(hipStream_t s, const VT *p, VT *q, const CGState *S, const float *dlo,
            double *part) const {
    if constexpr (kPlanar) {
      if (planar()) {
        hipLaunchKernelGGL((k_cg_spmv_pl<IP>), dim3(gp), dim3(TileShapePlS::kThreads), 0, s, TP,
                           p, q, S, dlo, part);
        return;
      }
    }
    hipLaunchKernelGGL((k_cg_spmv<IP, BD>), dim3(gt), dim3(tile_threads_host(BD)), 0, s, T, p,
                       q, S, dlo, part);
  }
  void residual(hipStream_t s, const double *b, const double *x, const EpiResidual<VT, BD> &epi,
                double *part) const {
    hipLaunchKernelGGL((k_residual<RP, BD, VT>), dim3(grt), dim3(tile_threads_host(BD)), 0, s, TR,
                       b, x, epi, part);
  }
  static constexpr bool kU32 = false;
  void corr(hipStream_t s, const void *u, const VT *r, VT *z, const float *linv, const CGState *S,
            int check_done, double *part) const {
    hipLaunchKernelGGL((k_cg_lapply<LP, BD>), dim3(gc), dim3(tile_threads_host(BD)), 0, s, TC,
                       (const VT *)u, r, z, linv, S, check_done, part);
  }
  void lo_rowsum(hipStream_t s, float *dlo) const {
    hipLaunchKernelGGL(k_lo_rowsum, dim3(grid_for((uint64_t)A->nrows * BD * BD, 256)), dim3(256),
                       0, s, A->nrows, BD * BD, A->rowptr, (const float *)A->vals_lo, dlo);
  }
  void dinv(hipStream_t s, int pc, VT *out) const {
    hipLaunchKernelGGL((k_dinv<RP, VT>), dim3(grid_for(A->nrows, 256)), dim3(256), 0, s,
                       A->nrows, A->rowptr, A->colidx, TR.vals, TR.vals_lo, BD, pc, out);
  }
};

// the tile policy of A: the correction pass runs its precision's tile shape
// (one block per lane for FP8 3x3), so it may need a plan of its own
template <int IP, int RP, int BD>
int make_tile_ops(smh_ctx *ctx, const smh_bsr *A, const smh_tile_plan *plan,
                  const smh_tile_plan *rplan, const smh_cg_opts *o, void *out) {
  using VT = typename Val<IP>::T;
  constexpr int LP = BD == 3 ? (sizeof(VT) == 4 ? kFP8F : kFP8D)
                             : (sizeof(VT) == 4 ? SMH_F32 : kF32D);
  constexpr bool kIl = BD == 1 && SMH_TILE_INTERLEAVE1;
  TileOps<IP, RP, LP, BD> &ops = *(TileOps<IP, RP, LP, BD> *)out;
  ops.A = A;
  ops.T = tile_dev(A, plan, &ops.gt, kIl);
  // residuals (and the Jacobi inverse) over the caller's FP64 values when given
  smh_bsr Ar = *A;
  if (RP == SMH_F64 && o->vals_f64) {
    Ar.vals = o->vals_f64;
    Ar.vals_lo = nullptr;
  }
  ops.TR = tile_dev(&Ar, rplan, &ops.grt, kIl);
  smh_bsr Cb = *A;
  const bool learned = o->precond == SMH_PRECOND_LEARNED;
  Cb.vals = learned ? (const void *)o->learned_corr : A->vals;
  Cb.vals_lo = nullptr;
  smh_tile_plan lplan = *plan;
  if (learned && tile_blocks_of<LP, BD>() != plan->tile_blocks) {
    const int prc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, tile_blocks_of<LP, BD>(), &lplan);
    if (prc) return prc;
    SMH_CHECK(ctx, lplan.tile_row != nullptr, SMH_ERR_UNSUPPORTED,
              "smh_cg_solve: a row holds %u blocks (the learned correction supports %u)",
              lplan.maxrow, tile_blocks_of<LP, BD>() / 2);
  }
  // the 3x3 correction pass strides over its XCD's tiles too (96^3 x 5:
  // 59 -> 57 us per pass)
  ops.TC = tile_dev(&Cb, &lplan, &ops.gc, kIl || BD == 3, SMH_TILE_GRID_C);
  return SMH_OK;
}

template <int IP, int RP, int BD, class OPS>
int cg_run(smh_ctx *ctx, const OPS &ops, uint32_t nrows, bool has_lo, const double *b, double *x,
           const smh_cg_opts *o, smh_cg_stats *st, Arena *ar = nullptr) {
  using VT = typename Val<IP>::T;
  const uint64_t n = (uint64_t)nrows * BD;
  const unsigned nbe = grid_for(nrows, 256, SMH_CG_UPDATE_GRID);
  const unsigned nbv = grid_for(n, 256, sizeof(VT) == 8 ? SMH_CG_VEC_GRID64 : SMH_CG_VEC_GRID);
  const int pc = o->precond;
  const bool learned = pc == SMH_PRECOND_LEARNED;
  const unsigned gt = ops.parts_spmv(), grt = ops.parts_res(), gc = ops.parts_corr();
  const int check = o->check_every > 0 ? o->check_every : 16;
  const void *pdata = nullptr;  // the preconditioner data K2 reads
  const float *linv = nullptr;   // the learned preconditioner's Linv blocks
  Work w;
  int rc = SMH_OK;
  auto T0 = std::chrono::steady_clock::now();
  CGState hs{};
  int nru = 0, batch = 0;
  const int z32 = learned && OPS::kU32;  // the learned correction's u stored as float
  EpiResidual<VT, BD> epi{};
  // the solve's buffers: device allocations, or pieces of the arena
  auto walloc = [&](void **pp, size_t bytes) -> int {
    if (!ar) return dmalloc(ctx, pp, bytes);
    *pp = ar->take(bytes);
    return *pp ? SMH_OK
               : fail(ctx, SMH_ERR_UNSUPPORTED, "CG: the %zu-byte scratch array is too small",
                      ar->cap);
  };
#define TRY(x)            \
  do {                    \
    rc = (x);             \
    if (rc) goto done;    \
  } while (0)
#define LCHECK()                                                                  \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                       \
      rc = fail(ctx, SMH_ERR_HIP, "CG launch failed: %s", hipGetErrorString(e_)); \
      goto done;                                                                  \
    }                                                                             \
  } while (0)
  TRY(walloc(&w.r, sizeof(VT) * n));
  TRY(walloc(&w.z, sizeof(VT) * n));
  TRY(walloc(&w.p, sizeof(VT) * n));
  TRY(walloc(&w.q, sizeof(VT) * n));
  if (pc != SMH_PRECOND_NONE && !learned)
    TRY(walloc(&w.dinv, sizeof(VT) * (uint64_t)nrows * BD * BD));
  if (learned) {
    TRY(walloc(&w.u, sizeof(VT) * n));
    w.dinv = nullptr;
    if (BD == 3) {
      // Linv's six lower-triangular entries per row: what the update and the
      // correction pass read (24 instead of 36 bytes a row)
      TRY(walloc(&w.linv, sizeof(float) * 6 * (uint64_t)nrows));
      hipLaunchKernelGGL(k_pack_linv, dim3(grid_for((uint64_t)nrows * 6, 256)), dim3(256), 0,
                         ctx->stream, nrows, (const float *)o->learned, (float *)w.linv);
    }
  }
  linv = learned ? (BD == 3 ? (const float *)w.linv : (const float *)o->learned) : nullptr;
  // mixed precision: the inner operator keeps the exact block row sums
  if (o->mixed && has_lo) {
    TRY(walloc(&w.dlo, sizeof(float) * (uint64_t)nrows * BD * BD));
    ops.lo_rowsum(ctx->stream, (float *)w.dlo);
  }
  pdata = learned ? (const void *)linv : (const void *)w.dinv;
  TRY(walloc((void **)&w.S, sizeof(CGState)));
  TRY(walloc((void **)&w.part, sizeof(double) * 2 * kPartSlots + 64));
  // b'b, state, preconditioner
  hipLaunchKernelGGL(k_dot_wg, dim3(nbv), dim3(256), 0, ctx->stream, n, b, w.part);
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, ctx->stream, w.part, (int)nbv,
                     w.part + 2 * kPartSlots);
  hipLaunchKernelGGL(k_cg_init_state, dim3(1), dim3(1), 0, ctx->stream, w.S, o->tol, o->rel,
                     o->norm_form, o->maxiters, (pc != SMH_PRECOND_NONE && o->zr_stop) ? 1 : 0,
                     o->reliable_delta, o->mixed, o->replace_every,
                     pc != SMH_PRECOND_NONE ? 1 : 0, (const double *)(w.part + 2 * kPartSlots));
  if (pc != SMH_PRECOND_NONE && !learned) ops.dinv(ctx->stream, pc, (VT *)w.dinv);
  epi.b = b;
  epi.r = (VT *)w.r;
  epi.z = learned ? w.u : w.z;  // learned: u = Linv r, z follows in k_cg_lapply
  epi.z32 = z32;
  epi.p = (VT *)w.p;
  epi.dinv = pdata;
  epi.pc = pc;
  // r = b - A x, z = M r, p = z
  epi.init = learned ? 0 : 1;
  ops.residual(ctx->stream, b, (const double *)x, epi, w.part);
  if (learned) {
    ops.corr(ctx->stream, w.u, (const VT *)w.r, (VT *)w.z, linv,
             (const CGState *)w.S, 0, w.part);
    hipLaunchKernelGGL(k_cg_rres, dim3(1), dim3(256), 0, ctx->stream, w.S,
                       (const double *)w.part, (int)gc, 0);
    hipLaunchKernelGGL((k_cg_restart<VT>), dim3(nbv), dim3(256), 0, ctx->stream, n, (VT *)w.p,
                       (const VT *)w.z, (const CGState *)w.S);
  } else {
    hipLaunchKernelGGL(k_cg_rres, dim3(1), dim3(256), 0, ctx->stream, w.S,
                       (const double *)w.part, (int)grt, 0);
  }
  LCHECK();
  epi.init = 0;
  // Iterations are queued in batches between host checks of the state.  A
  // replacement pauses the system until the host runs it, and the paused
  // rest of a batch still costs its launches, so the batch shrinks after a
  // replacement and doubles back towards check_every after a clean batch.
  batch = check;
  for (;;) {
    SMH_HIP(ctx, hipMemcpyAsync(&hs, w.S, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
    SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hs.status == SMH_OK && (hs.want_ru || hs.claimed)) {
      ++nru;
      batch = std::max(2, batch / 4);
      ops.residual(ctx->stream, b, (const double *)x, epi, w.part);
      if (learned)
        ops.corr(ctx->stream, w.u, (const VT *)w.r, (VT *)w.z, linv,
                 (const CGState *)w.S, 0, w.part);
      hipLaunchKernelGGL(k_cg_rres, dim3(1), dim3(256), 0, ctx->stream, w.S,
                         (const double *)w.part, (int)(learned ? gc : grt), 0);
      hipLaunchKernelGGL((k_cg_restart<VT>), dim3(nbv), dim3(256), 0, ctx->stream, n, (VT *)w.p,
                         (const VT *)w.z, (const CGState *)w.S);
      LCHECK();
      continue;
    }
    if (hs.done || hs.status != SMH_OK) break;
    for (int k = 0; k < batch; ++k) {
      ops.spmv(ctx->stream, (const VT *)w.p, (VT *)w.q, (const CGState *)w.S,
               (const float *)w.dlo, w.part);
      hipLaunchKernelGGL((k_cg_update<VT, BD>), dim3(nbe), dim3(256), 0, ctx->stream, nrows, pc,
                         pdata, x, (VT *)w.r, learned ? w.u : w.z, z32, (const VT *)w.p,
                         (const VT *)w.q, w.S, (const double *)w.part, (int)gt,
                         w.part + kPartSlots);
      if (learned)
        ops.corr(ctx->stream, w.u, (const VT *)w.r, (VT *)w.z, linv,
                 (const CGState *)w.S, 1, w.part + kPartSlots);
      hipLaunchKernelGGL(k_cg_r2, dim3(1), dim3(256), 0, ctx->stream, w.S,
                         (const double *)(w.part + kPartSlots), (int)(learned ? gc : nbe));
      hipLaunchKernelGGL((k_cg_pupd<VT>), dim3(nbv), dim3(256), 0, ctx->stream, n, (VT *)w.p,
                         (const VT *)w.z, (const CGState *)w.S);
    }
    LCHECK();
    batch = std::min(check, 2 * batch);
  }
  // certify: FP64 residual of the final x (r, z are scratch from here on)
  ops.residual(ctx->stream, b, (const double *)x, epi, w.part);
  hipLaunchKernelGGL(k_cg_rres, dim3(1), dim3(256), 0, ctx->stream, w.S, (const double *)w.part,
                     (int)grt, 1);
  LCHECK();
  SMH_HIP(ctx, hipMemcpyAsync(&hs, w.S, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (st) {
    st->iters = hs.iter;
    st->reliable_updates = nru;
    st->rr = hs.rr;
    st->true_rel = hs.bb > 0 ? std::sqrt(hs.rr_true / hs.bb) : std::sqrt(hs.rr_true);
    int status = hs.status;
    if (status == SMH_OK && !(hs.rr_true <= hs.thr * (1.0 + 1e-12)) &&
        !(hs.precond_stop && hs.rz <= hs.thr))
      status = SMH_ERR_NOT_CONVERGED;
    st->status = status;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - T0).count();
  }
done:
#undef TRY
#undef LCHECK
  hipStreamSynchronize(ctx->stream);
  if (!ar) w.release();
  return rc;
}

// ---- batched CG: many independent systems in one block-diagonal matrix ---------
// System m owns rows [rows[m], rows[m+1]).  Every system runs its own CG
// recurrence (own alpha, beta and stopping test) in lockstep: the SpMV and the
// vector updates stream the whole union once per iteration, and the scalar
// reductions are segmented by system in two fixed-order levels: each system's
// rows are cut into chunks of kBChunk rows (a chunk never spans two systems),
// one workgroup sums a chunk, one thread per system sums its chunks.  One
// iteration is nine launches, all on the device and captured once into a
// hipGraph that the host replays check_every times between state reads:
//   K1  q = A_in p, per-row p'q (row tiles)   S1  chunk sums of p'q
//   R1  alpha per system
//   K2  x, r, z update (a workgroup per chunk) with chunk sums of z'r, r'r
//   R2  stopping test, then beta or a reliable-update request
//   RU  r = b - A x in FP64 for the tiles of systems that claimed convergence
//       or asked for a reliable update (other workgroups exit at once)
//   S2  chunk sums of the replaced z'r, r'r     RR  certify claims, beta
//   K3  p = z + beta p, or p = z after a restart (a workgroup per chunk)
// The replacement thus sits between the x/r update and beta, as in the
// reference loop's periodic r = b - A x.
enum BState : int {
  kBActive = 0,
  kBClaimed = 1,   // r'r <= thr in the iteration: certified by RR
  kBDoneOk = 2,
  kBDoneMaxit = 3,
  kBBreakdown = 4,
  kBWant = 5       // asked for a residual replacement: served by RU/RR
};

__device__ __forceinline__ bool bflagged(int st) { return st == kBClaimed || st == kBWant; }

struct BDev {
  double *rz, *alpha, *beta, *rr, *rr_true, *thr, *bb, *rr_ref;
  int *state, *iter, *restart;
};

__global__ void k_bcg_mesh_rows(uint32_t nb, const uint32_t *__restrict__ rows,
                                uint32_t *__restrict__ mesh) {
  const uint32_t m = blockIdx.x;
  if (m >= nb) return;
  for (uint32_t i = rows[m] + threadIdx.x; i < rows[m + 1]; i += blockDim.x) mesh[i] = m;
}

// every block of a row of system m must stay inside system m
__global__ void k_bcg_check(uint32_t n, const uint32_t *__restrict__ rowptr,
                            const uint32_t *__restrict__ colidx, const uint32_t *__restrict__ mesh,
                            const uint32_t *__restrict__ rows, int *__restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = mesh[i], lo = rows[m], hi = rows[m + 1];
  for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (colidx[k] < lo || colidx[k] >= hi) {
      atomicOr(err, 1);
      return;
    }
}

// K1: q = A_in p; prow = p'q per row
template <typename VT>
struct EpiPqRow {
  VT *q;
  const VT *p;
  const float *dlo;
  double *prow;
  int bd;
  template <typename AT>  // the row sums' type, as EpiPq
  __device__ void operator()(uint32_t r, const AT *s) {
    AT sv[3];
    for (int a = 0; a < bd; ++a) sv[a] = s[a];
    if (dlo) {
      const float *D = dlo + (uint64_t)r * bd * bd;
      for (int a = 0; a < bd; ++a)
        for (int b = 0; b < bd; ++b) sv[a] += (AT)D[bd * a + b] * (AT)p[(uint64_t)r * bd + b];
    }
    double d = 0.0;
    for (int a = 0; a < bd; ++a) {
      const uint64_t o = (uint64_t)r * bd + a;
      const VT qv = (VT)sv[a];
      q[o] = qv;
      d += (double)qv * (double)p[o];
    }
    prow[r] = d;
  }
};

template <int IP, int BD, class SH = TileShape<BD>>
__global__ void __launch_bounds__(SH::kThreads)
k_bcg_spmv(TileDev A, const typename Val<IP>::T *__restrict__ p, typename Val<IP>::T *__restrict__ q,
           const float *dlo, double *prow) {
  using VT = typename Val<IP>::T;
  constexpr int AP = cg_acc_prec<IP>();
  __shared__ TileSmem<AP, BD, SH> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  if (k0 < k1 && A.row_sys) {
    // a workgroup whose rows all belong to stopped systems has nothing to do
    const uint32_t R0 = A.tile_row[k0], R1 = A.tile_row[k1];
    bool any = false;
    if (R0 < R1)
      for (uint32_t m = A.row_sys[R0]; m <= A.row_sys[R1 - 1] && !any; ++m)
        any = A.sys_state[m] == 0;
    if (!any) return;  // the same answer on every thread
  }
  XWide<VT, typename Val<AP>::T> xg{p, BD};
  EpiPqRow<VT> epi{q, p, dlo, prow, BD};
  spmv_tiles<AP, BD>(A, k0, k1, sm, xg, epi);
}

// K1 over lane-planar FP32 3x3 values (the batched mixed solver's planar
// shape)
template <int IP>
__global__ void __launch_bounds__(TileShapePlS::kThreads)
k_bcg_spmv_pl(TileDev A, const typename Val<IP>::T *__restrict__ p,
              typename Val<IP>::T *__restrict__ q, const float *dlo, double *prow) {
  using VT = typename Val<IP>::T;
  constexpr int AP = cg_acc_prec<IP>();
  __shared__ TileSmem<AP, 3, TileShapePlS> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  XWide<VT, typename Val<AP>::T> xg{p, 3};
  EpiPqRow<VT> epi{q, p, dlo, prow, 3};
  spmv_tiles<AP, 3>(A, k0, k1, sm, xg, epi);
}

// Chunks: crow[c]..crow[c+1] are the rows of chunk c, all in system cmesh[c];
// system m owns chunks mchunk[m]..mchunk[m+1].
constexpr uint32_t kBChunk = 2048;

struct BChunks {
  const uint32_t *crow, *cmesh, *mchunk;
  uint32_t nchunks, nb;
};

// S1/S2: per-chunk sums of one or two per-row vectors (flagged_only: only the
// chunks of systems being replaced)
__global__ void __launch_bounds__(256) k_bcg_csum(BChunks C, const double *__restrict__ v1,
                                                  const double *__restrict__ v2,
                                                  const int *__restrict__ state,
                                                  int flagged_only, double *cp1, double *cp2) {
  __shared__ double s4[4];
  const uint32_t c = blockIdx.x;
  if (flagged_only && !bflagged(state[C.cmesh[c]])) return;  // uniform
  double a = 0.0, b = 0.0;
  for (uint32_t i = C.crow[c] + threadIdx.x; i < C.crow[c + 1]; i += 256) {
    a += v1[i];
    if (v2) b += v2[i];
  }
  a = block_sum256(a, s4);
  if (v2) b = block_sum256(b, s4);
  if (threadIdx.x == 0) {
    cp1[c] = a;
    if (v2) cp2[c] = b;
  }
}

// a system's total over its chunks, in chunk order (one thread)
// (loads issued eight at a time, added in chunk order)
__device__ __forceinline__ double sys_sum(const BChunks &C, uint32_t m, const double *cp) {
  double t = 0.0;
  uint32_t c = C.mchunk[m];
  const uint32_t e = C.mchunk[m + 1];
  for (; c + 8 <= e; c += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = cp[c + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += v[u];
  }
  for (; c < e; ++c) t += cp[c];
  return t;
}

// R1: alpha per system (a thread per system)
__global__ void __launch_bounds__(256) k_bcg_alpha(BChunks C, const double *cp, BDev S) {
  const uint32_t m = blockIdx.x * 256 + threadIdx.x;
  if (m >= C.nb || S.state[m] != kBActive) return;
  const double pq = sys_sum(C, m, cp);
  if (finite_pos(pq)) {
    S.alpha[m] = S.rz[m] / pq;
  } else {
    S.alpha[m] = 0.0;
    S.state[m] = kBBreakdown;
  }
}

// K2: x += alpha p, r -= alpha q, z = M r; a workgroup per chunk (one system),
// with the chunk's z'r and r'r
template <typename VT, int BD>
__global__ void __launch_bounds__(256)
k_bcg_update(BChunks C, int pc, const void *dinv, double *__restrict__ x, VT *__restrict__ r,
             VT *__restrict__ z, const VT *__restrict__ p, const VT *__restrict__ q, BDev S,
             double *cpz, double *cpr) {
  __shared__ double s4[4];
  const uint32_t c = blockIdx.x, m = C.cmesh[c];
  if (S.state[m] != kBActive) return;  // uniform; R2 skips this system
  const double alpha = S.alpha[m];
  double drz = 0.0, drr = 0.0;
  for (uint64_t i = C.crow[c] + threadIdx.x; i < C.crow[c + 1]; i += 256) {
    double rv[BD], zv[BD];
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = i * BD + a;
      x[o] = x[o] + alpha * (double)p[o];
      const VT rn = (VT)((double)r[o] - alpha * (double)q[o]);
      r[o] = rn;
      rv[a] = (double)rn;
    }
    apply_precond<BD>(pc, rv, zv, dinv, sizeof(VT) == 4, i);
    for (int a = 0; a < BD; ++a) {
      const VT zz = (VT)zv[a];
      z[i * BD + a] = zz;
      drz += (double)zz * rv[a];
      drr += rv[a] * rv[a];
    }
  }
  drz = block_sum256(drz, s4);
  drr = block_sum256(drr, s4);
  if (threadIdx.x == 0) {
    cpz[c] = drz;
    cpr[c] = drr;
  }
}

// R2: the stopping test, then beta or a reliable-update request (a thread per
// system)
__global__ void __launch_bounds__(256) k_bcg_beta(BChunks C, const double *cpz,
                                                  const double *cpr, BDev S, int maxiters,
                                                  int mixed, double delta2, int replace_every,
                                                  int precond) {
  const uint32_t m = blockIdx.x * 256 + threadIdx.x;
  if (m >= C.nb || S.state[m] != kBActive) return;
  const double rz = sys_sum(C, m, cpz), rr = sys_sum(C, m, cpr);
  const int it = S.iter[m] + 1;
  S.iter[m] = it;
  S.rr[m] = rr;
  S.restart[m] = 0;
  if (!(rr < INFINITY) || !(rz == rz)) {
    S.state[m] = kBBreakdown;
  } else if (!(rr > S.thr[m])) {
    S.state[m] = kBClaimed;
  } else if (precond && !(rz > 0.0)) {
    S.state[m] = kBBreakdown;  // an indefinite preconditioner
  } else if (it >= maxiters) {
    S.state[m] = kBDoneMaxit;
  } else if ((mixed && rr < delta2 * S.rr_ref[m]) ||
             (replace_every > 0 && it % replace_every == 0)) {
    S.state[m] = kBWant;  // S.rz keeps the previous z'r for RR's beta
  } else {
    S.beta[m] = rz / S.rz[m];
    S.rz[m] = rz;
  }
  if (rr > S.rr_ref[m]) S.rr_ref[m] = rr;
}

// K3: p = z + beta p (restart 0) or p = z (restart 1); a workgroup per chunk
template <typename VT, int BD>
__global__ void __launch_bounds__(256)
k_bcg_pupd(BChunks C, VT *__restrict__ p, const VT *__restrict__ z, BDev S) {
  const uint32_t c = blockIdx.x, m = C.cmesh[c];
  if (S.state[m] != kBActive) return;  // uniform
  const bool restart = S.restart[m] != 0;
  const double beta = S.beta[m];
  for (uint64_t o = (uint64_t)C.crow[c] * BD + threadIdx.x; o < (uint64_t)C.crow[c + 1] * BD;
       o += 256)
    p[o] = restart ? z[o] : (VT)((double)z[o] + beta * (double)p[o]);
}

// RU epilogue: r = b - A x (FP64, A in precision RP), z = M r for the rows of
// the systems being replaced (all rows on the host-driven passes); per-row
// z'r and r'r of every row the workgroup computed
template <typename VT, int BD>
struct EpiResRow {
  const double *b;
  VT *r, *z;
  const void *dinv;
  int pc, all;
  const uint32_t *mesh;
  const int *state;
  double *prz, *prr;
  __device__ void operator()(uint32_t row, const double *s) {
    double rv[BD], zv[BD], drr = 0.0, drz = 0.0;
    for (int a = 0; a < BD; ++a) {
      const double v = b[(uint64_t)row * BD + a] - s[a];
      drr += v * v;
      rv[a] = (double)(VT)v;
    }
    apply_precond<BD>(pc, rv, zv, dinv, sizeof(VT) == 4, row);
    const bool mine = all || bflagged(state[mesh[row]]);
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = (uint64_t)row * BD + a;
      const VT zz = (VT)zv[a];
      if (mine) {
        r[o] = (VT)rv[a];
        z[o] = zz;
      }
      drz += (double)zz * rv[a];
    }
    prz[row] = drz;
    prr[row] = drr;
  }
};

template <int RP, int BD, typename VT>
__global__ void __launch_bounds__(TileShape<BD>::kThreads)
k_bcg_residual(TileDev A, const double *__restrict__ x, EpiResRow<VT, BD> epi) {
  __shared__ TileSmem<RP, BD> sm;
  uint32_t k0, k1;
  wg_tiles(A, k0, k1);
  if (!epi.all) {
    // skip the workgroup unless one of its rows belongs to a system being
    // replaced (the same answer on every thread)
    if (k0 >= k1) return;
    const uint32_t R0 = A.tile_row[k0], R1 = A.tile_row[k1];
    if (R0 >= R1) return;
    bool any = false;
    for (uint32_t m = epi.mesh[R0]; m <= epi.mesh[R1 - 1] && !any; ++m)
      any = bflagged(epi.state[m]);
    if (!any) return;
  }
  XPlain<double> xg{x, BD};
  spmv_tiles<RP, BD>(A, k0, k1, sm, xg, epi);
}

// RR: after RU and S2.  mode 0 (in the iteration): certify the claims, and
// beta from the replaced residual for the requests; mode 1: start of the
// solve; mode 2: the final certification of every system.  A thread per system.
__global__ void __launch_bounds__(256) k_bcg_rres(BChunks C, const double *cpz,
                                                  const double *cpr, BDev S, int mode,
                                                  int precond) {
  const uint32_t m = blockIdx.x * 256 + threadIdx.x;
  if (m >= C.nb) return;
  const int st = S.state[m];
  if (mode == 0 && !bflagged(st)) return;
  const double rz = sys_sum(C, m, cpz), rr = sys_sum(C, m, cpr);
  S.rr_true[m] = rr;
  if (mode == 2) return;
  if (mode == 1) {
    S.restart[m] = 1;
  } else if (st == kBClaimed) {
    if (rr <= S.thr[m] * (1.0 + 1e-12)) {
      S.state[m] = kBDoneOk;
      return;
    }
    S.restart[m] = 1;  // a false claim: restart from the true residual
  } else if (!(rr <= 100.0 * S.rr[m])) {
    S.restart[m] = 1;  // the old direction belongs to a far smaller residual
  } else {
    S.restart[m] = 0;
    S.beta[m] = rz / S.rz[m];
  }
  S.state[m] = kBActive;
  if (!(rr < INFINITY) || !(rz == rz)) {
    S.state[m] = kBBreakdown;
  } else if (!(rr > S.thr[m])) {
    S.state[m] = kBDoneOk;
  } else if (precond && !(rz > 0.0)) {
    S.state[m] = kBBreakdown;
  }
  S.rz[m] = rz;
  S.rr[m] = rr;
  S.rr_ref[m] = rr;
}

// per-chunk b'b
template <int BD>
__global__ void __launch_bounds__(256) k_bcg_bb(BChunks C, const double *__restrict__ b,
                                                double *cp) {
  __shared__ double s4[4];
  const uint32_t c = blockIdx.x;
  double t = 0.0;
  for (uint64_t o = (uint64_t)C.crow[c] * BD + threadIdx.x; o < (uint64_t)C.crow[c + 1] * BD;
       o += 256)
    t += b[o] * b[o];
  t = block_sum256(t, s4);
  if (threadIdx.x == 0) cp[c] = t;
}

// per-system b'b, thresholds and initial state
__global__ void __launch_bounds__(256) k_bcg_init(BChunks C, const double *cp, BDev S,
                                                  double tol, int rel) {
  const uint32_t m = blockIdx.x * 256 + threadIdx.x;
  if (m >= C.nb) return;
  const double bb = sys_sum(C, m, cp);
  S.bb[m] = bb;
  S.thr[m] = rel ? tol * tol * bb : tol;
  S.state[m] = kBActive;
  S.iter[m] = 0;
  S.restart[m] = 1;
  S.rz[m] = S.rr[m] = S.rr_ref[m] = S.rr_true[m] = 0.0;
  S.alpha[m] = S.beta[m] = 0.0;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// The last kernel of every replay of the captured iterations: bumps the
// replay counter and publishes (replay number, systems still iterating) to
// pinned host memory, so the host learns of convergence without a copy or a
// stream synchronisation (the count is written first, the number after a
// system-scope fence).
__global__ void __launch_bounds__(256) k_bcg_signal(const int *__restrict__ state, uint32_t nb,
                                                    unsigned long long *counter,
                                                    volatile unsigned long long *host) {
  __shared__ unsigned s4[4];
  unsigned n = 0;
  for (uint32_t m = threadIdx.x; m < nb; m += 256) n += state[m] == kBActive || bflagged(state[m]);
  for (int off = 32; off > 0; off >>= 1) n += (unsigned)__shfl_xor((int)n, off, 64);
  if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x) return;
  const unsigned long long c = *counter + 1;
  *counter = c;
  host[1] = s4[0] + s4[1] + s4[2] + s4[3];
  __threadfence_system();
  host[0] = c;
}

// The lane-planar inner SpMV's buffer and tile plan when the caller (the
// allocation-free solver) provides them
struct PlanarIn {
  const smh_tile_plan *plan = nullptr;
  float *vals = nullptr;
};

// CG over the blocked-CSR row tiles of A
template <int IP, int RP, int BD>
int cg_tiles(smh_ctx *ctx, const smh_bsr *A, const smh_tile_plan *plan, const smh_tile_plan *rplan,
             const double *b, double *x, const smh_cg_opts *o, smh_cg_stats *st,
             Arena *ar = nullptr, const PlanarIn *pin = nullptr) {
  using VT = typename Val<IP>::T;
  constexpr int LP = BD == 3 ? (sizeof(VT) == 4 ? kFP8F : kFP8D)
                             : (sizeof(VT) == 4 ? SMH_F32 : kF32D);
  TileOps<IP, RP, LP, BD> ops;
  int rc = make_tile_ops<IP, RP, BD>(ctx, A, plan, rplan, o, &ops);
  if (rc) return rc;
  if constexpr (decltype(ops)::kPlanar) {
    // the inner SpMV streams a lane-planar copy of the FP32 values, kept per
    // pattern in the context (refilled by every solve), or in the buffer the
    // allocation-free solver carves from its scratch array
    if (A->nnz > 0 && (!ar || (pin && pin->vals))) {
      float *pv = nullptr;
      smh_tile_plan pplan;
      if (ar) {
        pv = pin->vals;
        pplan = *pin->plan;
      } else {
        rc = planar_buffer(ctx, A, &pv);
        if (rc) return rc;
        rc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, (uint32_t)TileShapePlS::kBlocks,
                       &pplan);
        if (rc) return rc;
      }
      if (pplan.tile_row) {
        hipLaunchKernelGGL(k_to_planar, dim3(grid_for((A->nnz + 63) & ~63ull, 256)), dim3(256),
                           0, ctx->stream, A->nnz, (const float *)A->vals, pv);
        SMH_LAUNCH_CHECK(ctx);
        smh_bsr Pb = *A;
        Pb.vals = pv;
        Pb.vals_lo = nullptr;
        ops.TP = tile_dev(&Pb, &pplan, &ops.gp, true, SMH_TILE_GRID_PLS);
      }
    }
  }
  return cg_run<IP, RP, BD>(ctx, ops, A->nrows, A->precision == SMH_F32X2 && A->vals_lo, b, x, o,
                            st, ar);
}

// The batched SpMV's tile shape and grid: for 3x3 blocks, every (rounds per
// lane, threads per workgroup) shape below with 1,024, 2,048 and 4,096
// workgroups, and for FP32 inner products the lane-planar shape with 4,096,
// 8,192 and 12,288; the kernel itself (epilogue included) timed on the
// pattern (five launches after a warm-up); scalar systems keep the default
// shape and tune the grid only.  Chosen once per pattern and precision,
// cached in the context.
using BShape0 = TileShape<3>;
using BShape1 = TileShapeR<3, 8, 64>;
using BShape2 = TileShapeR<3, 24, 64>;
using BShape3 = TileShapeR<3, 8, 128>;
using BShape4 = TileShapeR<3, 16, 128>;
// shape 5: lane-planar FP32 values (TileShapePlS; FP32 inner products only),
// with workgroups striding over their XCD's tiles
constexpr int kBShapes = 5, kBShapePlanar = 5;
inline uint32_t bshape_blocks(int k) {
  switch (k) {
    case 1: return BShape1::kBlocks;
    case 2: return BShape2::kBlocks;
    case 3: return BShape3::kBlocks;
    case 4: return BShape4::kBlocks;
    case kBShapePlanar: return TileShapePlS::kBlocks;
  }
  return BShape0::kBlocks;
}
inline unsigned bshape_threads(int k) {
  switch (k) {
    case 1: return BShape1::kThreads;
    case 2: return BShape2::kThreads;
    case 3: return BShape3::kThreads;
    case 4: return BShape4::kThreads;
    case kBShapePlanar: return TileShapePlS::kThreads;
  }
  return BShape0::kThreads;
}
template <int IP, int BD>
void launch_bcg_spmv(int shape, unsigned g, hipStream_t s, const TileDev &T,
                     const typename Val<IP>::T *p, typename Val<IP>::T *q, const float *dlo,
                     double *prow) {
  if constexpr (BD == 3) {
    switch (shape) {
      case 1:
        hipLaunchKernelGGL((k_bcg_spmv<IP, 3, BShape1>), dim3(g), dim3(BShape1::kThreads), 0, s, T,
                           p, q, dlo, prow);
        return;
      case 2:
        hipLaunchKernelGGL((k_bcg_spmv<IP, 3, BShape2>), dim3(g), dim3(BShape2::kThreads), 0, s, T,
                           p, q, dlo, prow);
        return;
      case 3:
        hipLaunchKernelGGL((k_bcg_spmv<IP, 3, BShape3>), dim3(g), dim3(BShape3::kThreads), 0, s, T,
                           p, q, dlo, prow);
        return;
      case 4:
        hipLaunchKernelGGL((k_bcg_spmv<IP, 3, BShape4>), dim3(g), dim3(BShape4::kThreads), 0, s, T,
                           p, q, dlo, prow);
        return;
      case kBShapePlanar:
        if constexpr (IP == SMH_F32) {
          hipLaunchKernelGGL((k_bcg_spmv_pl<IP>), dim3(g), dim3(TileShapePlS::kThreads), 0, s, T,
                             p, q, dlo, prow);
          return;
        }
        break;
    }
  }
  hipLaunchKernelGGL((k_bcg_spmv<IP, BD>), dim3(g), dim3(tile_threads_host(BD)), 0, s, T, p, q,
                     dlo, prow);
}

struct BTune {
  int shape = 0;
  unsigned grid = 0;
  smh_tile_plan plan{};
};
template <int IP, int BD>
BTune tuned_bcg(smh_ctx *ctx, hipStream_t stream, const smh_bsr *A,
                const smh_tile_plan *default_plan, void *pbuf, void *qbuf, const float *dlo,
                double *prow, const float *planar) {
  using VT = typename Val<IP>::T;
  for (const auto &t : ctx->tuned)
    if (t.rowptr == A->rowptr && t.nrows == A->nrows && t.nnz == A->nnz && t.bdim == A->bdim &&
        t.prec == IP) {
      BTune r;
      r.shape = t.shape;
      r.grid = t.grid;
      if (tile_plan(ctx, A->rowptr, A->nrows, A->nnz,
                    BD == 3 ? bshape_blocks(t.shape) : tile_blocks_host(BD), &r.plan))
        r.plan = *default_plan;
      return r;
    }
  BTune best;
  best.plan = *default_plan;
  float best_ms = 1e30f;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return best;
  const unsigned cands[] = {1024, 2048, 4096};
  const unsigned cands_pl[] = {4096, 8192, 12288};
  const int nshapes = BD == 3 ? kBShapes + (planar ? 1 : 0) : 1;
  for (int k = 0; k < nshapes; ++k) {
    smh_tile_plan pl;
    if (tile_plan(ctx, A->rowptr, A->nrows, A->nnz,
                  BD == 3 ? bshape_blocks(k) : tile_blocks_host(BD), &pl) != SMH_OK ||
        !pl.tile_row)
      continue;
    const bool is_pl = k == kBShapePlanar;
    for (unsigned g : is_pl ? cands_pl : cands) {
      unsigned gg = 0;
      smh_bsr Ap = *A;
      if (is_pl) {
        Ap.vals = planar;
        Ap.vals_lo = nullptr;
      }
      TileDev T = tile_dev(&Ap, &pl, &gg, is_pl);
      gg = std::min<unsigned>(g, pl.ntiles);
      if (gg >= 8) gg &= ~7u;
      if (gg < 1) gg = 1;
      T.interleave = is_pl && gg >= 8 ? 1u : 0u;
      T.tiles_per_wg = (uint32_t)((pl.ntiles + gg - 1) / gg);
      float ms = 0.0f;
      for (int rep = 0; rep < 6; ++rep) {
        (void)hipEventRecord(e0, stream);
        launch_bcg_spmv<IP, BD>(k, gg, stream, T, (const VT *)pbuf, (VT *)qbuf, dlo, prow);
        (void)hipEventRecord(e1, stream);
        (void)hipEventSynchronize(e1);
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (rep) ms += t;  // the first launch warms up
      }
      if (ms < best_ms) {
        best_ms = ms;
        best.shape = k;
        best.grid = gg;
        best.plan = pl;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  ctx->tuned.push_back({A->rowptr, A->nrows, A->bdim, A->nnz, best.grid, best.shape, IP});
  return best;
}

// A batched solve built once and replayed: its device workspace, chunk
// tables and three captured graphs (prologue + set-up + the first chunk of
// iterations; a chunk of iterations; the final certification), cached in
// the context under everything the graphs hold by value.
struct BPlan : smh_cached {
  // key
  const uint32_t *colidx = nullptr;
  const void *vals = nullptr, *vals_lo = nullptr;
  uint32_t nrows = 0, bdim = 0;
  uint64_t nnz = 0;
  int prec = 0, ip = 0;
  const double *b = nullptr;
  double *x = nullptr;
  std::vector<uint32_t> rows;
  smh_cg_opts o{};
  std::string pro_key;
  // state
  std::vector<void *> bufs;
  hipGraph_t g[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t ge[3] = {nullptr, nullptr, nullptr};
  BDev S{};
  int check = 16;
  ~BPlan() override {
    for (int k = 0; k < 3; ++k) {
      if (ge[k]) (void)hipGraphExecDestroy(ge[k]);
      if (g[k]) (void)hipGraphDestroy(g[k]);
    }
    for (void *p : bufs) dfree(p);
  }
  bool same(const smh_bsr *A, const double *bb, const double *xx, uint32_t nb,
            const uint32_t *hrows, const smh_cg_opts *oo, const std::string &pk, int ipp) const {
    return rowptr == A->rowptr && colidx == A->colidx && vals == A->vals &&
           vals_lo == A->vals_lo && nrows == A->nrows && nnz == A->nnz && bdim == A->bdim &&
           prec == A->precision && ip == ipp && b == bb && x == xx && rows.size() == nb + 1u &&
           std::memcmp(rows.data(), hrows, sizeof(uint32_t) * (nb + 1)) == 0 &&
           std::memcmp(&o, oo, sizeof(smh_cg_opts)) == 0 && pro_key == pk;
  }
};

template <int IP, int RP, int BD>
int bcg_build(smh_ctx *ctx, BPlan &P, const smh_bsr *Am, const smh_tile_plan *plan, uint32_t nb,
              const uint32_t *hrows, const double *b, double *x, const smh_cg_opts *o,
              const std::function<int(hipStream_t)> &prologue) {
  using VT = typename Val<IP>::T;
  const uint32_t nrows = Am->nrows;
  const uint64_t n = (uint64_t)nrows * BD;
  const int pc = o->precond;
  const int check = o->check_every > 0 ? o->check_every : 16;
  P.check = check;
  hipStream_t s = ctx->own_stream;
  auto alloc = [&](void **pp, size_t bytes) {
    const int r = dmalloc(ctx, pp, bytes);
    if (!r) P.bufs.push_back(*pp);
    return r;
  };
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr, *dlo = nullptr;
  uint32_t *rows = nullptr, *mesh = nullptr, *ctab = nullptr;
  double *prz = nullptr, *prr = nullptr, *sbuf = nullptr, *cp1 = nullptr, *cp2 = nullptr;
  int *ibuf = nullptr, *err = nullptr;
  unsigned long long *counter = nullptr;
  int rc = SMH_OK;
  // chunk tables: [crow (nchunks+1) | cmesh (nchunks) | mchunk (nb+1)]
  std::vector<uint32_t> hcrow, hcmesh, hmchunk(nb + 1);
  for (uint32_t m = 0; m < nb; ++m) {
    hmchunk[m] = (uint32_t)hcrow.size();
    for (uint32_t i = hrows[m]; i < hrows[m + 1]; i += std::min(kBChunk, hrows[m + 1] - i)) {
      hcrow.push_back(i);
      hcmesh.push_back(m);
    }
  }
  const uint32_t nch = (uint32_t)hcrow.size();
  hmchunk[nb] = nch;
  hcrow.push_back(nrows);
  std::vector<uint32_t> htab(hcrow);
  htab.insert(htab.end(), hcmesh.begin(), hcmesh.end());
  htab.insert(htab.end(), hmchunk.begin(), hmchunk.end());
  if ((rc = alloc(&r, sizeof(VT) * n)) || (rc = alloc(&z, sizeof(VT) * n)) ||
      (rc = alloc(&p, sizeof(VT) * n)) || (rc = alloc(&q, sizeof(VT) * n)) ||
      (rc = alloc((void **)&rows, sizeof(uint32_t) * (nb + 1))) ||
      (rc = alloc((void **)&mesh, sizeof(uint32_t) * (nrows ? nrows : 1))) ||
      (rc = alloc((void **)&prz, sizeof(double) * (nrows ? nrows : 1))) ||
      (rc = alloc((void **)&prr, sizeof(double) * (nrows ? nrows : 1))) ||
      (rc = alloc((void **)&sbuf, sizeof(double) * 8 * nb)) ||
      (rc = alloc((void **)&ibuf, sizeof(int) * 3 * nb)) || (rc = alloc((void **)&err, 64)) ||
      (rc = alloc((void **)&counter, 64)) ||
      (rc = alloc((void **)&ctab, sizeof(uint32_t) * htab.size())) ||
      (rc = alloc((void **)&cp1, sizeof(double) * (nch ? nch : 1))) ||
      (rc = alloc((void **)&cp2, sizeof(double) * (nch ? nch : 1))))
    return rc;
  const BChunks CH{ctab, ctab + nch + 1, ctab + 2 * nch + 1, nch, nb};
  if (pc != SMH_PRECOND_NONE && (rc = alloc(&dinv, sizeof(VT) * (uint64_t)nrows * BD * BD)))
    return rc;
  if (o->mixed && Am->precision == SMH_F32X2 && Am->vals_lo &&
      (rc = alloc(&dlo, sizeof(float) * (uint64_t)nrows * BD * BD)))
    return rc;
  // FP32 3x3 inner products may stream a lane-planar copy of the values
  // (made from the prologue's values by the set-up of every replayed solve)
  void *planar = nullptr;
  if (BD == 3 && IP == SMH_F32 && Am->nnz > 0 &&
      (rc = alloc(&planar, sizeof(float) * planar_floats(Am->nnz))))
    return rc;
  auto to_planar = [&]() {
    if (planar)
      hipLaunchKernelGGL(k_to_planar, dim3(grid_for((Am->nnz + 63) & ~63ull, 256)), dim3(256), 0,
                         s, Am->nnz, (const float *)Am->vals, (float *)planar);
  };
  const BDev S{sbuf, sbuf + nb, sbuf + 2 * nb, sbuf + 3 * nb, sbuf + 4 * nb, sbuf + 5 * nb,
               sbuf + 6 * nb, sbuf + 7 * nb, ibuf, ibuf + nb, ibuf + 2 * nb};
  P.S = S;
  SMH_HIP(ctx, hipMemcpyAsync(rows, hrows, sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice, s));
  SMH_HIP(ctx, hipMemcpyAsync(ctab, htab.data(), sizeof(uint32_t) * htab.size(),
                              hipMemcpyHostToDevice, s));
  SMH_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_bcg_mesh_rows, dim3(nb), dim3(256), 0, s, nb, (const uint32_t *)rows, mesh);
  hipLaunchKernelGGL(k_bcg_check, dim3(grid_for(nrows, 256)), dim3(256), 0, s, nrows, Am->rowptr,
                     Am->colidx, (const uint32_t *)mesh, (const uint32_t *)rows, err);
  int herr = 0;
  SMH_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipStreamSynchronize(s));
  SMH_CHECK(ctx, !herr, SMH_ERR_INVALID,
            "smh_cg_solve_batched: a block couples two systems (the matrix is not "
            "block-diagonal by the given row ranges)");
  // the autotuned SpMV (tile shape and grid), then the other kernels' shapes
  bool cached = false;
  for (const auto &t : ctx->tuned)
    cached |= t.rowptr == Am->rowptr && t.nrows == Am->nrows && t.nnz == Am->nnz &&
              t.bdim == Am->bdim && t.prec == IP;
  if (dlo && !cached)  // the tuning runs the kernel with its epilogue: finite row sums
    SMH_HIP(ctx, hipMemsetAsync(dlo, 0, sizeof(float) * (uint64_t)nrows * BD * BD, s));
  to_planar();
  const BTune tune =
      tuned_bcg<IP, BD>(ctx, s, Am, plan, p, q, (const float *)dlo, prz, (const float *)planar);
  unsigned g = 0, gsp = tune.grid;
  const TileDev T = tile_dev(Am, plan, &g);  // the residual kernel's tiles (default shape)
  unsigned gdef = 0;
  const bool use_pl = tune.shape == kBShapePlanar && planar;
  smh_bsr Asp = *Am;
  if (use_pl) {
    Asp.vals = planar;
    Asp.vals_lo = nullptr;
  }
  TileDev Ts = tile_dev(&Asp, &tune.plan, &gdef, use_pl);
  if (gsp) Ts.tiles_per_wg = (uint32_t)((tune.plan.ntiles + gsp - 1) / gsp);
  else gsp = gdef;
  Ts.interleave = use_pl && gsp >= 8 ? 1u : 0u;
  // the SpMV streams the tiles of systems still iterating only
  Ts.row_sys = mesh;
  Ts.sys_state = S.state;
  const unsigned gs = grid_for(nb, 256);  // a thread per system
  EpiResRow<VT, BD> epi{b, (VT *)r, (VT *)z, dinv, pc, 0, (const uint32_t *)mesh, S.state,
                        prz, prr};
  EpiResRow<VT, BD> epi_all = epi;
  epi_all.all = 1;
  // RU + S2 + RR (all: every system, else only those being replaced)
  auto replace = [&](int all, int mode) {
    hipLaunchKernelGGL((k_bcg_residual<RP, BD, VT>), dim3(g), dim3(tile_threads_host(BD)), 0, s,
                       T, (const double *)x, all ? epi_all : epi);
    hipLaunchKernelGGL(k_bcg_csum, dim3(nch), dim3(256), 0, s, CH, (const double *)prz,
                       (const double *)prr, (const int *)S.state, all ? 0 : 1, cp1, cp2);
    hipLaunchKernelGGL(k_bcg_rres, dim3(gs), dim3(256), 0, s, CH, (const double *)cp1,
                       (const double *)cp2, S, mode, pc != SMH_PRECOND_NONE ? 1 : 0);
  };
  auto pupd = [&]() {
    hipLaunchKernelGGL((k_bcg_pupd<VT, BD>), dim3(nch), dim3(256), 0, s, CH, (VT *)p,
                       (const VT *)z, S);
  };
  // the set-up, after the caller's prologue (e.g. the assembly):
  // preconditioner, low-part row sums, b'b, the states, then r = b - A x,
  // z = M r, p = z for every system; the replay counter restarts at zero
  auto setup = [&]() {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, counter, 8, 0ull);
    if (use_pl) to_planar();
    if (pc != SMH_PRECOND_NONE)
      hipLaunchKernelGGL((k_dinv<RP, VT>), dim3(grid_for(nrows, 256)), dim3(256), 0, s, nrows,
                         Am->rowptr, Am->colidx, Am->vals, Am->vals_lo, BD, pc, (VT *)dinv);
    if (dlo)
      hipLaunchKernelGGL(k_lo_rowsum, dim3(grid_for((uint64_t)nrows * BD * BD, 256)), dim3(256),
                         0, s, nrows, BD * BD, Am->rowptr, (const float *)Am->vals_lo,
                         (float *)dlo);
    hipLaunchKernelGGL(k_bcg_bb<BD>, dim3(nch), dim3(256), 0, s, CH, b, cp1);
    hipLaunchKernelGGL(k_bcg_init, dim3(gs), dim3(256), 0, s, CH, (const double *)cp1, S, o->tol,
                       o->rel);
    replace(1, 1);
    pupd();
  };
  auto signal = [&]() {
    hipLaunchKernelGGL(k_bcg_signal, dim3(1), dim3(256), 0, s, (const int *)S.state, nb,
                       counter, ctx->signal_dev);
  };
  // one chunk of `check` iterations
  auto iterate = [&]() {
    for (int k = 0; k < check; ++k) {
      launch_bcg_spmv<IP, BD>(tune.shape, gsp, s, Ts, (const VT *)p, (VT *)q, (const float *)dlo,
                              prz);
      hipLaunchKernelGGL(k_bcg_csum, dim3(nch), dim3(256), 0, s, CH, (const double *)prz,
                         (const double *)nullptr, (const int *)S.state, 0, cp1, cp2);
      hipLaunchKernelGGL(k_bcg_alpha, dim3(gs), dim3(256), 0, s, CH, (const double *)cp1, S);
      hipLaunchKernelGGL((k_bcg_update<VT, BD>), dim3(nch), dim3(256), 0, s, CH, pc,
                         (const void *)dinv, x, (VT *)r, (VT *)z, (const VT *)p, (const VT *)q, S,
                         cp1, cp2);
      hipLaunchKernelGGL(k_bcg_beta, dim3(gs), dim3(256), 0, s, CH, (const double *)cp1,
                         (const double *)cp2, S, o->maxiters, o->mixed,
                         o->reliable_delta * o->reliable_delta, o->replace_every,
                         pc != SMH_PRECOND_NONE ? 1 : 0);
      replace(0, 0);
      pupd();
    }
  };
  SMH_HIP(ctx, hipGetLastError());
  for (int k = 0; k < 3; ++k) {
    SMH_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int prc = SMH_OK;
    if (k == 0) {
      if (prologue) prc = prologue(s);
      setup();
    }
    if (k < 2) {
      iterate();
      signal();
    } else {
      replace(1, 2);  // certify every system against its FP64 residual
    }
    const hipError_t ce = hipStreamEndCapture(s, &P.g[k]);
    if (prc) return prc;
    if (ce != hipSuccess || hipGraphInstantiate(&P.ge[k], P.g[k], nullptr, nullptr, 0) !=
                                hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, SMH_ERR_HIP, "smh_cg_solve_batched: graph instantiation failed");
    }
  }
  return SMH_OK;
}

// Replays go out two ahead of the host, which watches the signal in pinned
// memory (no copies, no stream synchronisation) and stops once every system
// has left the iteration (by maxiters at the latest); replays queued past
// that exit at once.
int bcg_replay(smh_ctx *ctx, BPlan &P, uint32_t nb, const smh_cg_opts *o,
               smh_cg_batch_stats *st) {
  hipStream_t s = ctx->own_stream;
  const int check = P.check;
  volatile unsigned long long *sig = ctx->signal_host;
  const uint64_t max_replays = (uint64_t)(o->maxiters + check - 1) / check + 2;
  sig[1] = ~0ull;  // this solve's replays count from zero
  sig[0] = 0;
  uint64_t launched = 0;
  auto launch = [&](hipGraphExec_t ge) {
    if (hipGraphLaunch(ge, s) != hipSuccess) return false;
    ++launched;
    return true;
  };
  bool ok = launch(P.ge[0]) && (max_replays < 2 || launch(P.ge[1]));
  while (ok) {
    // wait until all but the newest replay have finished
    uint64_t spins = 0;
    while (sig[0] + 1 < launched) {
      if (++spins % 4096) continue;
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess && sig[0] + 1 < launched)
        return fail(ctx, SMH_ERR_HIP, "smh_cg_solve_batched: lost the replay signal");
      if (q != hipSuccess && q != hipErrorNotReady)
        return fail(ctx, SMH_ERR_HIP, "smh_cg_solve_batched: replay failed: %s",
                    hipGetErrorString(q));
    }
    if (sig[1] == 0 || launched >= max_replays) break;
    ok = launch(P.ge[1]);
  }
  SMH_CHECK(ctx, ok, SMH_ERR_HIP, "smh_cg_solve_batched: graph launch failed");
  SMH_CHECK(ctx, hipGraphLaunch(P.ge[2], s) == hipSuccess, SMH_ERR_HIP,
            "smh_cg_solve_batched: graph launch failed");
  std::vector<int> hstate(nb), hiter(nb);
  std::vector<double> hrr(nb), hbb(nb), hthr(nb);
  SMH_HIP(ctx, hipMemcpyAsync(hstate.data(), P.S.state, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipMemcpyAsync(hiter.data(), P.S.iter, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipMemcpyAsync(hrr.data(), P.S.rr_true, sizeof(double) * nb, hipMemcpyDeviceToHost,
                              s));
  SMH_HIP(ctx, hipMemcpyAsync(hbb.data(), P.S.bb, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipMemcpyAsync(hthr.data(), P.S.thr, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipStreamSynchronize(s));
  if (st)
    for (uint32_t m = 0; m < nb; ++m) {
      st[m].iters = hiter[m];
      st[m].true_rel = hbb[m] > 0 ? std::sqrt(hrr[m] / hbb[m]) : std::sqrt(hrr[m]);
      int status = SMH_OK;
      if (hstate[m] == kBBreakdown) status = SMH_ERR_BREAKDOWN;
      else if (!(hrr[m] <= hthr[m] * (1.0 + 1e-12))) status = SMH_ERR_NOT_CONVERGED;
      st[m].status = status;
    }
  return SMH_OK;
}

template <int IP, int RP, int BD>
int bcg_run(smh_ctx *ctx, const smh_bsr *Am, const smh_tile_plan *plan, uint32_t nb,
            const uint32_t *hrows, const double *b, double *x, const smh_cg_opts *o,
            smh_cg_batch_stats *st, double *seconds,
            const std::function<int(hipStream_t)> &prologue, const std::string &pro_key) {
  auto T0 = std::chrono::steady_clock::now();
  // the graph runs on the context's own stream (capture is not allowed on the
  // legacy default stream), ordered after the caller's work
  (void)hipStreamSynchronize(ctx->stream);
  BPlan *P = nullptr;
  for (auto &c : ctx->cached) {
    BPlan *bp = dynamic_cast<BPlan *>(c.get());
    if (bp && bp->same(Am, b, x, nb, hrows, o, pro_key, IP)) P = bp;
  }
  int rc = SMH_OK;
  if (!P) {
    auto np = std::make_unique<BPlan>();
    np->rowptr = Am->rowptr;
    np->colidx = Am->colidx;
    np->vals = Am->vals;
    np->vals_lo = Am->vals_lo;
    np->nrows = Am->nrows;
    np->nnz = Am->nnz;
    np->bdim = Am->bdim;
    np->prec = Am->precision;
    np->ip = IP;
    np->b = b;
    np->x = x;
    np->rows.assign(hrows, hrows + nb + 1);
    np->o = *o;
    np->pro_key = pro_key;
    rc = bcg_build<IP, RP, BD>(ctx, *np, Am, plan, nb, hrows, b, x, o, prologue);
    (void)hipStreamSynchronize(ctx->own_stream);
    if (rc == SMH_OK) {
      P = np.get();
      ctx->cached.push_back(std::move(np));
    }
  }
  if (P) rc = bcg_replay(ctx, *P, nb, o, st);
  (void)hipStreamSynchronize(ctx->own_stream);
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - T0).count();
  return rc;
}

}  // namespace

extern "C" int smh_bsr_forget(smh_ctx *ctx, const uint32_t *rowptr) {
  if (!ctx) return SMH_ERR_INVALID;
  forget_plans(ctx, rowptr);
  return SMH_OK;
}

extern "C" int smh_bsr_planar(smh_ctx *ctx, const smh_bsr *A, float *d_out) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && d_out && A->rowptr && (A->nnz == 0 || A->vals), SMH_ERR_INVALID,
            "smh_bsr_planar: null argument");
  SMH_CHECK(ctx, A->bdim == 3 && (A->precision == SMH_F32 || A->precision == SMH_F32X2),
            SMH_ERR_INVALID,
            "smh_bsr_planar: needs 3x3 float values (SMH_F32, or the hi part of SMH_F32X2)");
  if (A->nnz == 0) return SMH_OK;
  hipLaunchKernelGGL(k_to_planar, dim3(grid_for((A->nnz + 63) & ~63ull, 256)), dim3(256), 0,
                     ctx->stream, A->nnz, (const float *)A->vals, d_out);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" int smh_bsrmv(smh_ctx *ctx, const smh_bsr *A, const void *x, void *y,
                         int32_t exact_order) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && x && y && A->rowptr && (A->nnz == 0 || (A->colidx && A->vals)),
            SMH_ERR_INVALID, "smh_bsrmv: null argument");
  SMH_CHECK(ctx, A->bdim == 1 || A->bdim == 3 || (exact_order && A->bdim >= 1 && A->bdim <= 8),
            SMH_ERR_UNSUPPORTED,
            "smh_bsrmv: block size %u (1 or 3; 1..8 with exact_order)", A->bdim);
  SMH_CHECK(ctx, A->precision >= SMH_F64 && A->precision <= SMH_F32P, SMH_ERR_INVALID,
            "smh_bsrmv: unknown precision %d", A->precision);
  SMH_CHECK(ctx, A->precision != SMH_F32X2 || A->vals_lo, SMH_ERR_INVALID,
            "smh_bsrmv: SMH_F32X2 needs vals_lo");
  SMH_CHECK(ctx, A->precision != SMH_F32P || (A->bdim == 3 && !exact_order), SMH_ERR_INVALID,
            "smh_bsrmv: SMH_F32P holds 3x3 blocks (and has no exact_order form)");
  if (A->nrows == 0) return SMH_OK;
  if (A->precision == SMH_F32P) {
    smh_tile_plan pplan;
    const uint32_t nb = (uint32_t)TileShapePl::kBlocks;
    int rc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, nb, &pplan);
    if (rc) return rc;
    SMH_CHECK(ctx, pplan.tile_row != nullptr, SMH_ERR_UNSUPPORTED,
              "smh_bsrmv: a row holds %u blocks (at most %u with SMH_F32P)", pplan.maxrow,
              nb / 2);
    unsigned g = 0;
    const TileDev T = tile_dev(A, &pplan, &g, true, SMH_TILE_GRID_PL);
    hipLaunchKernelGGL(k_bsrmv_planar, dim3(g), dim3(TileShapePl::kThreads), 0, ctx->stream, T,
                       (const float *)x, (float *)y);
    SMH_LAUNCH_CHECK(ctx);
    return SMH_OK;
  }
  if (exact_order) {
    SMH_CHECK(ctx, A->precision == SMH_F64, SMH_ERR_INVALID,
              "smh_bsrmv: exact_order needs SMH_F64 values");
    const uint64_t n = (uint64_t)A->nrows * A->bdim;
    hipLaunchKernelGGL(k_bsrmv_exact, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, A->nrows,
                       (int)A->bdim, A->rowptr, A->colidx, (const double *)A->vals,
                       (const double *)x, (double *)y);
    SMH_LAUNCH_CHECK(ctx);
    return SMH_OK;
  }
  smh_tile_plan plan;
  const uint32_t nb = tile_blocks_host((int)A->bdim);
  int rc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, nb, &plan);
  if (rc) return rc;
  SMH_CHECK(ctx, plan.tile_row != nullptr, SMH_ERR_UNSUPPORTED,
            "smh_bsrmv: a row holds %u blocks (at most %u outside exact_order)", plan.maxrow,
            nb / 2);
  unsigned g = 0;
  const TileDev T = tile_dev(A, &plan, &g, A->bdim == 1 && SMH_TILE_INTERLEAVE1);
#define SMH_MV(P)                                                                            \
  do {                                                                                       \
    using VT = typename Val<P>::T;                                                           \
    if (A->bdim == 3)                                                                        \
      hipLaunchKernelGGL((k_bsrmv_tiles<P, 3>), dim3(g), dim3(tile_threads_host(3)), 0,   \
                         ctx->stream, T,                                                    \
                         (const VT *)x, (VT *)y);                                            \
    else                                                                                     \
      hipLaunchKernelGGL((k_bsrmv_tiles<P, 1>), dim3(g), dim3(256), 0, ctx->stream, T,     \
                         (const VT *)x, (VT *)y);                                            \
  } while (0)
  if (A->precision == SMH_F64) SMH_MV(SMH_F64);
  else if (A->precision == SMH_F32) SMH_MV(SMH_F32);
  else if (A->precision == SMH_F32D) SMH_MV(kF32D);
  else SMH_MV(SMH_F32X2);
#undef SMH_MV
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" int smh_block_jacobi(smh_ctx *ctx, const smh_bsr *A, double *d_dinv) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && d_dinv && (A->bdim == 1 || A->bdim == 3), SMH_ERR_INVALID,
            "smh_block_jacobi: bad argument");
  SMH_CHECK(ctx, A->precision >= SMH_F64 && A->precision <= SMH_F32D, SMH_ERR_INVALID,
            "smh_block_jacobi: unknown precision %d", A->precision);
  const int pc = A->bdim == 3 ? SMH_PRECOND_BLOCK_JACOBI : SMH_PRECOND_JACOBI;
  const unsigned g = grid_for(A->nrows, 256);
  if (A->precision == SMH_F64)
    hipLaunchKernelGGL((k_dinv<SMH_F64, double>), dim3(g), dim3(256), 0, ctx->stream, A->nrows,
                       A->rowptr, A->colidx, A->vals, A->vals_lo, (int)A->bdim, pc,
                       d_dinv);
  else if (A->precision == SMH_F32X2)
    hipLaunchKernelGGL((k_dinv<SMH_F32X2, double>), dim3(g), dim3(256), 0, ctx->stream, A->nrows,
                       A->rowptr, A->colidx, A->vals, A->vals_lo, (int)A->bdim, pc,
                       d_dinv);
  else
    hipLaunchKernelGGL((k_dinv<kF32D, double>), dim3(g), dim3(256), 0, ctx->stream, A->nrows,
                       A->rowptr, A->colidx, A->vals, A->vals_lo, (int)A->bdim, pc,
                       d_dinv);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" void smh_cg_default_opts(smh_cg_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->maxiters = 1000;
  o->tol = 1e-8;
  o->rel = 1;
  o->precond = SMH_PRECOND_NONE;
  o->reliable_delta = 0.3;
  o->check_every = 16;
  o->fallback = 1;
}

namespace {
// The fallback chain of the single-system solvers: the requested
// preconditioner, then (block) Jacobi, then none.  A rung is retried on the
// next one when it breaks down (p'q <= 0, z'r <= 0, NaN) or, for the learned
// preconditioner only, when it ends without convergence.  Every rung starts
// from the caller's initial guess and is certified by its own FP64 residual;
// the stats report the total iterations, the preconditioner that produced x
// and the first rung's status.
template <class RUN>
int solve_chain(smh_ctx *ctx, uint64_t n, int bdim, double *x, const smh_cg_opts *o,
                smh_cg_stats *st, RUN &&run, Arena *ar = nullptr) {
  smh_cg_stats s{};
  int chain[3] = {o->precond, 0, 0}, nchain = 1;
  if (o->fallback) {
    if (o->precond == SMH_PRECOND_LEARNED)
      chain[nchain++] = bdim == 3 ? SMH_PRECOND_BLOCK_JACOBI : SMH_PRECOND_JACOBI;
    if (o->precond != SMH_PRECOND_NONE) chain[nchain++] = SMH_PRECOND_NONE;
  }
  double *x0 = nullptr;
  int rc = SMH_OK;
  if (nchain > 1) {
    if (ar) {
      x0 = (double *)ar->take(sizeof(double) * (n ? n : 1));
      if (!x0)
        return fail(ctx, SMH_ERR_UNSUPPORTED, "CG: the %zu-byte scratch array is too small",
                    ar->cap);
    } else {
      rc = dmalloc(ctx, (void **)&x0, sizeof(double) * (n ? n : 1));
      if (rc) return rc;
    }
    if (hipMemcpyAsync(x0, x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess) {
      if (!ar) dfree(x0);
      return fail(ctx, SMH_ERR_HIP, "smh_cg_solve: copy of the initial guess failed");
    }
  }
  smh_cg_opts oo = *o;
  int total = 0, nru = 0, first = SMH_OK;
  double secs = 0.0;
  int k = 0;
  for (;; ++k) {
    oo.precond = chain[k];
    if (k > 0 &&
        hipMemcpyAsync(x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream) !=
            hipSuccess) {
      rc = fail(ctx, SMH_ERR_HIP, "smh_cg_solve: reset to the initial guess failed");
      break;
    }
    s = smh_cg_stats{};
    const size_t mark = ar ? ar->used : 0;  // every rung reuses the same pieces
    rc = run(&oo, &s);
    if (ar) ar->used = mark;
    if (rc) break;
    total += s.iters;
    nru += s.reliable_updates;
    secs += s.seconds;
    if (k == 0) first = s.status;
    const bool retry = s.status == SMH_ERR_BREAKDOWN ||
                       (chain[k] == SMH_PRECOND_LEARNED && s.status == SMH_ERR_NOT_CONVERGED);
    if (!retry || k + 1 >= nchain) break;
  }
  if (x0 && !ar) {
    (void)hipStreamSynchronize(ctx->stream);
    dfree(x0);
  }
  if (!rc && st) {
    *st = s;
    st->iters = total;
    st->reliable_updates = nru;
    st->seconds = secs;
    st->precond_used = chain[k];
    st->fallbacks = k;
    st->first_status = first;
  }
  return rc;
}
}  // namespace

namespace {
int cg_validate(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x, const smh_cg_opts *o,
                const char *who) {
  SMH_CHECK(ctx, A && b && x && o && A->rowptr, SMH_ERR_INVALID, "%s: null argument", who);
  SMH_CHECK(ctx, A->bdim == 1 || A->bdim == 3, SMH_ERR_UNSUPPORTED,
            "%s: block size %u (1 or 3 supported)", who, A->bdim);
  SMH_CHECK(ctx, o->precond >= 0 && o->precond <= 3, SMH_ERR_INVALID,
            "%s: unknown preconditioner %d", who, o->precond);
  SMH_CHECK(ctx,
            o->precond != SMH_PRECOND_LEARNED ||
                (o->learned && o->learned_corr && ((uintptr_t)o->learned_corr & 3u) == 0),
            SMH_ERR_INVALID, "%s: the learned preconditioner needs opts.learned and "
                             "opts.learned_corr (smh_gnn_precond)", who);
  SMH_CHECK(ctx, A->bdim == 3 || o->precond != SMH_PRECOND_BLOCK_JACOBI, SMH_ERR_INVALID,
            "%s: block Jacobi needs 3x3 blocks", who);
  SMH_CHECK(ctx, !o->mixed || A->precision == SMH_F32X2, SMH_ERR_INVALID,
            "%s: mixed precision needs an SMH_F32X2 matrix", who);
  SMH_CHECK(ctx, A->precision >= SMH_F64 && A->precision <= SMH_F32D, SMH_ERR_INVALID,
            "%s: unknown precision %d", who, A->precision);
  SMH_CHECK(ctx, A->precision != SMH_F32X2 || A->vals_lo, SMH_ERR_INVALID,
            "%s: SMH_F32X2 needs vals_lo", who);
  SMH_CHECK(ctx, o->maxiters >= 0, SMH_ERR_INVALID, "%s: negative maxiters", who);
  SMH_CHECK(ctx, o->mixed >= 0 && o->mixed <= 2, SMH_ERR_INVALID,
            "%s: mixed must be 0, 1 or 2 (got %d)", who, o->mixed);
  SMH_CHECK(ctx, !o->vals_f64 || (o->mixed && ((uintptr_t)o->vals_f64 & 7u) == 0),
            SMH_ERR_INVALID, "%s: vals_f64 (8-byte aligned) goes with mixed precision", who);
  return SMH_OK;
}

// the solve of smh_cg_solve / smh_cg_solve_noalloc over a tile plan
int cg_dispatch(smh_ctx *ctx, const smh_bsr *A, const smh_tile_plan *plan, const double *b,
                double *x, const smh_cg_opts *o, smh_cg_stats *st, Arena *ar,
                const PlanarIn *pin = nullptr) {
  const int bd = (int)A->bdim;
  auto run = [&](const smh_cg_opts *oo, smh_cg_stats *so) {
#define SMH_CG(IP, RP)                                                                  \
  (A->bdim == 3 ? cg_tiles<IP, RP, 3>(ctx, A, plan, plan, b, x, oo, so, ar, pin) \
                : cg_tiles<IP, RP, 1>(ctx, A, plan, plan, b, x, oo, so, ar, pin))
    if (oo->mixed == 2) return oo->vals_f64 ? SMH_CG(kF32D, SMH_F64) : SMH_CG(kF32D, SMH_F32X2);
    if (oo->mixed) return oo->vals_f64 ? SMH_CG(SMH_F32, SMH_F64) : SMH_CG(SMH_F32, SMH_F32X2);
    if (A->precision == SMH_F64) return SMH_CG(SMH_F64, SMH_F64);
    if (A->precision == SMH_F32X2) return SMH_CG(SMH_F32X2, SMH_F32X2);
    return SMH_CG(kF32D, kF32D);
#undef SMH_CG
  };
  return solve_chain(ctx, (uint64_t)A->nrows * bd, bd, x, o, st, run, ar);
}
}  // namespace

extern "C" int smh_cg_solve(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x,
                            const smh_cg_opts *o, smh_cg_stats *st) {
  if (!ctx) return SMH_ERR_INVALID;
  int rc = cg_validate(ctx, A, b, x, o, "smh_cg_solve");
  if (rc) return rc;
  if (A->nrows == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    return SMH_OK;
  }
  const int bd = (int)A->bdim;
  smh_tile_plan plan;
  rc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, tile_blocks_host(bd), &plan);
  if (rc) return rc;
  SMH_CHECK(ctx, plan.tile_row != nullptr, SMH_ERR_UNSUPPORTED,
            "smh_cg_solve: a row holds %u blocks (at most %u supported)", plan.maxrow,
            tile_blocks_host(bd) / 2);
  return cg_dispatch(ctx, A, &plan, b, x, o, st, nullptr);
}

extern "C" int smh_cg_solve_noalloc(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x,
                                    const smh_cg_opts *o, smh_cg_stats *st) {
  if (!ctx) return SMH_ERR_INVALID;
  const char *who = "smh_cg_solve_noalloc";
  int rc = cg_validate(ctx, A, b, x, o, who);
  if (rc) return rc;
  SMH_CHECK(ctx, o->precond != SMH_PRECOND_LEARNED, SMH_ERR_UNSUPPORTED,
            "%s: the learned preconditioner is not supported (none, Jacobi, block Jacobi)", who);
  SMH_CHECK(ctx, ctx->scratch != nullptr, SMH_ERR_HIP,
            "%s: the scratch array of this device is not available", who);
  if (A->nrows == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    return SMH_OK;
  }
  // everything the solve needs, bounded before any work is queued
  const int bd = (int)A->bdim;
  const uint64_t nrows = A->nrows, n = nrows * bd, bb = (uint64_t)bd * bd;
  const uint32_t tb = tile_blocks_host(bd);
  const uint64_t vt = o->mixed == 1 ? 4 : 8;
  const uint64_t max_tiles = A->nnz / (tb / 2 + 1) + 1;
  uint64_t need = 256 + 4 * (max_tiles + 1) + 256;            // row-max counter, tile rows
  if (o->fallback && o->precond != SMH_PRECOND_NONE) need += 8 * n + 256;  // initial guess
  need += 4 * (vt * n + 256);                                  // r, z, p, q
  if (o->precond != SMH_PRECOND_NONE) need += vt * nrows * bb + 256;  // (block) Jacobi
  if (o->mixed && A->precision == SMH_F32X2) need += 4 * nrows * bb + 256;  // low row sums
  need += sizeof(CGState) + 256 + sizeof(double) * 2 * kPartSlots + 64 + 256;
  // mixed 3x3: the lane-planar copy of the values and its tile rows, as
  // smh_cg_solve streams them
  const bool want_pl = bd == 3 && (o->mixed == 1 || o->mixed == 2);
  const uint32_t tbp = (uint32_t)TileShapePlS::kBlocks;
  const uint64_t max_tiles_pl = A->nnz / (tbp / 2 + 1) + 1;
  if (want_pl) need += sizeof(float) * planar_floats(A->nnz) + 256 + 4 * (max_tiles_pl + 1) + 256;
  SMH_CHECK(ctx, need <= ctx->scratch_bytes, SMH_ERR_UNSUPPORTED,
            "%s: this system needs %llu bytes of scratch; the maximum is %zu (use smh_cg_solve)",
            who, (unsigned long long)need, ctx->scratch_bytes);
  Arena ar{(unsigned char *)ctx->scratch, ctx->scratch_bytes, 0};
  // the row-tile plan, built into the scratch array on every call
  unsigned *dmax = (unsigned *)ar.take(sizeof(unsigned));
  SMH_HIP(ctx, hipMemsetAsync(dmax, 0, sizeof(unsigned), ctx->stream));
  hipLaunchKernelGGL(k_rowmax, dim3(grid_for(A->nrows, 256, 2048)), dim3(256), 0, ctx->stream,
                     A->rowptr, A->nrows, dmax);
  unsigned maxrow = 0;
  SMH_HIP(ctx, hipMemcpyAsync(&maxrow, dmax, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SMH_CHECK(ctx, maxrow <= tb / 2, SMH_ERR_UNSUPPORTED,
            "%s: a row holds %u blocks (at most %u supported)", who, maxrow, tb / 2);
  smh_tile_plan plan;
  plan.rowptr = A->rowptr;
  plan.nrows = A->nrows;
  plan.nnz = A->nnz;
  plan.maxrow = maxrow;
  plan.tile_blocks = tb;
  plan.tprime = tb - (maxrow ? maxrow : 1) + 1;
  const uint64_t nt = std::max<uint64_t>(1, (A->nnz + plan.tprime - 1) / plan.tprime);
  plan.ntiles = (uint32_t)nt;
  plan.tile_row = (uint32_t *)ar.take(sizeof(uint32_t) * (nt + 1));
  SMH_CHECK(ctx, nt <= max_tiles && plan.tile_row, SMH_ERR_UNSUPPORTED,
            "%s: the tile plan does not fit the scratch array", who);
  hipLaunchKernelGGL(k_tile_rows, dim3(grid_for(nt + 1, 256)), dim3(256), 0, ctx->stream,
                     A->rowptr, A->nrows, plan.tprime, plan.ntiles, plan.tile_row);
  SMH_LAUNCH_CHECK(ctx);
  smh_tile_plan pplan;
  PlanarIn pin;
  if (want_pl && A->nnz > 0 && maxrow <= tbp / 2) {
    pplan = plan;
    pplan.tile_blocks = tbp;
    pplan.tprime = tbp - (maxrow ? maxrow : 1) + 1;
    const uint64_t ntp = std::max<uint64_t>(1, (A->nnz + pplan.tprime - 1) / pplan.tprime);
    pplan.ntiles = (uint32_t)ntp;
    pplan.tile_row = (uint32_t *)ar.take(sizeof(uint32_t) * (ntp + 1));
    pin.vals = (float *)ar.take(sizeof(float) * planar_floats(A->nnz));
    SMH_CHECK(ctx, ntp <= max_tiles_pl && pplan.tile_row && pin.vals, SMH_ERR_UNSUPPORTED,
              "%s: the planar values do not fit the scratch array", who);
    hipLaunchKernelGGL(k_tile_rows, dim3(grid_for(ntp + 1, 256)), dim3(256), 0, ctx->stream,
                       A->rowptr, A->nrows, pplan.tprime, pplan.ntiles, pplan.tile_row);
    SMH_LAUNCH_CHECK(ctx);
    pin.plan = &pplan;
  }
  return cg_dispatch(ctx, A, &plan, b, x, o, st, &ar, &pin);
}

namespace {
int cg_batched(smh_ctx *ctx, const smh_bsr *A, uint32_t nbatch, const uint32_t *rows,
               const double *b, double *x, const smh_cg_opts *o, smh_cg_batch_stats *stats,
               double *seconds, const std::function<int(hipStream_t)> &prologue,
               const std::string &pro_key) {
  SMH_CHECK(ctx, A && rows && b && x && o && A->rowptr && nbatch > 0, SMH_ERR_INVALID,
            "smh_cg_solve_batched: null argument");
  SMH_CHECK(ctx, A->bdim == 1 || A->bdim == 3, SMH_ERR_UNSUPPORTED,
            "smh_cg_solve_batched: block size %u (1 or 3 supported)", A->bdim);
  SMH_CHECK(ctx, o->precond >= 0 && o->precond <= SMH_PRECOND_BLOCK_JACOBI, SMH_ERR_INVALID,
            "smh_cg_solve_batched: preconditioner %d (none, Jacobi or block Jacobi)", o->precond);
  SMH_CHECK(ctx, A->bdim == 3 || o->precond <= SMH_PRECOND_JACOBI, SMH_ERR_INVALID,
            "smh_cg_solve_batched: block Jacobi needs 3x3 blocks");
  SMH_CHECK(ctx, !o->mixed || A->precision == SMH_F32X2, SMH_ERR_INVALID,
            "smh_cg_solve_batched: mixed precision needs an SMH_F32X2 matrix");
  SMH_CHECK(ctx, o->mixed == 0 || o->mixed == 1, SMH_ERR_INVALID,
            "smh_cg_solve_batched: mixed must be 0 or 1");
  SMH_CHECK(ctx, !o->vals_f64, SMH_ERR_UNSUPPORTED,
            "smh_cg_solve_batched: vals_f64 is not supported (residuals use hi + lo)");
  SMH_CHECK(ctx, A->precision >= SMH_F64 && A->precision <= SMH_F32D, SMH_ERR_INVALID,
            "smh_cg_solve_batched: unknown precision %d", A->precision);
  SMH_CHECK(ctx, A->precision != SMH_F32X2 || A->vals_lo, SMH_ERR_INVALID,
            "smh_cg_solve_batched: SMH_F32X2 needs vals_lo");
  SMH_CHECK(ctx, rows[0] == 0 && rows[nbatch] == A->nrows, SMH_ERR_INVALID,
            "smh_cg_solve_batched: rows must run from 0 to nrows");
  for (uint32_t m = 0; m < nbatch; ++m)
    SMH_CHECK(ctx, rows[m] <= rows[m + 1], SMH_ERR_INVALID,
              "smh_cg_solve_batched: rows must not decrease (system %u)", m);
  if (A->nrows == 0) return SMH_OK;
  const int bd = (int)A->bdim;
  smh_tile_plan plan;
  int rc = tile_plan(ctx, A->rowptr, A->nrows, A->nnz, tile_blocks_host(bd), &plan);
  if (rc) return rc;
  SMH_CHECK(ctx, plan.tile_row != nullptr, SMH_ERR_UNSUPPORTED,
            "smh_cg_solve_batched: a row holds %u blocks (at most %u supported)", plan.maxrow,
            tile_blocks_host(bd) / 2);
#define SMH_BCG(IP, RP)                                                                      \
  (bd == 3 ? bcg_run<IP, RP, 3>(ctx, A, &plan, nbatch, rows, b, x, o, stats, seconds,   \
                                 prologue, pro_key)                                          \
           : bcg_run<IP, RP, 1>(ctx, A, &plan, nbatch, rows, b, x, o, stats, seconds,   \
                                 prologue, pro_key))
  if (o->mixed) rc = SMH_BCG(SMH_F32, SMH_F32X2);
  else if (A->precision == SMH_F64) rc = SMH_BCG(SMH_F64, SMH_F64);
  else if (A->precision == SMH_F32X2) rc = SMH_BCG(SMH_F32X2, SMH_F32X2);
  else rc = SMH_BCG(kF32D, kF32D);
#undef SMH_BCG
  return rc;
}
}  // namespace

extern "C" int smh_cg_solve_batched(smh_ctx *ctx, const smh_bsr *A, uint32_t nbatch,
                                    const uint32_t *rows, const double *b, double *x,
                                    const smh_cg_opts *o, smh_cg_batch_stats *stats,
                                    double *seconds) {
  if (!ctx) return SMH_ERR_INVALID;
  return cg_batched(ctx, A, nbatch, rows, b, x, o, stats, seconds, nullptr, std::string());
}

extern "C" int smh_springs_solve_batched(smh_ctx *ctx, const smh_index *idx,
                                         const smh_springs_params *p, const smh_springs_set *set,
                                         void *vals, void *vals_lo, double *rhs, uint32_t nbatch,
                                         const uint32_t *rows, double *x, const smh_cg_opts *o,
                                         smh_cg_batch_stats *stats, double *seconds) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, idx && p && set && vals && rhs && o, SMH_ERR_INVALID,
            "smh_springs_solve_batched: null argument");
  SMH_CHECK(ctx, o->mixed == 0 || vals_lo, SMH_ERR_INVALID,
            "smh_springs_solve_batched: mixed precision needs vals_lo (SMH_F32X2 values)");
  const int prec = o->mixed ? SMH_F32X2 : SMH_F64;
  const smh_bsr A{idx->nrows, 3, idx->nnz, idx->rowptr, idx->colidx, prec, vals,
                  o->mixed ? vals_lo : nullptr};
  // the assembly is the first work of the first captured graph (its plan is
  // built before the capture)
  if (const int pr = smh::prepare_assembly(ctx, idx, set->endpoints, 2)) return pr;
  auto prologue = [&](hipStream_t s) {
    hipStream_t caller = ctx->stream;
    ctx->stream = s;
    const int r = smh_assemble_springs(ctx, idx, p, set, prec, vals, vals_lo, rhs);
    ctx->stream = caller;
    return r;
  };
  // the graphs hold the assembly's arguments by value: they are part of the key
  std::string key((const char *)p, sizeof(*p));
  key.append((const char *)set, sizeof(*set));
  key.append((const char *)&vals, sizeof(vals));
  key.append((const char *)&vals_lo, sizeof(vals_lo));
  key.append((const char *)&rhs, sizeof(rhs));
  key.append((const char *)&idx, sizeof(idx));
  return cg_batched(ctx, &A, nbatch, rows, rhs, x, o, stats, seconds, prologue, key);
}

// ---- scalar 27-point stencil storage: conversion, product and CG -------------------------
namespace {
int check_stencil(smh_ctx *ctx, const smh_stencil27 *S, const char *who) {
  SMH_CHECK(ctx, S && S->vals, SMH_ERR_INVALID, "%s: null matrix", who);
  SMH_CHECK(ctx, S->n0 >= 1 && S->n1 >= 1 && S->n2 >= 1, SMH_ERR_INVALID, "%s: empty grid", who);
  SMH_CHECK(ctx, (uint64_t)S->n0 * S->n1 * S->n2 < (1ull << 31), SMH_ERR_UNSUPPORTED,
            "%s: grid too large", who);
  SMH_CHECK(ctx, S->precision >= SMH_F64 && S->precision <= SMH_F32D, SMH_ERR_INVALID,
            "%s: unknown precision %d", who, S->precision);
  SMH_CHECK(ctx, S->precision != SMH_F32X2 || S->vals_lo, SMH_ERR_INVALID,
            "%s: SMH_F32X2 needs vals_lo", who);
  SMH_CHECK(ctx, (S->flags & ~SMH_STENCIL_PERIODIC) == 0, SMH_ERR_INVALID, "%s: unknown flags %u",
            who, S->flags);
  return SMH_OK;
}
}  // namespace

extern "C" int smh_stencil27_from_bsr(smh_ctx *ctx, const smh_bsr *A, uint32_t n0, uint32_t n1,
                                      uint32_t n2, void *vals, void *vals_lo) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && A->rowptr && vals, SMH_ERR_INVALID, "smh_stencil27_from_bsr: null argument");
  SMH_CHECK(ctx, A->bdim == 1, SMH_ERR_UNSUPPORTED,
            "smh_stencil27_from_bsr: 1x1 blocks only (got %u)", A->bdim);
  SMH_CHECK(ctx, (uint64_t)n0 * n1 * n2 == A->nrows, SMH_ERR_INVALID,
            "smh_stencil27_from_bsr: %u x %u x %u grid for %u rows", n0, n1, n2, A->nrows);
  SMH_CHECK(ctx, A->precision >= SMH_F64 && A->precision <= SMH_F32D, SMH_ERR_INVALID,
            "smh_stencil27_from_bsr: unknown precision %d", A->precision);
  SMH_CHECK(ctx, A->nnz == 0 || (A->colidx && A->vals), SMH_ERR_INVALID,
            "smh_stencil27_from_bsr: null values");
  const bool lo = A->precision == SMH_F32X2;
  SMH_CHECK(ctx, !lo || (A->vals_lo && vals_lo), SMH_ERR_INVALID,
            "smh_stencil27_from_bsr: SMH_F32X2 needs vals_lo in and out");
  if (A->nrows == 0) return SMH_OK;
  smh_stencil27 S{n0, n1, n2, A->precision, vals, vals_lo};
  const StDev D = st_dev(&S, vals, vals_lo);
  int *err = (int *)ctx->red_count;
  SMH_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
  const unsigned g = (unsigned)((A->nrows + 63) / 64);
  if (A->precision == SMH_F64)
    hipLaunchKernelGGL((k_st_convert<double, double>), dim3(g), dim3(256), 0, ctx->stream, D,
                       A->rowptr,
                       A->colidx, (const double *)A->vals, (const float *)nullptr,
                       (double *)vals, (float *)nullptr, err);
  else
    hipLaunchKernelGGL((k_st_convert<float, float>), dim3(g), dim3(256), 0, ctx->stream, D,
                       A->rowptr,
                       A->colidx, (const float *)A->vals,
                       lo ? (const float *)A->vals_lo : (const float *)nullptr, (float *)vals,
                       lo ? (float *)vals_lo : (float *)nullptr, err);
  SMH_LAUNCH_CHECK(ctx);
  int herr = 0;
  SMH_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SMH_CHECK(ctx, herr == 0, SMH_ERR_INVALID,
            "smh_stencil27_from_bsr: a column lies outside the 27-point stencil of its row");
  return SMH_OK;
}

extern "C" int smh_stencil27_correction(smh_ctx *ctx, const smh_bsr *C, uint32_t n0,
                                        uint32_t n1, uint32_t n2, void *out) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, C && C->rowptr && out, SMH_ERR_INVALID,
            "smh_stencil27_correction: null argument");
  SMH_CHECK(ctx, C->bdim == 1 && C->precision == SMH_F32, SMH_ERR_INVALID,
            "smh_stencil27_correction: the correction is a 1x1 SMH_F32 matrix");
  SMH_CHECK(ctx, (uint64_t)n0 * n1 * n2 == C->nrows, SMH_ERR_INVALID,
            "smh_stencil27_correction: %u x %u x %u grid for %u rows", n0, n1, n2, C->nrows);
  SMH_CHECK(ctx, ((uintptr_t)out & 3u) == 0, SMH_ERR_INVALID,
            "smh_stencil27_correction: out must be 4-byte aligned");
  if (C->nrows == 0) return SMH_OK;
  SMH_CHECK(ctx, C->nnz == 0 || (C->colidx && C->vals), SMH_ERR_INVALID,
            "smh_stencil27_correction: null values");
  smh_stencil27 S{n0, n1, n2, SMH_F32, out, nullptr};
  const StDev D = st_dev(&S, out, nullptr);
  int *err = (int *)ctx->red_count;
  SMH_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL((k_st_convert<float, uint16_t>), dim3((unsigned)((C->nrows + 63) / 64)),
                     dim3(256), 0, ctx->stream, D, C->rowptr, C->colidx, (const float *)C->vals,
                     (const float *)nullptr, (uint16_t *)out, (float *)nullptr, err);
  SMH_LAUNCH_CHECK(ctx);
  int herr = 0;
  SMH_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SMH_CHECK(ctx, herr == 0, SMH_ERR_INVALID,
            "smh_stencil27_correction: a column lies outside the 27-point stencil of its row");
  return SMH_OK;
}

extern "C" int smh_stencil27_mv(smh_ctx *ctx, const smh_stencil27 *S, const void *x, void *y) {
  if (!ctx) return SMH_ERR_INVALID;
  int rc = check_stencil(ctx, S, "smh_stencil27_mv");
  if (rc) return rc;
  SMH_CHECK(ctx, x && y, SMH_ERR_INVALID, "smh_stencil27_mv: null vector");
  const StDev D = st_dev(S, S->vals, S->vals_lo);
  const unsigned g = grid_for(D.n, 256, SMH_ST_GRID);
  if (S->flags & SMH_STENCIL_PERIODIC) {
    SMH_CHECK(ctx, S->precision == SMH_F64 || S->precision == SMH_F32, SMH_ERR_UNSUPPORTED,
              "smh_stencil27_mv: periodic stencils hold SMH_F64 or SMH_F32 values");
    if (S->precision == SMH_F64)
      hipLaunchKernelGGL((k_stmv<SMH_F64, true>), dim3(g), dim3(256), 0, ctx->stream, D,
                         (const double *)x, (double *)y);
    else
      hipLaunchKernelGGL((k_stmv<SMH_F32, true>), dim3(g), dim3(256), 0, ctx->stream, D,
                         (const float *)x, (float *)y);
    SMH_LAUNCH_CHECK(ctx);
    return SMH_OK;
  }
  if (S->precision == SMH_F64)
    hipLaunchKernelGGL((k_stmv<SMH_F64>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (const double *)x, (double *)y);
  else if (S->precision == SMH_F32)
    hipLaunchKernelGGL((k_stmv<SMH_F32>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (const float *)x, (float *)y);
  else if (S->precision == SMH_F32D)
    hipLaunchKernelGGL((k_stmv<kF32D>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (const double *)x, (double *)y);
  else
    hipLaunchKernelGGL((k_stmv<SMH_F32X2>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (const double *)x, (double *)y);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

namespace {
template <int IP, int RP>
int cg_stencil(smh_ctx *ctx, const smh_stencil27 *S, const double *b, double *x,
               const smh_cg_opts *o, smh_cg_stats *st) {
  using VT = typename Val<IP>::T;
  // the correction's FP16 slots, gathered u in FP32 unless the solve is all FP64
  constexpr int LP = IP == SMH_F64 ? kF16D : kF16F;
  StencilOps<IP, RP, LP> ops;
  ops.A = st_dev(S, S->vals, S->vals_lo);
  ops.R = RP == SMH_F64 && o->vals_f64 ? st_dev(S, o->vals_f64, nullptr) : ops.A;
  ops.C = st_dev(S, o->precond == SMH_PRECOND_LEARNED ? o->learned_corr : S->vals, nullptr);
  ops.g = grid_for(ops.A.n, 256, SMH_ST_GRID);
  return cg_run<IP, RP, 1>(ctx, ops, ops.A.n, S->precision == SMH_F32X2 && S->vals_lo, b, x, o,
                           st);
}
}  // namespace

extern "C" int smh_cg_solve_stencil27(smh_ctx *ctx, const smh_stencil27 *S, const double *b,
                                      double *x, const smh_cg_opts *o, smh_cg_stats *st) {
  if (!ctx) return SMH_ERR_INVALID;
  int rc = check_stencil(ctx, S, "smh_cg_solve_stencil27");
  if (rc) return rc;
  SMH_CHECK(ctx, b && x && o, SMH_ERR_INVALID, "smh_cg_solve_stencil27: null argument");
  SMH_CHECK(ctx, !(S->flags & SMH_STENCIL_PERIODIC), SMH_ERR_UNSUPPORTED,
            "smh_cg_solve_stencil27: periodic stencils are not supported by the solver");
  SMH_CHECK(ctx, o->precond >= 0 && o->precond <= 3, SMH_ERR_INVALID,
            "smh_cg_solve_stencil27: unknown preconditioner %d", o->precond);
  SMH_CHECK(ctx,
            o->precond != SMH_PRECOND_LEARNED ||
                (o->learned && o->learned_corr && ((uintptr_t)o->learned_corr & 3u) == 0),
            SMH_ERR_INVALID, "smh_cg_solve_stencil27: the learned preconditioner needs "
                             "opts.learned (float[n]) and opts.learned_corr (FP16 [27 n], "
                             "smh_stencil27_correction)");
  SMH_CHECK(ctx, !o->mixed || S->precision == SMH_F32X2, SMH_ERR_INVALID,
            "smh_cg_solve_stencil27: mixed precision needs an SMH_F32X2 matrix");
  SMH_CHECK(ctx, o->maxiters >= 0, SMH_ERR_INVALID, "smh_cg_solve_stencil27: negative maxiters");
  SMH_CHECK(ctx, o->mixed >= 0 && o->mixed <= 2, SMH_ERR_INVALID,
            "smh_cg_solve_stencil27: mixed must be 0, 1 or 2 (got %d)", o->mixed);
  SMH_CHECK(ctx, !o->vals_f64 || (o->mixed && ((uintptr_t)o->vals_f64 & 7u) == 0),
            SMH_ERR_INVALID,
            "smh_cg_solve_stencil27: vals_f64 (8-byte aligned) goes with mixed precision");
  auto run = [&](const smh_cg_opts *oo, smh_cg_stats *so) {
    if (oo->mixed == 2)
      return oo->vals_f64 ? cg_stencil<kF32D, SMH_F64>(ctx, S, b, x, oo, so)
                          : cg_stencil<kF32D, SMH_F32X2>(ctx, S, b, x, oo, so);
    if (oo->mixed)
      return oo->vals_f64 ? cg_stencil<SMH_F32, SMH_F64>(ctx, S, b, x, oo, so)
                          : cg_stencil<SMH_F32, SMH_F32X2>(ctx, S, b, x, oo, so);
    if (S->precision == SMH_F64) return cg_stencil<SMH_F64, SMH_F64>(ctx, S, b, x, oo, so);
    if (S->precision == SMH_F32X2) return cg_stencil<SMH_F32X2, SMH_F32X2>(ctx, S, b, x, oo, so);
    return cg_stencil<kF32D, kF32D>(ctx, S, b, x, oo, so);
  };
  return solve_chain(ctx, (uint64_t)S->n0 * S->n1 * S->n2, 1, x, o, st, run);
}
