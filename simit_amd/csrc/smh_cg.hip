// Blocked SpMV entry point and the device-resident CG / PCG solver.
//
// The CG loop is what Simit programs write inline (apps/springs/isprings.sim:
// 77-93, apps/fem/fem_linear.sim:211-230, test/input/program/cg.sim:25-53,
// apps/thermal/thermal.sim:75-130).  One iteration here is three kernels:
//   K1  q = A p  (wavefront per block row) + FP64 partial of p'q; the last
//       workgroup to finish reduces the partials in a fixed order and
//       computes alpha = r'z / p'q on the device;
//   K2  x += alpha p; r -= alpha q; z = M r; FP64 partials of z'r and r'r;
//       the last workgroup computes beta and the stopping test;
//   K3  p = z + beta p.
// Scalars never leave the device; the host only looks at the state every
// check_every iterations.  Every reduction has a fixed shape, so the solve
// is bitwise repeatable.
//
// Mixed precision: the inner iteration streams the FP32 matrix (hi part)
// with FP32 vectors and FP64 dots; x is accumulated in FP64.  Whenever the
// iterated residual fell by reliable_delta since the last update, or claims
// convergence, a reliable update recomputes r = b - A x in FP64 from
// hi + lo (~2^-48 relative to the FP64 matrix).  Convergence is always
// certified by that FP64 residual.
#include "smh_linalg.h"

#include <chrono>
#include <cmath>
#include <cstring>

namespace {
using namespace smh;

// float values promoted to double (an FP32 matrix driven by FP64 vectors)
constexpr int kF32D = 3;
}  // namespace

namespace smh {
template <> struct Val<kF32D> {
  using T = double;
  __device__ static __forceinline__ double ld(const void *v, const void *, uint64_t i) {
    return (double)__builtin_nontemporal_load(((const float *)v) + i);
  }
};
}  // namespace smh

namespace {

// ---- plain SpMV ---------------------------------------------------------------
#pragma clang fp contract(off)
// thread per scalar row, ascending ij then bj, separate multiply and add:
// bit-identical to the CPU loop nest.
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

template <int PREC>
__global__ void __launch_bounds__(256)
k_bsrmv3(BsrDev A, uint32_t per, const typename Val<PREC>::T *__restrict__ x,
         typename Val<PREC>::T *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  uint32_t r0, r1;
  wave_rows(A.nrows, per, r0, r1);
  for (uint32_t r = r0; r < r1; ++r) {
    const auto s = row3<PREC>(A, x, r, lane);
    if (lane == 0 || lane == 3 || lane == 6) y[(uint64_t)r * 3 + lane / 3] = s;
  }
}

template <int PREC>
__global__ void __launch_bounds__(256)
k_bsrmv1(BsrDev A, uint32_t per, const typename Val<PREC>::T *__restrict__ x,
         typename Val<PREC>::T *__restrict__ y) {
  const int lane = threadIdx.x & 63, hl = lane & 31, half = lane >> 5;
  uint32_t r0, r1;
  wave_rows(A.nrows, per, r0, r1);
  for (uint32_t r = r0; r < r1; r += 2) {
    const uint32_t rr = r + half;
    const bool act = rr < r1;
    const auto s = row1<PREC>(A, x, rr, act, hl);
    if (act && hl == 0) y[rr] = s;
  }
}

struct Launch {
  unsigned blocks;
  uint32_t per;  // rows per wavefront (3x3) or per wavefront pair of rows (1x1)
};
Launch spmv_launch(uint32_t nrows, int bd) {
  const uint64_t waves_wanted = bd == 3 ? nrows : (nrows + 1) / 2;
  uint64_t blocks = (waves_wanted + 3) / 4;
  const uint64_t cap = 2048;  // 8 workgroups of 4 waves per CU on 256 CUs
  if (blocks > cap) blocks = cap;
  if (blocks >= 8) blocks &= ~7ull;
  if (blocks < 1) blocks = 1;
  const uint64_t waves = blocks * 4;
  uint64_t per = (nrows + waves - 1) / waves;
  if (bd == 1 && (per & 1)) ++per;
  if (per < 1) per = 1;
  return {(unsigned)blocks, (uint32_t)per};
}

// ---- CG device state -------------------------------------------------------------
struct CGState {
  double rz, rz_old, pq, alpha, beta, rr, rr_true, thr, rr_ref, bb, delta2;
  int iter, done, status, want_ru, claimed, maxiters, precond_stop;
};

__device__ bool last_block(unsigned *counter) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// fixed-order sum of np partials by one 256-thread block; valid on all threads
__device__ double sum_partials(const double *part, int np, double *s4) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += __builtin_nontemporal_load(part + i);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = acc;
  __syncthreads();
  const double t = ((s4[0] + s4[1]) + s4[2]) + s4[3];
  __syncthreads();
  return t;
}

__device__ __forceinline__ bool finite_pos(double v) { return v > 0.0 && v < INFINITY; }

// K1: q = A p, p'q, alpha
template <int IP, int BD>
__global__ void __launch_bounds__(256)
k_cg_spmv(BsrDev A, uint32_t per, const typename Val<IP>::T *__restrict__ p,
          typename Val<IP>::T *__restrict__ q, CGState *S, double *part, unsigned *cnt) {
  __shared__ double s4[4];
  if (S->done) return;
  const int lane = threadIdx.x & 63;
  uint32_t r0, r1;
  wave_rows(A.nrows, per, r0, r1);
  double d = 0.0;
  if (BD == 3) {
    for (uint32_t r = r0; r < r1; ++r) {
      const auto s = row3<IP>(A, p, r, lane);
      if (lane == 0 || lane == 3 || lane == 6) {
        const uint64_t o = (uint64_t)r * 3 + lane / 3;
        q[o] = s;
        d += (double)s * (double)p[o];
      }
    }
  } else {
    const int hl = lane & 31, half = lane >> 5;
    for (uint32_t r = r0; r < r1; r += 2) {
      const uint32_t rr = r + half;
      const bool act = rr < r1;
      const auto s = row1<IP>(A, p, rr, act, hl);
      if (act && hl == 0) {
        q[rr] = s;
        d += (double)s * (double)p[rr];
      }
    }
  }
  const double t = block_sum256(d, s4);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
  if (last_block(cnt)) {
    const double pq = sum_partials(part, gridDim.x, s4);
    if (threadIdx.x == 0) {
      S->pq = pq;
      if (!finite_pos(pq)) {
        S->status = SMH_ERR_BREAKDOWN;
        S->done = 1;
      } else {
        S->alpha = S->rz / pq;
      }
      *cnt = 0;
    }
  }
}

__device__ __forceinline__ void apply_precond(int pc, int BD, const double *r, double *z,
                                              const void *dinv, bool dinv_f32, uint64_t i) {
  if (pc == SMH_PRECOND_NONE) {
    for (int a = 0; a < BD; ++a) z[a] = r[a];
  } else if (BD == 1 || pc == SMH_PRECOND_JACOBI) {
    for (int a = 0; a < BD; ++a)
      z[a] = r[a] * (dinv_f32 ? (double)((const float *)dinv)[i * BD + a]
                              : ((const double *)dinv)[i * BD + a]);
  } else {
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b)
        s += (dinv_f32 ? (double)((const float *)dinv)[i * 9 + 3 * a + b]
                       : ((const double *)dinv)[i * 9 + 3 * a + b]) * r[b];
      z[a] = s;
    }
  }
}

// stopping rule of the reference loops: r'r > thr (and z'r > thr for PCG)
__device__ __forceinline__ bool stop_test(const CGState *S, double rr, double rz) {
  return !(rr > S->thr) || (S->precond_stop && !(rz > S->thr));
}

// K2: x += alpha p; r -= alpha q; z = M r; z'r, r'r; beta; stopping test
template <typename VT, int BD>
__global__ void __launch_bounds__(256)
k_cg_update(uint32_t nrows, int pc, const void *dinv, double *__restrict__ x, VT *__restrict__ r,
            VT *__restrict__ z, const VT *__restrict__ p, const VT *__restrict__ q, CGState *S,
            double *part, unsigned *cnt, int mixed) {
  __shared__ double s4[4];
  if (S->done) return;
  const double alpha = S->alpha;
  double drz = 0.0, drr = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nrows;
       i += (uint64_t)gridDim.x * 256) {
    double rv[BD], zv[BD];
    for (int a = 0; a < BD; ++a) {
      const uint64_t o = i * BD + a;
      const double pv = (double)p[o];
      x[o] = x[o] + alpha * pv;
      const VT rn = (VT)((double)r[o] - alpha * (double)q[o]);
      r[o] = rn;
      rv[a] = (double)rn;
    }
    apply_precond(pc, BD, rv, zv, dinv, sizeof(VT) == 4, i);
    for (int a = 0; a < BD; ++a) {
      const VT zz = (VT)zv[a];
      z[i * BD + a] = zz;
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
  if (last_block(cnt)) {
    double a1 = 0.0, a2 = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += 256) {
      a1 += __builtin_nontemporal_load(part + 2 * k);
      a2 += __builtin_nontemporal_load(part + 2 * k + 1);
    }
    a1 = wave_sum(a1);
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = a1;
    __syncthreads();
    const double rz = ((s4[0] + s4[1]) + s4[2]) + s4[3];
    __syncthreads();
    a2 = wave_sum(a2);
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = a2;
    __syncthreads();
    const double rr = ((s4[0] + s4[1]) + s4[2]) + s4[3];
    if (threadIdx.x == 0) {
      S->iter += 1;
      S->rz_old = S->rz;
      S->beta = rz / S->rz;
      S->rz = rz;
      S->rr = rr;
      if (!(rr < INFINITY) || !(rz == rz)) {
        S->status = SMH_ERR_BREAKDOWN;
        S->done = 1;
      } else if (stop_test(S, rr, rz)) {
        S->claimed = 1;  // verified against the FP64 residual by the host
        S->done = 1;
      } else if (S->iter >= S->maxiters) {
        S->done = 1;
      } else if (mixed && rr < S->delta2 * S->rr_ref) {
        S->want_ru = 1;
      }
      if (rr > S->rr_ref) S->rr_ref = rr;
      *cnt = 0;
    }
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

// r = b - A x in FP64 (A in precision RP), stored as VT; partial ||r||^2
template <int RP, int BD, typename VT>
__global__ void __launch_bounds__(256)
k_residual(BsrDev A, uint32_t per, const double *__restrict__ b, const double *__restrict__ x,
           VT *__restrict__ r, CGState *S, double *part, unsigned *cnt) {
  __shared__ double s4[4];
  const int lane = threadIdx.x & 63;
  uint32_t r0, r1;
  wave_rows(A.nrows, per, r0, r1);
  double d = 0.0;
  if (BD == 3) {
    for (uint32_t rw = r0; rw < r1; ++rw) {
      const double s = row3<RP>(A, x, rw, lane);
      if (lane == 0 || lane == 3 || lane == 6) {
        const uint64_t o = (uint64_t)rw * 3 + lane / 3;
        const double rv = b[o] - s;
        r[o] = (VT)rv;
        d += rv * rv;
      }
    }
  } else {
    const int hl = lane & 31, half = lane >> 5;
    for (uint32_t rw = r0; rw < r1; rw += 2) {
      const uint32_t rr = rw + half;
      const bool act = rr < r1;
      const double s = row1<RP>(A, x, rr, act, hl);
      if (act && hl == 0) {
        const double rv = b[rr] - s;
        r[rr] = (VT)rv;
        d += rv * rv;
      }
    }
  }
  const double t = block_sum256(d, s4);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
  if (last_block(cnt)) {
    const double rr = sum_partials(part, gridDim.x, s4);
    if (threadIdx.x == 0) {
      S->rr_true = rr;
      *cnt = 0;
    }
  }
}

// after a residual replacement: z = M r, z'r, r'r; reset the update reference;
// init=1 also sets p = z (start of the solve)
template <typename VT, int BD>
__global__ void __launch_bounds__(256)
k_cg_restart(uint32_t nrows, int pc, const void *dinv, const VT *__restrict__ r,
             VT *__restrict__ z, VT *__restrict__ p, int init, CGState *S, double *part,
             unsigned *cnt) {
  __shared__ double s4[4];
  double drz = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nrows;
       i += (uint64_t)gridDim.x * 256) {
    double rv[BD], zv[BD];
    for (int a = 0; a < BD; ++a) rv[a] = (double)r[i * BD + a];
    apply_precond(pc, BD, rv, zv, dinv, sizeof(VT) == 4, i);
    for (int a = 0; a < BD; ++a) {
      const VT zz = (VT)zv[a];
      z[i * BD + a] = zz;
      if (init) p[i * BD + a] = zz;
      drz += (double)zz * rv[a];
    }
  }
  const double t = block_sum256(drz, s4);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
  if (last_block(cnt)) {
    const double rz = sum_partials(part, gridDim.x, s4);
    if (threadIdx.x == 0) {
      S->rz = rz;
      S->rr = S->rr_true;
      S->rr_ref = S->rr_true;
      S->want_ru = 0;
      S->claimed = 0;
      if (!(S->rr_true < INFINITY)) {
        S->status = SMH_ERR_BREAKDOWN;
        S->done = 1;
      } else if (stop_test(S, S->rr_true, rz)) {
        S->done = 1;
      } else if (S->iter < S->maxiters && S->status == SMH_OK) {
        S->done = 0;
      }
      *cnt = 0;
    }
  }
}

// Jacobi / block-Jacobi inverse of the diagonal (thermal.sim:229-233 and
// inv3 of runtime.cpp:135-160), from the matrix in its best precision,
// plus an optional learned symmetric correction per block.
template <int RP, typename OT>
__global__ void k_dinv(BsrDev A, int bd, int pc, const float *learned, OT *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nrows) return;
  const int bb = bd * bd;
  double d[9];
  for (int q = 0; q < bb; ++q) d[q] = 0.0;
  for (uint32_t ij = A.rowptr[i]; ij < A.rowptr[i + 1]; ++ij)
    if (A.colidx[ij] == i) {
      for (int q = 0; q < bb; ++q) d[q] = Val<RP>::ld(A.vals, A.vals_lo, (uint64_t)ij * bb + q);
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
  for (int q = 0; q < 9; ++q) {
    double v = inv[q];
    if (pc == SMH_PRECOND_LEARNED && learned) v += (double)learned[(uint64_t)i * 9 + q];
    out[(uint64_t)i * 9 + q] = (OT)v;
  }
}

__global__ void k_cg_init_state(CGState *S, double bb, double tol, int rel, int norm_form,
                                int maxiters, int precond_stop, double delta) {
  if (threadIdx.x || blockIdx.x) return;
  memset(S, 0, sizeof(CGState));
  S->bb = bb;
  double thr = rel ? tol * tol * bb : tol;
  if (norm_form && !rel) thr = tol * tol;
  S->thr = thr;
  S->maxiters = maxiters;
  S->precond_stop = precond_stop;
  S->delta2 = delta * delta;
}

__global__ void k_fill_zero(double *p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    p[i] = 0.0;
}

// ---- host driver ---------------------------------------------------------------
struct Work {
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
  CGState *S = nullptr;
  double *part = nullptr;
  unsigned *cnt = nullptr;
  void release() {
    dfree(r); dfree(z); dfree(p); dfree(q); dfree(dinv); dfree(S); dfree(part); dfree(cnt);
  }
};

template <int IP, int RP, int BD>
int cg_run(smh_ctx *ctx, const smh_bsr *Am, const double *b, double *x, const smh_cg_opts *o,
           smh_cg_stats *st) {
  using VT = typename Val<IP>::T;
  const uint32_t nrows = Am->nrows;
  const uint64_t n = (uint64_t)nrows * BD;
  BsrDev Ain{nrows, Am->rowptr, Am->colidx, Am->vals, Am->vals_lo};
  BsrDev Ares = Ain;
  const Launch L = spmv_launch(nrows, BD);
  const unsigned nbe = grid_for(nrows, 256, 2048);
  const int pc = o->precond;
  const int mixed = o->mixed;
  Work w;
  int rc = SMH_OK;
  auto T0 = std::chrono::steady_clock::now();
#define TRY(x)            \
  do {                    \
    rc = (x);             \
    if (rc) goto done;    \
  } while (0)
  TRY(dmalloc(ctx, &w.r, sizeof(VT) * n));
  TRY(dmalloc(ctx, &w.z, sizeof(VT) * n));
  TRY(dmalloc(ctx, &w.p, sizeof(VT) * n));
  TRY(dmalloc(ctx, &w.q, sizeof(VT) * n));
  if (pc != SMH_PRECOND_NONE) TRY(dmalloc(ctx, &w.dinv, sizeof(VT) * (uint64_t)nrows * BD * BD));
  TRY(dmalloc(ctx, (void **)&w.S, sizeof(CGState)));
  TRY(dmalloc(ctx, (void **)&w.part, sizeof(double) * 2 * 4096));
  TRY(dmalloc(ctx, (void **)&w.cnt, sizeof(unsigned) * 8));
  hipMemsetAsync(w.cnt, 0, sizeof(unsigned) * 8, ctx->stream);
  {
    double bb = 0.0;
    TRY(smh_dot(ctx, SMH_F64, n, b, b, &bb));
    hipLaunchKernelGGL(k_cg_init_state, dim3(1), dim3(1), 0, ctx->stream, w.S, bb, o->tol, o->rel,
                       o->norm_form, o->maxiters,
                       (pc != SMH_PRECOND_NONE && o->zr_stop) ? 1 : 0,
                       o->reliable_delta);
    if (pc != SMH_PRECOND_NONE)
      hipLaunchKernelGGL((k_dinv<RP, VT>), dim3(grid_for(nrows, 256)), dim3(256), 0, ctx->stream,
                         Ares, BD, pc, o->learned, (VT *)w.dinv);
    // r = b - A x, z = M r, p = z
    hipLaunchKernelGGL((k_residual<RP, BD, VT>), dim3(L.blocks), dim3(256), 0, ctx->stream, Ares,
                       L.per, b, (const double *)x, (VT *)w.r, w.S, w.part, w.cnt);
    hipLaunchKernelGGL((k_cg_restart<VT, BD>), dim3(nbe), dim3(256), 0, ctx->stream, nrows, pc,
                       (const void *)w.dinv, (const VT *)w.r, (VT *)w.z, (VT *)w.p, 1, w.S,
                       w.part, w.cnt + 1);
    SMH_LAUNCH_CHECK(ctx);
    CGState hs{};
    const int check = o->check_every > 0 ? o->check_every : 16;
    int launched = 0, nru = 0;
    for (;;) {
      SMH_HIP(ctx, hipMemcpyAsync(&hs, w.S, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
      SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const bool need_ru =
          hs.status == SMH_OK && (hs.want_ru || hs.claimed ||
                                  (o->replace_every > 0 && hs.iter > 0 &&
                                   hs.iter / o->replace_every > nru &&
                                   hs.iter % o->replace_every == 0));
      if (need_ru) {
        ++nru;
        hipLaunchKernelGGL((k_residual<RP, BD, VT>), dim3(L.blocks), dim3(256), 0, ctx->stream,
                           Ares, L.per, b, (const double *)x, (VT *)w.r, w.S, w.part, w.cnt);
        hipLaunchKernelGGL((k_cg_restart<VT, BD>), dim3(nbe), dim3(256), 0, ctx->stream, nrows,
                           pc, (const void *)w.dinv, (const VT *)w.r, (VT *)w.z, (VT *)w.p, 0,
                           w.S, w.part, w.cnt + 1);
        SMH_LAUNCH_CHECK(ctx);
        continue;
      }
      if (hs.done || hs.status != SMH_OK) break;
      const int batch = o->replace_every > 0
                            ? std::min(check, o->replace_every - hs.iter % o->replace_every)
                            : check;
      for (int k = 0; k < batch; ++k) {
        hipLaunchKernelGGL((k_cg_spmv<IP, BD>), dim3(L.blocks), dim3(256), 0, ctx->stream, Ain,
                           L.per, (const VT *)w.p, (VT *)w.q, w.S, w.part, w.cnt + 2);
        hipLaunchKernelGGL((k_cg_update<VT, BD>), dim3(nbe), dim3(256), 0, ctx->stream, nrows, pc,
                           (const void *)w.dinv, x, (VT *)w.r, (VT *)w.z, (const VT *)w.p,
                           (const VT *)w.q, w.S, w.part + 4096, w.cnt + 3, mixed);
        hipLaunchKernelGGL((k_cg_pupd<VT>), dim3(grid_for(n, 256, 2048)), dim3(256), 0,
                           ctx->stream, n, (VT *)w.p, (const VT *)w.z, (const CGState *)w.S);
      }
      SMH_LAUNCH_CHECK(ctx);
      launched += batch;
    }
    (void)launched;
    // certify: FP64 residual of the final x
    hipLaunchKernelGGL((k_residual<RP, BD, VT>), dim3(L.blocks), dim3(256), 0, ctx->stream, Ares,
                       L.per, b, (const double *)x, (VT *)w.r, w.S, w.part, w.cnt);
    SMH_LAUNCH_CHECK(ctx);
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
      st->seconds =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - T0).count();
    }
  }
done:
#undef TRY
  hipStreamSynchronize(ctx->stream);
  w.release();
  return rc;
}

}  // namespace

extern "C" int smh_bsrmv(smh_ctx *ctx, const smh_bsr *A, const void *x, void *y,
                         int32_t exact_order) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && x && y && A->rowptr && (A->nnz == 0 || (A->colidx && A->vals)),
            SMH_ERR_INVALID, "smh_bsrmv: null argument");
  SMH_CHECK(ctx, A->bdim == 1 || A->bdim == 3 || (exact_order && A->bdim >= 1 && A->bdim <= 8),
            SMH_ERR_UNSUPPORTED,
            "smh_bsrmv: block size %u (1 or 3; 1..8 with exact_order)", A->bdim);
  SMH_CHECK(ctx, A->precision != SMH_F32X2 || A->vals_lo, SMH_ERR_INVALID,
            "smh_bsrmv: SMH_F32X2 needs vals_lo");
  if (A->nrows == 0) return SMH_OK;
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
  BsrDev D{A->nrows, A->rowptr, A->colidx, A->vals, A->vals_lo};
  const Launch L = spmv_launch(A->nrows, (int)A->bdim);
#define SMH_MV(P)                                                                             \
  do {                                                                                        \
    using T = typename Val<P>::T;                                                             \
    if (A->bdim == 3)                                                                         \
      hipLaunchKernelGGL(k_bsrmv3<P>, dim3(L.blocks), dim3(256), 0, ctx->stream, D, L.per,    \
                         (const T *)x, (T *)y);                                               \
    else                                                                                      \
      hipLaunchKernelGGL(k_bsrmv1<P>, dim3(L.blocks), dim3(256), 0, ctx->stream, D, L.per,    \
                         (const T *)x, (T *)y);                                               \
  } while (0)
  if (A->precision == SMH_F64) SMH_MV(SMH_F64);
  else if (A->precision == SMH_F32) SMH_MV(SMH_F32);
  else SMH_MV(SMH_F32X2);
#undef SMH_MV
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

extern "C" int smh_block_jacobi(smh_ctx *ctx, const smh_bsr *A, double *d_dinv) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && d_dinv && (A->bdim == 1 || A->bdim == 3), SMH_ERR_INVALID,
            "smh_block_jacobi: bad argument");
  BsrDev D{A->nrows, A->rowptr, A->colidx, A->vals, A->vals_lo};
  const int pc = A->bdim == 3 ? SMH_PRECOND_BLOCK_JACOBI : SMH_PRECOND_JACOBI;
  const unsigned g = grid_for(A->nrows, 256);
  if (A->precision == SMH_F64)
    hipLaunchKernelGGL((k_dinv<SMH_F64, double>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (int)A->bdim, pc, (const float *)nullptr, d_dinv);
  else if (A->precision == SMH_F32X2)
    hipLaunchKernelGGL((k_dinv<SMH_F32X2, double>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (int)A->bdim, pc, (const float *)nullptr, d_dinv);
  else
    hipLaunchKernelGGL((k_dinv<kF32D, double>), dim3(g), dim3(256), 0, ctx->stream, D,
                       (int)A->bdim, pc, (const float *)nullptr, d_dinv);
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
  o->reliable_delta = 0.1;
  o->check_every = 16;
}

extern "C" int smh_cg_solve(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x,
                            const smh_cg_opts *o, smh_cg_stats *st) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, A && b && x && o && A->rowptr, SMH_ERR_INVALID, "smh_cg_solve: null argument");
  SMH_CHECK(ctx, A->bdim == 1 || A->bdim == 3, SMH_ERR_UNSUPPORTED,
            "smh_cg_solve: block size %u (1 or 3 supported)", A->bdim);
  SMH_CHECK(ctx, o->precond >= 0 && o->precond <= 3, SMH_ERR_INVALID,
            "smh_cg_solve: unknown preconditioner %d", o->precond);
  SMH_CHECK(ctx, o->precond != SMH_PRECOND_LEARNED || (A->bdim == 3 && o->learned),
            SMH_ERR_INVALID, "smh_cg_solve: the learned preconditioner needs 3x3 blocks and "
                             "opts.learned");
  SMH_CHECK(ctx, A->bdim == 3 || o->precond <= SMH_PRECOND_JACOBI, SMH_ERR_INVALID,
            "smh_cg_solve: block Jacobi needs 3x3 blocks");
  SMH_CHECK(ctx, !o->mixed || A->precision == SMH_F32X2, SMH_ERR_INVALID,
            "smh_cg_solve: mixed precision needs an SMH_F32X2 matrix");
  SMH_CHECK(ctx, A->precision != SMH_F32X2 || A->vals_lo, SMH_ERR_INVALID,
            "smh_cg_solve: SMH_F32X2 needs vals_lo");
  SMH_CHECK(ctx, o->maxiters >= 0, SMH_ERR_INVALID, "smh_cg_solve: negative maxiters");
  if (A->nrows == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    return SMH_OK;
  }
  int rc;
  if (o->mixed) {
    rc = A->bdim == 3 ? cg_run<SMH_F32, SMH_F32X2, 3>(ctx, A, b, x, o, st)
                      : cg_run<SMH_F32, SMH_F32X2, 1>(ctx, A, b, x, o, st);
  } else if (A->precision == SMH_F64) {
    rc = A->bdim == 3 ? cg_run<SMH_F64, SMH_F64, 3>(ctx, A, b, x, o, st)
                      : cg_run<SMH_F64, SMH_F64, 1>(ctx, A, b, x, o, st);
  } else if (A->precision == SMH_F32X2) {
    rc = A->bdim == 3 ? cg_run<SMH_F32X2, SMH_F32X2, 3>(ctx, A, b, x, o, st)
                      : cg_run<SMH_F32X2, SMH_F32X2, 1>(ctx, A, b, x, o, st);
  } else {
    rc = A->bdim == 3 ? cg_run<kF32D, kF32D, 3>(ctx, A, b, x, o, st)
                      : cg_run<kF32D, kF32D, 1>(ctx, A, b, x, o, st);
  }
  return rc;
}
