// Blocked-CSR SpMV building blocks shared by smh_bsrmv and the CG kernels.
//
// Reference loop (src/lower/index_expressions/lower_indexexprs.cpp:630-667):
//   for i: tmp = 0; for ij in [rowptr[i], rowptr[i+1]): j = colidx[ij];
//          for bi, bj: tmp[bi] += A[ij*b*b + bi*b + bj] * x[j*b + bj]
// The reference GPU path ran one thread per row over that serial loop, with
// 16-thread blocks (gpu_backend.h:49-51).
//
// MI355X layout of the work (3x3 blocks): one 64-lane wavefront per block
// row; lane l < 63 has the fixed role (slot = l/9, e = l%9): it reads value
// e of blocks rowptr[i] + slot, + 7, + 14, ...  A row's values are
// contiguous, so each load instruction reads 63 consecutive values (fully
// coalesced), and no lane ever divides.  The 3 row sums come out of a
// 5-step shuffle tree on lanes 0, 3, 6.  1x1 blocks: a half-wave per row.
#pragma once
#include "smh_internal.h"

namespace smh {

template <int PREC> struct Val;
template <> struct Val<SMH_F64> {
  using T = double;
  __device__ static __forceinline__ double ld(const void *v, const void *, uint64_t i) {
    return __builtin_nontemporal_load(((const double *)v) + i);
  }
};
template <> struct Val<SMH_F32> {
  using T = float;
  __device__ static __forceinline__ float ld(const void *v, const void *, uint64_t i) {
    return __builtin_nontemporal_load(((const float *)v) + i);
  }
};
template <> struct Val<SMH_F32X2> {
  using T = double;
  __device__ static __forceinline__ double ld(const void *v, const void *lo, uint64_t i) {
    return (double)__builtin_nontemporal_load(((const float *)v) + i) +
           (double)__builtin_nontemporal_load(((const float *)lo) + i);
  }
};

struct BsrDev {
  uint32_t nrows;
  const uint32_t *rowptr;
  const uint32_t *colidx;
  const void *vals;
  const void *vals_lo;
};

__device__ __forceinline__ double shfl_down_zero(double v, int d) {
  const int lane = threadIdx.x & 63;
  const int src = lane + d;
  const double o = __shfl(v, src & 63, 64);
  return src < 64 ? o : 0.0;
}
__device__ __forceinline__ float shfl_down_zero(float v, int d) {
  const int lane = threadIdx.x & 63;
  const int src = lane + d;
  const float o = __shfl(v, src & 63, 64);
  return src < 64 ? o : 0.0f;
}

// One 3x3 block row with the whole wavefront.  Returns the partial sum; the
// row sums y[bi] end up on lanes 3*bi (bi = 0, 1, 2).
template <int PREC>
__device__ __forceinline__ typename Val<PREC>::T
row3(const BsrDev &A, const typename Val<PREC>::T *__restrict__ x, uint32_t r, int lane) {
  using T = typename Val<PREC>::T;
  const uint32_t rb = A.rowptr[r], re = A.rowptr[r + 1];
  const int slot = lane / 9, e = lane - 9 * slot, bj = e % 3;
  T acc = 0;
  if (lane < 63) {
    for (uint32_t t = rb + slot; t < re; t += 7) {
      const T v = Val<PREC>::ld(A.vals, A.vals_lo, (uint64_t)t * 9 + e);
      const uint32_t j = A.colidx[t];
      acc += v * x[(uint64_t)j * 3 + bj];
    }
  }
  acc += shfl_down_zero(acc, 1) + shfl_down_zero(acc, 2);
  acc += shfl_down_zero(acc, 9);
  acc += shfl_down_zero(acc, 18);
  acc += shfl_down_zero(acc, 36);
  return acc;
}

// One 1x1 row per half-wave; the sum ends up on lane 0 of the half.
template <int PREC>
__device__ __forceinline__ typename Val<PREC>::T
row1(const BsrDev &A, const typename Val<PREC>::T *__restrict__ x, uint32_t r, bool active,
     int hl) {
  using T = typename Val<PREC>::T;
  T acc = 0;
  if (active) {
    const uint32_t rb = A.rowptr[r], re = A.rowptr[r + 1];
    for (uint32_t t = rb + hl; t < re; t += 32) {
      const T v = Val<PREC>::ld(A.vals, A.vals_lo, t);
      acc += v * x[A.colidx[t]];
    }
  }
  for (int off = 16; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  return acc;
}

// Contiguous row range of wavefront w out of W, with the workgroup order
// remapped so that each XCD (blocks are dealt round-robin over 8 XCDs)
// works on one contiguous eighth of the rows: its L2 then holds the x
// entries its rows gather.
__device__ __forceinline__ void wave_rows(uint32_t nrows, uint32_t per, uint32_t &r0,
                                          uint32_t &r1) {
  const uint32_t nb = gridDim.x;
  uint32_t lb = blockIdx.x;
  if ((nb & 7) == 0) lb = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
  const uint32_t w = lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t a = (uint64_t)w * per, b = a + per;
  r0 = (uint32_t)min<uint64_t>(a, nrows);
  r1 = (uint32_t)min<uint64_t>(b, nrows);
}

// block-wide FP64 sum (256 threads), fixed order; result valid on thread 0
__device__ __forceinline__ double block_sum256(double v, double *s4) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = ((s4[0] + s4[1]) + s4[2]) + s4[3];
  __syncthreads();
  return t;
}

}  // namespace smh
