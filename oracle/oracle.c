/*
 * CPU oracle: single-threaded FP64 restatement of the reference's hot-path
 * loop nests.  TEST INFRASTRUCTURE ONLY -- see oracle.h for who may use it
 * and what pins each function.
 *
 * The `omp` pragmas are inert in liboracle.so (built without -fopenmp: the
 * checker, strictly sequential).  liboracle_omp.so is the same source built
 * with -fopenmp, used only as the multi-core CPU baseline of bench.py: its
 * SpMV rows and vector updates are the same per element, its dot products
 * are parallel reductions (a different summation order).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Sparsity                                                                  */
/* ------------------------------------------------------------------------ */

/* path_indices.cpp:248-276 (Link::ve) + pack(sorted) :172-203.  Elements are
 * visited in ascending id, so each list comes out ascending; an element that
 * holds a vertex twice lists it twice, as push_back does there. */
int64_t or_index_ve(int32_t nv, int64_t ne, int32_t card, const int32_t *ep,
                    uint32_t *ptr, uint32_t *idx) {
  memset(ptr, 0, sizeof(uint32_t) * ((size_t)nv + 1));
  for (int64_t e = 0; e < ne; ++e)
    for (int32_t s = 0; s < card; ++s) ptr[ep[e * card + s] + 1]++;
  for (int32_t v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
  uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nv + 1));
  memcpy(fill, ptr, sizeof(uint32_t) * ((size_t)nv + 1));
  for (int64_t e = 0; e < ne; ++e)
    for (int32_t s = 0; s < card; ++s) idx[fill[ep[e * card + s]]++] = (uint32_t)e;
  free(fill);
  return (int64_t)ptr[nv];
}

/* path_indices.cpp:207-247 (Link::ev): fixed stride, endpoint slot order. */
void or_index_ev(int64_t ne, int32_t card, const int32_t *ep, uint32_t *ptr,
                 uint32_t *idx) {
  for (int64_t e = 0; e <= ne; ++e) ptr[e] = (uint32_t)(e * card);
  for (int64_t e = 0; e < ne * card; ++e) idx[e] = (uint32_t)ep[e];
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

/* path_indices.cpp:359-429: for each source vertex, the set of sinks reached
 * through any element it belongs to (std::set => sorted, unique), then
 * pack(sorted=true).  Several edge sets are united as Or does (:432-471). */
int64_t or_index_vev(int32_t nv, int32_t nsets, const int64_t *ne,
                     const int32_t *card, const int32_t *const *ep,
                     int32_t force_diag, uint32_t *rowptr, uint32_t *colidx,
                     int64_t cap) {
  uint32_t **vptr = (uint32_t **)calloc(nsets, sizeof(uint32_t *));
  uint32_t **vidx = (uint32_t **)calloc(nsets, sizeof(uint32_t *));
  for (int32_t s = 0; s < nsets; ++s) {
    vptr[s] = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nv + 1));
    vidx[s] = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)ne[s] * card[s] + 1));
    or_index_ve(nv, ne[s], card[s], ep[s], vptr[s], vidx[s]);
  }
  size_t scap = 64;
  uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * scap);
  int64_t nnz = 0;
  rowptr[0] = 0;
  for (int32_t v = 0; v < nv; ++v) {
    size_t n = 0;
    for (int32_t s = 0; s < nsets; ++s) {
      for (uint32_t q = vptr[s][v]; q < vptr[s][v + 1]; ++q) {
        uint32_t e = vidx[s][q];
        for (int32_t t = 0; t < card[s]; ++t) {
          if (n == scap) {
            scap *= 2;
            scratch = (uint32_t *)realloc(scratch, sizeof(uint32_t) * scap);
          }
          scratch[n++] = (uint32_t)ep[s][(int64_t)e * card[s] + t];
        }
      }
    }
    if (force_diag) {
      if (n == scap) {
        scap *= 2;
        scratch = (uint32_t *)realloc(scratch, sizeof(uint32_t) * scap);
      }
      scratch[n++] = (uint32_t)v;
    }
    qsort(scratch, n, sizeof(uint32_t), cmp_u32);
    size_t u = 0;
    for (size_t q = 0; q < n; ++q)
      if (u == 0 || scratch[q] != scratch[u - 1]) scratch[u++] = scratch[q];
    if (nnz + (int64_t)u > cap) {
      nnz = -1;
      break;
    }
    memcpy(colidx + nnz, scratch, u * sizeof(uint32_t));
    nnz += (int64_t)u;
    rowptr[v + 1] = (uint32_t)nnz;
  }
  free(scratch);
  for (int32_t s = 0; s < nsets; ++s) {
    free(vptr[s]);
    free(vidx[s]);
  }
  free(vptr);
  free(vidx);
  return nnz;
}

/* runtime.cpp:19-23 */
int32_t or_loc(const uint32_t *rowptr, const uint32_t *colidx, int32_t i,
               int32_t j) {
  uint32_t l = rowptr[i];
  while (colidx[l] != (uint32_t)j) l++;
  return (int32_t)l;
}

/* ------------------------------------------------------------------------ */
/* 3x3 determinant / inverse, runtime.cpp:91-95 and :135-160 (textbook      */
/* cofactor expansion, same association as written there)                  */
/* ------------------------------------------------------------------------ */
double or_det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

void or_inv3(const double *a, double *inv) {
  double cof00 = a[4] * a[8] - a[5] * a[7];
  double cof01 = -a[3] * a[8] + a[5] * a[6];
  double cof02 = a[3] * a[7] - a[4] * a[6];
  double cof10 = -a[1] * a[8] + a[2] * a[7];
  double cof11 = a[0] * a[8] - a[2] * a[6];
  double cof12 = -a[0] * a[7] + a[1] * a[6];
  double cof20 = a[1] * a[5] - a[2] * a[4];
  double cof21 = -a[0] * a[5] + a[2] * a[3];
  double cof22 = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * cof00 + a[1] * cof01 + a[2] * cof02;
  d = 1.0 / d;
  inv[0] = cof00 * d; inv[1] = cof10 * d; inv[2] = cof20 * d;
  inv[3] = cof01 * d; inv[4] = cof11 * d; inv[5] = cof21 * d;
  inv[6] = cof02 * d; inv[7] = cof12 * d; inv[8] = cof22 * d;
}

/* ------------------------------------------------------------------------ */
/* Springs: test/input/program/isprings.sim:19-69                            */
/* ------------------------------------------------------------------------ */

/* compute_elasticity (isprings.sim:38-69) for one spring: fills fe0 (force on
 * endpoint 0) and the 3x3 block k.  Association follows the source text
 * left to right. */
static void spring_terms(const or_springs_params *P, const double *x0,
                         const double *x1, double stiffness, double l0,
                         double fe0[3], double kb[9]) {
  const double h = P->h;
  double dx[3] = {x1[0] - x0[0], x1[1] - x0[1], x1[2] - x0[2]};
  double dxtdx = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
  double l = sqrt(dxtdx);
  double s = h * stiffness / (l0 * l0) * (l - l0);
  for (int q = 0; q < 3; ++q) fe0[q] = s * dx[q] / l;
  double c = h * h * stiffness / (l0 * l0 * l * l);
  double r = (l - l0) / l;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double o = dx[i] * dx[j];
      double d = (i == j ? dxtdx : 0.0) - o;
      kb[3 * i + j] = c * (o + r * d);
    }
}

void or_springs_assemble(const or_springs_params *P, int32_t np, int64_t ns,
                         const int32_t *ep, const double *x, const double *v,
                         const double *m, const uint8_t *fixed,
                         const double *k, const double *l0,
                         const uint32_t *rowptr, const uint32_t *colidx,
                         double *vals, double *f) {
  const int64_t nnz = rowptr[np];
  /* K and fe start at zero (inline.cpp:878-885); MD and fmg are separate
   * temporaries, added at the end: MDK = MD + K, f = fmg + fe. */
  double *md = (double *)calloc((size_t)np * 9, sizeof(double));
  double *fmg = (double *)calloc((size_t)np * 3, sizeof(double));
  memset(vals, 0, sizeof(double) * 9 * nnz);
  memset(f, 0, sizeof(double) * 3 * np);
  /* compute_mass_damping (isprings.sim:19-36), one point at a time */
  for (int32_t p = 0; p < np; ++p) {
    double mass = m[p];
    if (fixed[p]) {
      for (int q = 0; q < 9; ++q)
        md[9 * p + q] = (P->fixed_identity && (q % 4 != 0)) ? 0.0 : P->fixed_fill;
    } else {
      double dm = mass + P->h * P->viscous;
      for (int q = 0; q < 9; ++q) md[9 * p + q] = (q % 4 == 0) ? dm : 0.0;
      for (int q = 0; q < 3; ++q)
        fmg[3 * p + q] = mass * v[3 * p + q] + P->h * mass * P->grav[q];
    }
  }
  /* compute_elasticity over springs in ascending order (inline.cpp:821-888);
   * each write is A[loc] += (lower_maps.cpp:42) with loc from runtime.cpp:19 */
  for (int64_t s = 0; s < ns; ++s) {
    int32_t a = ep[2 * s], b = ep[2 * s + 1];
    int free0 = !fixed[a], free1 = !fixed[b];
    double fe0[3], kb[9];
    spring_terms(P, x + 3 * a, x + 3 * b, k[s], l0[s], fe0, kb);
    if (free0)
      for (int q = 0; q < 3; ++q) f[3 * a + q] += fe0[q];
    if (free1)
      for (int q = 0; q < 3; ++q) f[3 * b + q] += -fe0[q];
    if (free0 && free1) {
      int32_t laa = or_loc(rowptr, colidx, a, a), lab = or_loc(rowptr, colidx, a, b);
      int32_t lba = or_loc(rowptr, colidx, b, a), lbb = or_loc(rowptr, colidx, b, b);
      for (int q = 0; q < 9; ++q) {
        vals[9 * (int64_t)laa + q] += kb[q];
        vals[9 * (int64_t)lab + q] += -kb[q];
        vals[9 * (int64_t)lba + q] += -kb[q];
        vals[9 * (int64_t)lbb + q] += kb[q];
      }
    } else {
      if (free0) {
        int32_t l = or_loc(rowptr, colidx, a, a);
        for (int q = 0; q < 9; ++q) vals[9 * (int64_t)l + q] += kb[q];
      }
      if (free1) {
        int32_t l = or_loc(rowptr, colidx, b, b);
        for (int q = 0; q < 9; ++q) vals[9 * (int64_t)l + q] += kb[q];
      }
    }
  }
  /* MDK = MD + K: the diagonal storage is added onto the indexed storage
   * (lower_indexexprs.cpp:545-579); f = fmg + fe. */
  for (int32_t p = 0; p < np; ++p) {
    for (int q = 0; q < 3; ++q) f[3 * p + q] = fmg[3 * p + q] + f[3 * p + q];
    if (rowptr[p] == rowptr[p + 1]) continue; /* isolated point: no block */
    int32_t l = or_loc(rowptr, colidx, p, p);
    for (int q = 0; q < 9; ++q) vals[9 * (int64_t)l + q] = md[9 * p + q] + vals[9 * (int64_t)l + q];
  }
  free(md);
  free(fmg);
}

/* ------------------------------------------------------------------------ */
/* Linear tets: apps/fem/fem_linear.sim                                      */
/* ------------------------------------------------------------------------ */

/* precomputeTetMat (fem_linear.sim:176-194): M(jj,ii) = x_ii - x_3, B = inv(M),
 * W = -(1/6) det(M).  dN rows 0..2 = B rows, row 3 = -(sum) (:118-124). */
static void tet_geom(const double *xs, double dN[12], double *W) {
  double M[9], B[9];
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) M[3 * jj + ii] = xs[3 * ii + jj] - xs[9 + jj];
  or_inv3(M, B);
  *W = -(1.0 / 6.0) * or_det3(M);
  for (int di = 0; di < 3; ++di) dN[9 + di] = 0.0;
  for (int vi = 0; vi < 3; ++vi)
    for (int di = 0; di < 3; ++di) {
      dN[3 * vi + di] = B[3 * vi + di];
      dN[9 + di] = dN[9 + di] - B[3 * vi + di];
    }
}

/* Block (a,b) of s * Ke, Ke = W * B6' * E * B6 (fem_linear.sim:95-174), in
 * closed form.  With u = (s*W*lam) g_a and w = (s*W*mu) g_a:
 *   blk[i][j] = u_i gb_j + w_j gb_i + [i==j] (w . gb)
 * (the order of operations below is the one the device kernel uses). */
static void tet_block_from(const double dN[12], double W, double mu, double lam,
                           double s, int a, int b, double *blk) {
  const double *ga = dN + 3 * a, *gb = dN + 3 * b;
  const double sw = s * W, cl = sw * lam, cm = sw * mu;
  double u[3], w[3];
  for (int i = 0; i < 3; ++i) {
    u[i] = cl * ga[i];
    w[i] = cm * ga[i];
  }
  const double wg = w[0] * gb[0] + w[1] * gb[1] + w[2] * gb[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = u[i] * gb[j] + w[j] * gb[i];
      if (i == j) t = t + wg;
      blk[3 * i + j] = t;
    }
}

void or_tet_block(const double *xs, double mu, double lam, int a, int b,
                  double *blk, double *W) {
  double dN[12];
  tet_geom(xs, dN, W);
  tet_block_from(dN, *W, mu, lam, 1.0, a, b, blk);
}

void or_tet_assemble(const or_tet_params *P, int32_t nv, int64_t nt,
                     const int32_t *ep, const double *x, const int32_t *c,
                     const double *mu, const double *lam,
                     const uint32_t *rowptr, const uint32_t *colidx,
                     double *vals, double *f) {
  const int64_t nnz = rowptr[nv];
  memset(vals, 0, sizeof(double) * 9 * nnz);
  memset(f, 0, sizeof(double) * 3 * nv);
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t *v = ep + 4 * t;
    double xs[12], dN[12], W, blk[9];
    for (int a = 0; a < 4; ++a)
      for (int q = 0; q < 3; ++q) xs[3 * a + q] = x[3 * v[a] + q];
    tet_geom(xs, dN, &W);
    /* compute_stiffness assembly loop (fem_linear.sim:158-173): stiffness
     * blocks for free pairs, then the lumped mass on each diagonal */
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        if (c[v[a]] <= 0 && c[v[b]] <= 0) {
          tet_block_from(dN, W, mu[t], lam[t], P->stiff_scale, a, b, blk);
          int32_t l = or_loc(rowptr, colidx, v[a], v[b]);
          for (int q = 0; q < 9; ++q) vals[9 * (int64_t)l + q] += blk[q];
        }
      }
    if (P->rho != 0.0) {
      double mm = 0.25 * P->rho * W;
      for (int a = 0; a < 4; ++a) {
        int32_t l = or_loc(rowptr, colidx, v[a], v[a]);
        for (int q = 0; q < 9; ++q) vals[9 * (int64_t)l + q] += (q % 4 == 0) ? mm : 0.0;
      }
    }
    /* body load, compute_force's fg = m*grav term with m = W/4 times a
     * force density (fem_linear.sim:48-81) */
    double wq = 0.25 * W;
    for (int a = 0; a < 4; ++a)
      if (c[v[a]] <= 0)
        for (int q = 0; q < 3; ++q) f[3 * v[a] + q] += wq * P->load[q];
  }
  /* constrained rows: identity diagonal map over the vertex set, added as
   * MD + K is in isprings.sim:74 */
  if (P->fixed_diag != 0.0)
    for (int32_t p = 0; p < nv; ++p) {
      if (c[p] <= 0 || rowptr[p] == rowptr[p + 1]) continue;
      int32_t l = or_loc(rowptr, colidx, p, p);
      for (int q = 0; q < 9; ++q)
        vals[9 * (int64_t)l + q] = ((q % 4 == 0) ? P->fixed_diag : 0.0) + vals[9 * (int64_t)l + q];
    }
}

/* ------------------------------------------------------------------------ */
/* Neo-Hookean tets: test/input/program/femTet.sim                           */
/* ------------------------------------------------------------------------ */

/* Natural logarithm from x = m 2^k, m in [sqrt(1/2), sqrt(2)), and
 * log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, summed as the
 * series 2 s (1 + s^2/3 + s^4/5 + ... + s^24/25) by Horner's rule: a
 * textbook expansion, written so that the device kernels can repeat it
 * operation for operation (the femTet element's log(det(F)), femTet.sim:
 * 28, 36).  Pinned against libm's log in the tests (<= 2 ulp). */
static const double kLn2 = 0.69314718055994530942;
double or_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int k = 0;
  double m = frexp(x, &k);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    k = k - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double t = 1.0 / 25.0;
  for (int n = 11; n >= 0; --n) t = t * s2 + 1.0 / (double)(2 * n + 1);
  return (double)k * kLn2 + 2.0 * s * t;
}

/* 3x3 helpers in the association of the .sim expressions: C = A*B sums
 * k ascending; C = A*B' uses B's rows. */
static void mm3(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static void mm3t(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] =
          A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}

/* precomputeTetMat (femTet.sim:126-144) for one tet: B = 1.0*inv(M),
 * W = -(1/6) det(M), M(jj,ii) = x_ii(jj) - x_3(jj). */
void or_femtet_geom(const double *xs, double *B, double *W) {
  double M[9], Mi[9];
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) M[3 * jj + ii] = xs[3 * ii + jj] - xs[9 + jj];
  or_inv3(M, Mi);
  for (int q = 0; q < 9; ++q) B[q] = 1.0 * Mi[q];
  *W = -(1.0 / 6.0) * or_det3(M);
}

/* initializeTet (femTet.sim:146-148): verts.m = map precomputeTetMat to
 * tets, a map without a reduction: each write assigns, so a vertex keeps
 * the value of the last (highest) tet that holds it; m of a vertex in no
 * tet is left as it was. */
void or_femtet_precompute(double rho, int32_t nv, int64_t nt, const int32_t *ep,
                          const double *x, double *W, double *B, double *m) {
  (void)nv;
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t *v = ep + 4 * t;
    double xs[12];
    for (int a = 0; a < 4; ++a)
      for (int q = 0; q < 3; ++q) xs[3 * a + q] = x[3 * (int64_t)v[a] + q];
    or_femtet_geom(xs, B + 9 * t, W + t);
    for (int ii = 0; ii < 4; ++ii) m[v[ii]] = 0.25 * rho * W[t];
  }
}

/* F = Ds*B with Ds(jj,ii) = x_ii(jj) - x_3(jj) (femTet.sim:49-54, 80-85) */
static void femtet_F(const double *xs, const double *B, double *F) {
  double Ds[9];
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) Ds[3 * jj + ii] = xs[3 * ii + jj] - xs[9 + jj];
  mm3(Ds, B, F);
}

/* The element's internal force columns H(:,ii) (H = -W * PK1(F) * B',
 * femTet.sim:27-31, 56-57), H row-major. */
void or_femtet_H(const double *xs, double u, double l, double W, const double *B, double *H) {
  double F[9], Fi[9], FiT[9], P[9], Ps[9];
  femtet_F(xs, B, F);
  const double JJ = or_log(or_det3(F));
  or_inv3(F, Fi);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FiT[3 * i + j] = Fi[3 * j + i];
  const double lJ = l * JJ;
  for (int q = 0; q < 9; ++q) P[q] = u * (F[q] - FiT[q]) + lJ * FiT[q];
  const double nW = -W;
  for (int q = 0; q < 9; ++q) Ps[q] = nW * P[q];
  mm3t(Ps, B, H);
}

/* dH = -W * dPdF(F, dF) * B' (femTet.sim:33-39, 103-104) for the dF whose
 * only non-zero row kk is dFRow(row, :) */
static void femtet_dH(const double *F, const double *Fi, const double *FiT, double JJ, double u,
                      double l, double W, const double *B, const double *dFrow, int kk,
                      double *dH) {
  double dF[9], FidF[9], T1[9], T2[9], dP[9], dPs[9];
  for (int q = 0; q < 9; ++q) dF[q] = 0.0;
  for (int ll = 0; ll < 3; ++ll) dF[3 * kk + ll] = dFrow[ll];
  mm3(Fi, dF, FidF);
  const double s1 = u - l * JJ;
  for (int q = 0; q < 9; ++q) T1[q] = s1 * FiT[q];
  mm3t(T1, FidF, T2);
  const double s2 = l * (FidF[0] + FidF[4] + FidF[8]);
  for (int q = 0; q < 9; ++q) dP[q] = u * dF[q] + T2[q] + s2 * FiT[q];
  const double nW = -W;
  for (int q = 0; q < 9; ++q) dPs[q] = nW * dP[q];
  mm3t(dPs, B, dH);
  (void)F;
}

/* compute_stiffness (femTet.sim:74-124): block K(v(jj), v(row)) of one tet
 * without the h*h scaling and sign: Kb(jj,:,:) for dF directions
 * (row, kk = 0..2), out[ll*3 + kk]. */
void or_femtet_kb(const double *xs, double u, double l, double W, const double *B, int jj,
                  int row, double *out) {
  double F[9], Fi[9], FiT[9], dFrow[3], dH[9];
  femtet_F(xs, B, F);
  const double JJ = or_log(or_det3(F));
  or_inv3(F, Fi);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FiT[3 * i + j] = Fi[3 * j + i];
  if (row < 3) {
    for (int ll = 0; ll < 3; ++ll) dFrow[ll] = B[3 * row + ll];
  } else {
    for (int ii = 0; ii < 3; ++ii) dFrow[ii] = -(B[ii] + B[3 + ii] + B[6 + ii]);
  }
  for (int kk = 0; kk < 3; ++kk) {
    femtet_dH(F, Fi, FiT, JJ, u, l, W, B, dFrow, kk, dH);
    for (int ll = 0; ll < 3; ++ll)
      out[3 * ll + kk] = jj < 3 ? dH[3 * ll + jj]
                                : -(dH[3 * ll + 0] + dH[3 * ll + 1] + dH[3 * ll + 2]);
  }
}

/* main's two maps (femTet.sim:150-152): f = map compute_force(h) to tets
 * reduce +, A = map compute_stiffness(h) to tets reduce +, every element
 * write a += in ascending element order, in the order the element function
 * writes.  A is 3x3 blocked CSR on rowptr/colidx. */
void or_femtet_assemble(double h, double rho, const double *grav, int32_t nv, int64_t nt,
                        const int32_t *ep, const double *x, const double *v, const int32_t *c,
                        const double *u, const double *l, const double *W, const double *B,
                        const uint32_t *rowptr, const uint32_t *colidx, double *vals,
                        double *f) {
  const int64_t nnz = rowptr[nv];
  memset(vals, 0, sizeof(double) * 9 * nnz);
  memset(f, 0, sizeof(double) * 3 * nv);
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t *vt = ep + 4 * t;
    const double *Bt = B + 9 * t;
    double xs[12];
    for (int a = 0; a < 4; ++a)
      for (int q = 0; q < 3; ++q) xs[3 * a + q] = x[3 * (int64_t)vt[a] + q];
    /* compute_force */
    const double mq = 0.25 * rho * W[t];
    double fg[3], H[9];
    for (int q = 0; q < 3; ++q) fg[q] = mq * grav[q];
    or_femtet_H(xs, u[t], l[t], W[t], Bt, H);
    for (int ii = 0; ii < 3; ++ii) {
      if (c[vt[ii]] <= 0)
        for (int q = 0; q < 3; ++q) f[3 * (int64_t)vt[ii] + q] += h * H[3 * q + ii];
      if (c[vt[3]] <= 0)
        for (int q = 0; q < 3; ++q) f[3 * (int64_t)vt[3] + q] += -h * H[3 * q + ii];
    }
    for (int ii = 0; ii < 4; ++ii)
      if (c[vt[ii]] <= 0)
        for (int q = 0; q < 3; ++q)
          f[3 * (int64_t)vt[ii] + q] += h * fg[q] + mq * v[3 * (int64_t)vt[ii] + q];
    /* compute_stiffness: stiffness of free pairs (loops row, then jj), then
     * the lumped mass on every diagonal */
    const double hh = h * h;
    for (int row = 0; row < 4; ++row)
      for (int jj = 0; jj < 4; ++jj) {
        if (!(c[vt[jj]] <= 0 && c[vt[row]] <= 0)) continue;
        double kb[9];
        or_femtet_kb(xs, u[t], l[t], W[t], Bt, jj, row, kb);
        const int32_t loc = or_loc(rowptr, colidx, vt[jj], vt[row]);
        for (int q = 0; q < 9; ++q) vals[9 * (int64_t)loc + q] += -(hh * kb[q]);
      }
    for (int ii = 0; ii < 4; ++ii) {
      const int32_t loc = or_loc(rowptr, colidx, vt[ii], vt[ii]);
      for (int q = 0; q < 9; ++q) vals[9 * (int64_t)loc + q] += mq * ((q % 4 == 0) ? 1.0 : 0.0);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Trilinear hex, scalar Poisson                                             */
/* ------------------------------------------------------------------------ */

/* Unit-cube trilinear Laplacian stiffness, corner a = dx + 2 dy + 4 dz:
 * 1/3 on the diagonal, 0 for edge neighbours, -1/12 for face and body
 * diagonals (pinned against 2x2x2 Gauss quadrature in the tests). */
double or_hex_kref(int a, int b) {
  int d = __builtin_popcount((unsigned)(a ^ b));
  return d == 0 ? (1.0 / 3.0) : (d == 1 ? 0.0 : -(1.0 / 12.0));
}

void or_hex_assemble(const or_hex_params *P, int32_t nv, int64_t nh,
                     const int32_t *ep, const int32_t *dirichlet,
                     const double *kappa, const uint32_t *rowptr,
                     const uint32_t *colidx, double *vals, double *f) {
  const int64_t nnz = rowptr[nv];
  memset(vals, 0, sizeof(double) * nnz);
  memset(f, 0, sizeof(double) * nv);
  const double hs = P->h;
  const double src = P->source * (hs * hs * hs) * 0.125;
  for (int64_t e = 0; e < nh; ++e) {
    const int32_t *v = ep + 8 * e;
    double ke = kappa[e] * hs;
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        if (dirichlet[v[a]] <= 0 && dirichlet[v[b]] <= 0) {
          int32_t l = or_loc(rowptr, colidx, v[a], v[b]);
          vals[l] += ke * or_hex_kref(a, b);
        }
    for (int a = 0; a < 8; ++a)
      if (dirichlet[v[a]] <= 0) f[v[a]] += src;
  }
  for (int32_t p = 0; p < nv; ++p) {
    if (dirichlet[p] <= 0 || rowptr[p] == rowptr[p + 1]) continue;
    int32_t l = or_loc(rowptr, colidx, p, p);
    vals[l] = P->fixed_diag + vals[l];
  }
}

/* ------------------------------------------------------------------------ */
/* Linear algebra                                                            */
/* ------------------------------------------------------------------------ */

/* lower_indexexprs.cpp:630-667: row loop, ascending ij then bj, one
 * accumulator per block-row component, zero-initialised. */
void or_bsrmv(int32_t nrows, int32_t bdim, const uint32_t *rowptr,
              const uint32_t *colidx, const double *vals, const double *x,
              double *y) {
  const int bb = bdim * bdim;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < nrows; ++i) {
    double tmp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t ij = rowptr[i]; ij < rowptr[i + 1]; ++ij) {
      uint32_t j = colidx[ij];
      for (int bi = 0; bi < bdim; ++bi)
        for (int bj = 0; bj < bdim; ++bj)
          tmp[bi] += vals[(int64_t)ij * bb + bi * bdim + bj] * x[(int64_t)j * bdim + bj];
    }
    for (int bi = 0; bi < bdim; ++bi) y[(int64_t)i * bdim + bi] = tmp[bi];
  }
}

/* flatten.cpp:286-306 -> ir_builder.cpp:130-140: sequential inner product */
double or_dot(int64_t n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* ------------------------------------------------------------------------ */
/* CG                                                                        */
/* ------------------------------------------------------------------------ */

static void precond_apply(int32_t nrows, int32_t bdim, int32_t kind,
                          const double *dinv, const double *r, double *z) {
  const int64_t n = (int64_t)nrows * bdim;
  if (kind == 0) {
    memcpy(z, r, sizeof(double) * n);
  } else if (kind == 1) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
  } else {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < nrows; ++i)
      for (int a = 0; a < 3; ++a) {
        double s = 0.0;
        for (int b = 0; b < 3; ++b) s += dinv[9 * (int64_t)i + 3 * a + b] * r[3 * (int64_t)i + b];
        z[3 * (int64_t)i + a] = s;
      }
  }
}

/* Unpreconditioned form follows isprings.sim:84-103 / fem_linear.sim:211-230
 * (cg.sim:25-53 with norm_form); the relative/replacement/Jacobi options
 * follow thermal.sim:75-130 (Diag = 1/A(q,q) at :229-233). */
void or_cg(int32_t nrows, int32_t bdim, const uint32_t *rowptr,
           const uint32_t *colidx, const double *vals, const double *b,
           double *x, const or_cg_opts *o, or_cg_stats *st) {
  const int64_t n = (int64_t)nrows * bdim;
  const int bb = bdim * bdim;
  double *r = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n);
  double *Ap = (double *)malloc(sizeof(double) * n);
  double *dinv = NULL;
  if (o->precond == 1) {
    dinv = (double *)calloc(n, sizeof(double));
    for (int32_t i = 0; i < nrows; ++i)
      for (uint32_t ij = rowptr[i]; ij < rowptr[i + 1]; ++ij)
        if (colidx[ij] == (uint32_t)i)
          for (int a = 0; a < bdim; ++a)
            dinv[(int64_t)i * bdim + a] = 1.0 / vals[(int64_t)ij * bb + a * bdim + a];
  } else if (o->precond == 2) {
    dinv = (double *)calloc((size_t)n * 3, sizeof(double));
    for (int32_t i = 0; i < nrows; ++i)
      for (uint32_t ij = rowptr[i]; ij < rowptr[i + 1]; ++ij)
        if (colidx[ij] == (uint32_t)i) or_inv3(vals + (int64_t)ij * 9, dinv + 9 * (int64_t)i);
  }
  or_bsrmv(nrows, bdim, rowptr, colidx, vals, x, Ap);
  for (int64_t q = 0; q < n; ++q) r[q] = b[q] - Ap[q];
  precond_apply(nrows, bdim, o->precond, dinv, r, z);
  memcpy(p, z, sizeof(double) * n);
  double rz = or_dot(n, z, r);
  double rr = or_dot(n, r, r);
  double thr = o->tol;
  if (o->rel) thr = o->tol * o->tol * or_dot(n, b, b);
  int32_t iter = 0;
  double nr = o->norm_form ? sqrt(rr) : rr;
  while (nr > thr && (o->precond == 0 || rz > thr) && iter < o->maxiters) {
    or_bsrmv(nrows, bdim, rowptr, colidx, vals, p, Ap);
    double denom = or_dot(n, p, Ap);
    double alpha = (o->norm_form ? or_dot(n, r, r) : rz) / denom;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) x[q] = x[q] + alpha * p[q];
    iter = iter + 1;
    if (o->replace_every > 0 && iter % o->replace_every == 0) {
      or_bsrmv(nrows, bdim, rowptr, colidx, vals, x, Ap);
      for (int64_t q = 0; q < n; ++q) r[q] = b[q] - Ap[q];
    } else {
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < n; ++q) r[q] = r[q] - alpha * Ap[q];
    }
    precond_apply(nrows, bdim, o->precond, dinv, r, z);
    double rzn = or_dot(n, z, r);
    double beta = rzn / rz;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) p[q] = z[q] + beta * p[q];
    rz = rzn;
    rr = or_dot(n, r, r);
    nr = o->norm_form ? sqrt(rr) : rr;
  }
  if (st) {
    st->iters = iter;
    st->rr = rr;
    or_bsrmv(nrows, bdim, rowptr, colidx, vals, x, Ap);
    double e = 0.0;
    for (int64_t q = 0; q < n; ++q) {
      double d = b[q] - Ap[q];
      e += d * d;
    }
    double bn = or_dot(n, b, b);
    st->true_rel = bn > 0 ? sqrt(e / bn) : sqrt(e);
  }
  free(r); free(p); free(z); free(Ap); free(dinv);
}

/* test/input/program/isprings.sim:71-103 */
void or_springs_step(const or_springs_params *P, const or_cg_opts *o,
                     int32_t np, int64_t ns, const int32_t *ep, double *x,
                     double *v, const double *m, const uint8_t *fixed,
                     const double *k, const double *l0,
                     const uint32_t *rowptr, const uint32_t *colidx,
                     or_cg_stats *st) {
  const int64_t nnz = rowptr[np];
  double *vals = (double *)malloc(sizeof(double) * 9 * (nnz > 0 ? nnz : 1));
  double *f = (double *)malloc(sizeof(double) * 3 * np);
  or_springs_assemble(P, np, ns, ep, x, v, m, fixed, k, l0, rowptr, colidx, vals, f);
  or_cg(np, 3, rowptr, colidx, vals, f, v, o, st);
  for (int64_t q = 0; q < 3 * (int64_t)np; ++q) x[q] = x[q] + P->h * v[q];
  free(vals);
  free(f);
}
