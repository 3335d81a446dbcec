/*
 * CPU oracle for the Simit sparsity / assembly / blocked-SpMV / CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (simit_amd/, include/)
 * links, loads or calls this code.  Only tests/, the smoke() check and the
 * cpu_baseline leg of bench.py use it, always as the checker or as the
 * reported CPU baseline, never as the thing measured or shipped.
 *
 * Every routine is a single-threaded FP64 restatement of the loop nests the
 * reference's LLVM CPU backend emits for the hot path; each cites the
 * reference file:line it follows.  Compiled with -O2 -ffp-contract=off so
 * that every a*b+c is two roundings, like the reference's scalar code.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_index_*          reference fixtures test/path_indices-test.cpp (golden
 *                       vectors in tests/golden), the reference's own builder
 *                       compiled under oracle/_ref, closed-form row counts.
 *   or_loc              brute-force search + reference semantics runtime.cpp:19-23.
 *   or_springs_*        test/program-isprings.cpp:156-174 fixture, closed forms.
 *   or_tet_*            numpy W*B'*E*B literal product (apps/fem/fem_linear.sim:
 *                       95-174), rigid-mode null space, symmetry.
 *   or_hex_*            numpy 2x2x2 Gauss quadrature of the trilinear element.
 *   or_log              numpy/libm log on a sweep of arguments (<= 2 ulp).
 *   or_femtet_*         test/program-fem-test.cpp:97-105 (ten implicit steps of the
 *                       bar2k mesh reproduce the reference's positions to 1e-13),
 *                       central finite differences of the force (the stiffness is
 *                       its derivative), rigid translation/rotation invariance.
 *   or_bsrmv            dense numpy product; test/system-gemv-tests.cpp:854-903.
 *   or_dot              test/system-vector-tests.cpp:51-124 integer sums.
 *   or_cg_*             test/program-cg-test.cpp:41-43 fixture; scipy direct
 *                       solve residual.
 */
#ifndef SIMIT_ORACLE_H
#define SIMIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sparsity: path_indices.cpp ---------------------------------------- */
/* Link::ve (path_indices.cpp:248-276): for each vertex, the ids of the
 * elements it is an endpoint of, ascending, one entry per occurrence. */
int64_t or_index_ve(int32_t nv, int64_t ne, int32_t card, const int32_t *ep,
                    uint32_t *ptr, uint32_t *idx);
/* Link::ev (path_indices.cpp:207-247): element -> its endpoints in slot order */
void or_index_ev(int64_t ne, int32_t card, const int32_t *ep, uint32_t *ptr,
                 uint32_t *idx);
/* And(ve, ev) with an existential element (path_indices.cpp:359-429 plus
 * pack(sorted) :172-203): row v = sorted unique endpoints of every element v
 * belongs to.  Vertices with no element get an empty row.  With
 * force_diag != 0 every row also holds v itself (the Or of vev with the
 * diagonal path, path_indices.cpp:432-471).  nsets edge sets are united
 * (Or of several vev expressions).  Returns nnz, or -1 if cap is too small. */
int64_t or_index_vev(int32_t nv, int32_t nsets, const int64_t *ne,
                     const int32_t *card, const int32_t *const *ep,
                     int32_t force_diag, uint32_t *rowptr, uint32_t *colidx,
                     int64_t cap);

/* runtime.cpp:19-23: linear scan for column j in row i. */
int32_t or_loc(const uint32_t *rowptr, const uint32_t *colidx, int32_t i,
               int32_t j);

/* ---- element math: runtime.cpp:91-200 -------------------------------- */
double or_det3(const double *a);
void or_inv3(const double *a, double *out);

/* ---- springs: test/input/program/isprings.sim:19-69 ----------------- */
typedef struct {
  double h;          /* time step                          */
  double viscous;    /* damping coefficient in MD          */
  double grav[3];    /* gravity vector                     */
  double fixed_fill; /* value broadcast into a fixed point's MD block */
  int32_t fixed_identity; /* 1: fixed block = fixed_fill*I; 0: all nine entries */
} or_springs_params;

/* Assemble MDK = MD + K (3x3 BCSR on the vev index) and f = fmg + fe.
 * Zero-init, point map, then spring map in ascending spring order with
 * A[loc*9 + 3bi + bj] += (inline.cpp:821-888, lower_maps.cpp:35-120). */
void or_springs_assemble(const or_springs_params *P, int32_t np, int64_t ns,
                         const int32_t *ep, const double *x, const double *v,
                         const double *m, const uint8_t *fixed,
                         const double *k, const double *l0,
                         const uint32_t *rowptr, const uint32_t *colidx,
                         double *vals, double *f);

/* ---- linear tets: apps/fem/fem_linear.sim:95-174, :176-194 ---------- */
typedef struct {
  double stiff_scale; /* h*h in fem_linear's main(); 1 for statics  */
  double rho;         /* lumped mass density (0: no mass term)       */
  double load[3];     /* body force per unit volume (right-hand side) */
  double fixed_diag;  /* identity scale on constrained rows (0: none) */
} or_tet_params;

/* 3x3 block (a,b) of W*B'*E*B for one tet with corner coordinates xs[4][3]
 * and Lame parameters mu, lambda.  Also returns W. */
void or_tet_block(const double *xs, double mu, double lam, int a, int b,
                  double *blk, double *W);
void or_tet_assemble(const or_tet_params *P, int32_t nv, int64_t nt,
                     const int32_t *ep, const double *x, const int32_t *c,
                     const double *mu, const double *lam,
                     const uint32_t *rowptr, const uint32_t *colidx,
                     double *vals, double *f);

/* ---- neo-Hookean tets, test/input/program/femTet.sim --------------- */
double or_log(double x);
void or_femtet_geom(const double *xs, double *B, double *W);
void or_femtet_precompute(double rho, int32_t nv, int64_t nt, const int32_t *ep,
                          const double *x, double *W, double *B, double *m);
void or_femtet_H(const double *xs, double u, double l, double W, const double *B, double *H);
void or_femtet_kb(const double *xs, double u, double l, double W, const double *B, int jj,
                  int row, double *out);
void or_femtet_assemble(double h, double rho, const double *grav, int32_t nv, int64_t nt,
                        const int32_t *ep, const double *x, const double *v, const int32_t *c,
                        const double *u, const double *l, const double *W, const double *B,
                        const uint32_t *rowptr, const uint32_t *colidx, double *vals,
                        double *f);

/* ---- trilinear hex, scalar Poisson ---------------------------------- */
typedef struct {
  double h;          /* cube edge length                  */
  double source;     /* uniform volumetric heat source    */
  double fixed_diag; /* identity on Dirichlet rows        */
} or_hex_params;
double or_hex_kref(int a, int b);
void or_hex_assemble(const or_hex_params *P, int32_t nv, int64_t nh,
                     const int32_t *ep, const int32_t *dirichlet,
                     const double *kappa, const uint32_t *rowptr,
                     const uint32_t *colidx, double *vals, double *f);

/* ---- linear algebra: lower_indexexprs.cpp:630-667, flatten.cpp:286-306 */
void or_bsrmv(int32_t nrows, int32_t bdim, const uint32_t *rowptr,
              const uint32_t *colidx, const double *vals, const double *x,
              double *y);
double or_dot(int64_t n, const double *x, const double *y);

/* ---- CG: isprings.sim:84-103 / fem_linear.sim:211-230 / cg.sim:25-53 /
 *          thermal.sim:75-130 --------------------------------------------- */
typedef struct {
  int32_t maxiters;
  int32_t rel;          /* 0: stop on r'r > tol (absolute); 1: r'r > tol^2 b'b */
  int32_t replace_every;/* >0: r = b - A x every k-th iteration (thermal.sim) */
  int32_t precond;      /* 0 none, 1 scalar Jacobi, 2 3x3 block Jacobi        */
  int32_t norm_form;    /* 1: cg.sim form (norm(r) > tol, alpha = r'r/p'Ap)  */
  double tol;
} or_cg_opts;
typedef struct {
  int32_t iters;
  double rr;            /* final iterated r'r                                */
  double true_rel;      /* ||b - A x|| / ||b|| recomputed at the end         */
} or_cg_stats;
void or_cg(int32_t nrows, int32_t bdim, const uint32_t *rowptr,
           const uint32_t *colidx, const double *vals, const double *b,
           double *x, const or_cg_opts *o, or_cg_stats *st);

/* Full implicit springs step of test/input/program/isprings.sim:71-103:
 * assemble, CG from v, then x += h v.  Updates x and v in place. */
void or_springs_step(const or_springs_params *P, const or_cg_opts *o,
                     int32_t np, int64_t ns, const int32_t *ep, double *x,
                     double *v, const double *m, const uint8_t *fixed,
                     const double *k, const double *l0,
                     const uint32_t *rowptr, const uint32_t *colidx,
                     or_cg_stats *st);

#ifdef __cplusplus
}
#endif
#endif
