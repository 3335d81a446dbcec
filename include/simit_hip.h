/*
 * simit_hip.h -- C ABI of the MI355X execution backend for Simit's
 * sparsity / assembly / blocked-SpMV / CG path.
 *
 * Plain C: opaque handles, plain device pointers, sizes and int status
 * codes.  Every call is stream-ordered on the context's HIP stream unless it
 * says it synchronises.  All pointers marked "device" must be device memory
 * of the context's device (hipMalloc'd by the caller, or by smh_malloc).
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the simit-lang/simit source tree).  The reference's GPU path
 * generated these loops at run time as NVVM IR (src/backend/gpu/); here they
 * are fixed, hand-written gfx950 kernels behind this ABI, called by the C++
 * backend::Function work-alike in simit_amd/csrc/simit/.
 */
#ifndef SIMIT_HIP_H
#define SIMIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------
 * The C++ layer turns any non-zero status into simit::SimitException, as the
 * reference does for user/internal errors (src/error.cpp:87-97) and CUDA
 * driver errors (src/backend/gpu/gpu_codegen.h:15-51). */
#define SMH_OK 0
#define SMH_ERR_INVALID 1       /* bad argument (null, size, precision ...)  */
#define SMH_ERR_HIP 2           /* HIP runtime error                          */
#define SMH_ERR_UNSUPPORTED 3   /* valid but outside this backend's limits    */
#define SMH_ERR_NOMEM 4         /* device allocation failed                   */
#define SMH_ERR_BREAKDOWN 5     /* CG breakdown: p'Ap <= 0, z'r <= 0, NaN     */
#define SMH_ERR_NOT_CONVERGED 6 /* hit maxiters before the tolerance          */

/* value precisions of a blocked-CSR matrix */
#define SMH_F64 0   /* double values                                         */
#define SMH_F32 1   /* float values                                          */
#define SMH_F32X2 2 /* float hi + float lo = the FP64 value to ~2^-48        */
#define SMH_F32D 3  /* float values, double vectors (smh_bsrmv, smh_cg_solve) */
#define SMH_F32P 4  /* float 3x3 values in the lane-planar layout of
                       smh_bsr_planar, float vectors (smh_bsrmv only)       */

/* index build flags */
#define SMH_INDEX_FORCE_DIAG 1u /* every row holds its own vertex            */

typedef struct smh_ctx smh_ctx;
typedef struct smh_index smh_index;

/* ---- context -------------------------------------------------------------
 * Replaces GPUFunction's cuInit/cuCtxCreate on device 0
 * (src/backend/gpu/gpu_function.cpp:55-144). */
int smh_ctx_create(int device, smh_ctx **out);
int smh_ctx_destroy(smh_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *smh_ctx_last_error(const smh_ctx *ctx);
/* Order all later work on an existing hipStream_t (NULL: HIP's default
 * stream).  A new context runs on a stream of its own. */
int smh_ctx_set_stream(smh_ctx *ctx, void *hip_stream);
void *smh_ctx_get_stream(const smh_ctx *ctx);
int smh_ctx_synchronize(smh_ctx *ctx);
/* One line naming the context's device (name, architecture, compute units,
 * memory) into buf (NUL-terminated, truncated to len): what
 * backend::Function::printMachine reports (backend_function.h:59-62). */
int smh_ctx_describe(const smh_ctx *ctx, char *buf, size_t len);

/* ---- device memory: replaces cuMemAlloc / cuMemcpyHtoD / cuMemcpyDtoH in
 * GPUFunction::pushSetData, pushGlobalTensor, mapArgs and unmapArgs
 * (gpu_function.cpp:174-197, :219-418).  The copies synchronise. */
int smh_malloc(smh_ctx *ctx, size_t bytes, void **dptr);
int smh_free(smh_ctx *ctx, void *dptr);
int smh_memcpy_htod(smh_ctx *ctx, void *dst, const void *src, size_t bytes);
int smh_memcpy_dtoh(smh_ctx *ctx, void *dst, const void *src, size_t bytes);
int smh_memset(smh_ctx *ctx, void *dst, int value, size_t bytes);

/* ---- sparsity index ------------------------------------------------------
 * Replaces PathIndexBuilder::buildSegmented for the vev path expression
 * (src/path_indices.cpp:148-530; pexpr from path_expression_analysis.cpp:
 * 84-91), which the reference evaluates on the host with std::map/std::set
 * at every Function::init (llvm_function.cpp:438-465, gpu_function.cpp:651).
 * Built on the device, bit-identical: row v holds the sorted, unique
 * endpoints of every element v belongs to; a vertex with no element has an
 * empty row (unless SMH_INDEX_FORCE_DIAG).  nsets edge sets are united, as
 * an Or of their vev expressions is (path_indices.cpp:432-471).
 * d_endpoints[s] is the device copy of Set::getEndpointsData() of set s:
 * int32[ne[s] * card[s]] (graph.h:529).  Also keeps the vertex->element
 * incidence (Link::ve, path_indices.cpp:248-276) of every set, ascending. */
int smh_index_build(smh_ctx *ctx, uint32_t nverts, uint32_t nsets,
                    const uint64_t *ne, const uint32_t *card,
                    const int32_t *const *d_endpoints, uint32_t flags,
                    smh_index **out);
typedef struct {
  uint32_t nrows;
  uint64_t nnz;
  const uint32_t *rowptr; /* device, [nrows+1] ("<idx>.coords")            */
  const uint32_t *colidx; /* device, [nnz]     ("<idx>.sinks")             */
} smh_index_view;
int smh_index_get(const smh_index *idx, smh_index_view *view);
/* Link::ve of set s: device ve_ptr[nrows+1], ve_elem[ve_ptr[nrows]]. */
int smh_index_get_ve(const smh_index *idx, uint32_t set,
                     const uint32_t **ve_ptr, const uint32_t **ve_elem,
                     uint64_t *n_incidences);
int smh_index_destroy(smh_ctx *ctx, smh_index *idx);

/* gatherVVLocs (src/inline.cpp:524-577) with loc() (src/runtime.cpp:19-23):
 * locs[e*card*card + a*card + b] = position of endpoint b in endpoint a's
 * row, for every element of set s.  Device int32 output. */
int smh_index_locs(smh_ctx *ctx, const smh_index *idx, uint32_t set,
                   int32_t *d_locs);

/* ---- blocked-CSR matrix ----------------------------------------------------
 * The reference's Indexed tensor storage: vals[ij*b*b + bi*b + bj]
 * (src/runtime.h:21-44, src/ffi.h:23-74, lower_accesses.cpp:164-192). */
typedef struct {
  uint32_t nrows;
  uint32_t bdim;          /* 1 or 3                                        */
  uint64_t nnz;           /* blocks                                        */
  const uint32_t *rowptr; /* device                                        */
  const uint32_t *colidx; /* device                                        */
  int32_t precision;      /* SMH_F64 / SMH_F32 / SMH_F32X2 / SMH_F32D      */
  const void *vals;       /* device: double or float (hi part for F32X2)   */
  const void *vals_lo;    /* device: float lo part (SMH_F32X2 only)        */
} smh_bsr;

/* ---- assembly: map ... reduce + over an edge set ---------------------------
 * Replaces the element loop that inlineMap (src/inline.cpp:821-888) and
 * LowerMapFunctionRewriter (src/lower/lower_maps.cpp:35-120) emit, zero-init
 * included (inline.cpp:878-885), and the reference GPU path's f32 atomicAdd
 * scatter (src/backend/gpu/gpu_backend.cpp:501-520, :788-846).  Here each
 * output block is owned by one lane that sums its elements in ascending
 * element order: no atomics, deterministic, and with SMH_F64 bit-identical
 * to the CPU order.  Outputs: vals (precision as given; vals_lo only for
 * SMH_F32X2) on idx's pattern, and the dense right-hand side rhs (double). */

/* test/input/program/isprings.sim:19-69 (and apps/springs/isprings.sim):
 * A = MD + K, f = fmg + fe. */
typedef struct {
  double h;
  double viscous;
  double grav[3];
  double fixed_fill;      /* value a fixed point's MD block is set to      */
  int32_t fixed_identity; /* 1: fixed_fill*I; 0: broadcast to all nine     */
} smh_springs_params;
typedef struct {
  uint32_t npoints;
  uint64_t nsprings;
  const int32_t *endpoints; /* device int32[nsprings*2]                    */
  const double *x, *v;      /* device double[npoints*3]  (Point.x, .v)     */
  const double *m;          /* device double[npoints]    (Point.m)         */
  const uint8_t *fixed;     /* device bool[npoints]      (Point.fixed)     */
  const double *k, *l0;     /* device double[nsprings]   (Spring.k, .l0)   */
} smh_springs_set;
int smh_assemble_springs(smh_ctx *ctx, const smh_index *idx,
                         const smh_springs_params *p, const smh_springs_set *s,
                         int32_t precision, void *vals, void *vals_lo,
                         double *rhs);

/* apps/fem/fem_linear.sim:95-194: per-tet B = inv(M), W = -det(M)/6,
 * Ke = W*B'*E*B scaled by stiff_scale on free-free pairs, lumped mass
 * rho*W/4 on each diagonal, body load W/4*load on free rows, and
 * fixed_diag*I added on constrained rows (c > 0). */
typedef struct {
  double stiff_scale;
  double rho;
  double load[3];
  double fixed_diag;
} smh_tet_params;
typedef struct {
  uint32_t nverts;
  uint64_t ntets;
  const int32_t *endpoints; /* device int32[ntets*4]                       */
  const double *x;          /* device double[nverts*3]                     */
  const int32_t *c;         /* device int32[nverts]; > 0 means constrained */
  const double *mu, *lam;   /* device double[ntets] (Tet.u, Tet.l)         */
} smh_tet_set;
int smh_assemble_tet_linear(smh_ctx *ctx, const smh_index *idx,
                            const smh_tet_params *p, const smh_tet_set *s,
                            int32_t precision, void *vals, void *vals_lo,
                            double *rhs);

/* test/input/program/femTet.sim: neo-Hookean tets (PK1 and its derivative
 * dPdF with log(det F), :27-39).  main's two maps, A = map
 * compute_stiffness(h) to tets reduce + (:74-124: -(h*h) dH blocks on free
 * pairs, then the lumped mass 0.25*rho*W on every diagonal) and f = map
 * compute_force(h) to tets reduce + (:37-72: h times the element's internal
 * force columns on free vertices, then h*fg + m*v), in one pass.  The rest
 * shape B and volume W come from smh_femtet_precompute (initializeTet). */
typedef struct {
  double h;        /* time step (main: 0.01)                         */
  double rho;      /* density (1000)                                 */
  double grav[3];  /* gravity ((0, -10, 0))                          */
} smh_femtet_params;
typedef struct {
  uint32_t nverts;
  uint64_t ntets;
  const int32_t *endpoints; /* device int32[ntets*4]                       */
  const double *x, *v;      /* device double[nverts*3] (Vert.x, Vert.v)    */
  const int32_t *c;         /* device int32[nverts]; > 0 means constrained */
  const double *u, *l;      /* device double[ntets] (Tet.u, Tet.l)         */
  const double *W;          /* device double[ntets] rest volume (Tet.W)    */
  const double *B;          /* device double[ntets*9] (Tet.B, row-major)   */
} smh_femtet_set;
int smh_assemble_tet_neohookean(smh_ctx *ctx, const smh_index *idx,
                                const smh_femtet_params *p, const smh_femtet_set *s,
                                int32_t precision, void *vals, void *vals_lo, double *rhs);
/* initializeTet (femTet.sim:126-148), verts.m = map precomputeTetMat to
 * tets: W, B (device double[ntets], [ntets*9]) of every tet from s->x, and
 * m (device double[nverts]) = 0.25*rho*W of the last tet holding each
 * vertex (a map without a reduction assigns); vertices in no tet keep m. */
int smh_femtet_precompute(smh_ctx *ctx, const smh_index *idx, const smh_femtet_params *p,
                          const smh_femtet_set *s, double *W, double *B, double *m);

/* Scalar trilinear hex Poisson (1x1 blocks): kappa_e*h*Kref on free pairs,
 * source*h^3/8 on free rows, fixed_diag on Dirichlet rows.  Corner order
 * a = dx + 2*dy + 4*dz. */
typedef struct {
  double h;
  double source;
  double fixed_diag;
} smh_hex_params;
typedef struct {
  uint32_t nverts;
  uint64_t nhex;
  const int32_t *endpoints; /* device int32[nhex*8]                        */
  const int32_t *dirichlet; /* device int32[nverts]; > 0 means Dirichlet   */
  const double *kappa;      /* device double[nhex]                         */
} smh_hex_set;
int smh_assemble_hex_poisson(smh_ctx *ctx, const smh_index *idx,
                             const smh_hex_params *p, const smh_hex_set *s,
                             int32_t precision, void *vals, void *vals_lo,
                             double *rhs);

/* An edge set whose element function writes one per-edge block w_e into the
 * four blocks of its two endpoints with signs s[a*2+b] (all +1:
 * test/input/assembly/matrix_vv.sim, test/input/system/gemv_blocked.sim;
 * +1,-1,-1,+1: test/input/program/cg.sim:13-18), followed by the affine
 * system expression A = alpha*I + beta*A (cg.sim:57: I - 0.01*(map ...)).
 * bdim 1, 2 or 3; w is device double[nedges*bdim*bdim]; no right-hand side. */
typedef struct {
  uint32_t nverts;
  uint64_t nedges;
  uint32_t bdim;
  const int32_t *endpoints; /* device int32[nedges*2]                      */
  const double *w;          /* device double[nedges*bdim*bdim]             */
  double sign[4];
  double alpha, beta;
} smh_edge_blocks;
int smh_assemble_edge_blocks(smh_ctx *ctx, const smh_index *idx,
                             const smh_edge_blocks *e, int32_t precision,
                             void *vals, void *vals_lo);

/* ---- linear algebra --------------------------------------------------------
 * y = A*x: the blocked row loop of lower_indexexprs.cpp:630-667.  x and y
 * are float for SMH_F32 / SMH_F32P and double for SMH_F64 / SMH_F32X2 /
 * SMH_F32D.  exact_order=1
 * (SMH_F64 only) sums each row in the reference's ascending ij, bj order
 * with no fused multiply-add: bit-identical to the CPU loop. */
int smh_bsrmv(smh_ctx *ctx, const smh_bsr *A, const void *x, void *y,
              int32_t exact_order);
/* The SMH_F32X2 pair of n FP64 values: hi[i] = fl32(v[i]), lo[i] =
 * fl32(v[i] - hi[i]), the rounding the assemblies apply when they write
 * SMH_F32X2.  No counterpart in the reference (one global float type there):
 * a matrix assembled once in FP64 gives both the mixed solver's inner
 * operator and the FP64 values it certifies with (smh_cg_opts.vals_f64). */
int smh_split_hilo(smh_ctx *ctx, uint64_t n, const double *v, float *hi, float *lo);
/* r'r style inner product (flatten.cpp:286-306; the reference GPU path used
 * one global atomicAdd, gpu_backend.cpp:339-365).  Deterministic two-level
 * reduction in FP64; synchronises and writes *result on the host.
 * precision SMH_F64: double inputs; SMH_F32: float inputs. */
int smh_dot(smh_ctx *ctx, int32_t precision, uint64_t n, const void *x,
            const void *y, double *result);
/* The lane-planar layout of FP32 3x3 values (precision SMH_F32P), for the
 * SpMV's loads.  No counterpart in the reference, which keeps the row-major
 * block layout above (src/runtime.h:21-44): the same values, reordered.  The
 * blocks are cut into groups of 64 consecutive blocks; group g occupies the
 * 576 floats at d_out + 576*g: entries 0..3 of its 64 blocks (4 floats per
 * block), then entries 4..7, then entry 8 (entry 3*bi + bj of a block).
 * d_out holds 576*ceil(nnz/64) floats.  A: bdim 3, SMH_F32 (or the hi part
 * of SMH_F32X2).  smh_cg_solve with mixed = 1 builds such a copy of the hi
 * values itself (kept per pattern in the context until smh_bsr_forget). */
int smh_bsr_planar(smh_ctx *ctx, const smh_bsr *A, float *d_out);
/* smh_bsrmv and smh_cg_solve cut a pattern into row tiles on first use and
 * cache that plan per context, keyed by (rowptr pointer, nrows, nnz).  Call
 * this before a rowptr buffer is freed or refilled with another pattern of
 * the same size (smh_index_destroy does it for index-owned buffers). */
int smh_bsr_forget(smh_ctx *ctx, const uint32_t *rowptr);
/* y = a*x + b*y on double vectors: the dense vector statements of Simit
 * programs (e.g. points.x = points.x + h*points.v, isprings.sim:94; the
 * reference GPU path ran each as its own kernel, gpu_backend.cpp:1150-1215). */
int smh_axpby(smh_ctx *ctx, uint64_t n, double a, const double *x, double b,
              double *y);
/* Block-diagonal copy and 3x3 inverse per row (Jacobi, thermal.sim:229-233;
 * inv3 of runtime.cpp:135-160): d_dinv is double[nrows*bdim*bdim]. */
int smh_block_jacobi(smh_ctx *ctx, const smh_bsr *A, double *d_dinv);

/* ---- CG solve ---------------------------------------------------------------
 * The CG loops Simit programs write inline (apps/springs/isprings.sim:77-93,
 * apps/fem/fem_linear.sim:211-230, test/input/program/cg.sim:25-53,
 * apps/thermal/thermal.sim:75-130) as one device-resident solver: no
 * per-operation host sync (the reference ran the loop on a one-thread
 * harness with a device sync after every child launch,
 * gpu_backend.cpp:564-567, :985).  b and x are device double[nrows*bdim];
 * x holds the initial guess and receives the solution. */
#define SMH_PRECOND_NONE 0
#define SMH_PRECOND_JACOBI 1       /* scalar 1/A(i,i)                        */
#define SMH_PRECOND_BLOCK_JACOBI 2 /* 3x3 inverse of the diagonal block      */
#define SMH_PRECOND_LEARNED 3      /* block Jacobi + learned block-sparse
                                      correction (smh_gnn_precond)          */
typedef struct {
  int32_t maxiters;
  double tol;
  int32_t rel;           /* 0: stop when r'r <= tol; 1: ||r|| <= tol*||b||  */
  int32_t norm_form;     /* cg.sim form: norm(r) > tol                      */
  int32_t replace_every; /* > 0: r = b - A x every k iterations             */
  int32_t precond;
  int32_t mixed;         /* 1: FP32 inner iteration on vals (hi part) with
                            FP64 reliable updates r = b - A x on hi+lo
                            (needs SMH_F32X2); its SpMV multiplies FP32
                            values by FP32 vectors and sums each row in
                            FP64, rounding q once; 2: the same with the inner
                            vectors (p, q, r, z) kept in FP64 (over
                            stencil storage the learned correction's
                            input u = Linv r is gathered in FP32)         */
  double reliable_delta; /* mixed: update when ||r|| < delta*max ||r||
                            (default 0.3)                                   */
  int32_t check_every;   /* iterations between host convergence checks     */
  const float *learned;  /* SMH_PRECOND_LEARNED: device float[nrows*9],
                            Linv_i = L_i^-1 for A_ii = L_i L_i^T (row-major,
                            lower triangular); float[nrows], 1/sqrt(a_ii),
                            for 1x1 blocks; as smh_gnn_precond writes */
  int32_t zr_stop;       /* 1: also stop when z'r <= the threshold, as
                            thermal.sim's pre_conjgrad does (:115)          */
  const uint8_t *learned_corr; /* SMH_PRECOND_LEARNED: device bytes,
                            4-byte aligned: the normalised correction blocks
                            C_k on A's pattern (zero on the diagonal); 3x3
                            blocks: [nnz*12], FP8 e4m3 of 32 C_k, three
                            values and a zero pad byte per block row; 1x1
                            blocks: float[nnz]; the preconditioner is
                            z_i = Linv_i^T (Linv_i r_i + sum_k C_k Linv_j r_j) */
  int32_t fallback;      /* 1 (default): when the solve breaks down (p'Ap <= 0,
                            z'r <= 0 with a preconditioner, NaN) -- or, with
                            the learned preconditioner, ends unconverged --
                            solve again from the initial guess with (block)
                            Jacobi (after the learned one), then with no
                            preconditioner; each rung is certified by its
                            own FP64 residual.  0: report the first status */
  const double *vals_f64; /* mixed, optional: device FP64 values of the same
                            matrix in A's layout (blocked CSR: [nnz*b*b];
                            stencil: [27*n]), e.g. the assembly's FP64
                            output that hi and lo were split from
                            (smh_split_hilo).  When set, the reliable
                            updates, the (block) Jacobi inverse and the
                            final certification use these values instead
                            of hi + lo, which carries each value only to
                            ~2^-48: on stiff systems that difference alone
                            can move ||b - A x|| / ||b|| past 1e-8.  NULL:
                            hi + lo.  Not taken by the batched solver.    */
} smh_cg_opts;
typedef struct {
  int32_t iters;         /* iterations, summed over the rungs of the chain  */
  int32_t reliable_updates;
  int32_t status;        /* SMH_OK / SMH_ERR_BREAKDOWN / SMH_ERR_NOT_CONVERGED */
  double rr;             /* final iterated r'r                              */
  double true_rel;       /* ||b - A x|| / ||b|| in FP64 at exit             */
  double seconds;        /* wall time of the solve (host clock, synced)     */
  int32_t precond_used;  /* the preconditioner whose solve produced x       */
  int32_t fallbacks;     /* rungs of the fallback chain taken (0: none)      */
  int32_t first_status;  /* status of the requested preconditioner's solve   */
} smh_cg_stats;
void smh_cg_default_opts(smh_cg_opts *o);
int smh_cg_solve(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x,
                 const smh_cg_opts *opts, smh_cg_stats *stats);

/* smh_cg_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels, results (bit for bit) and statuses
 * as smh_cg_solve, fallback chain included; preconditioners none, Jacobi and
 * block Jacobi (not the learned one).  Every piece of scratch -- the row-tile
 * plan (rebuilt on each call), the CG vectors, the (block) Jacobi inverse,
 * the reductions and the fallback chain's copy of the initial guess -- lives
 * in one 32 MiB array of device memory that the library holds on each device
 * until the process ends; a system that needs more is rejected with
 * SMH_ERR_UNSUPPORTED before any work is queued (e.g. a 3x3 FP64
 * block-Jacobi solve fits up to ~170,000 rows, a scalar one ~690,000).
 * All calls on one device share that array and leave their contents in it:
 * calls that overlap on one device (two threads, or two contexts of one
 * device) race. */
int smh_cg_solve_noalloc(smh_ctx *ctx, const smh_bsr *A, const double *b, double *x,
                         const smh_cg_opts *opts, smh_cg_stats *stats);

/* ---- scalar 27-point stencil storage ---------------------------------------
 * Replaces TensorStorage::Stencil (src/storage.h:48) and its SpMV lowering
 * (src/lower/index_expressions/lower_indexexprs.cpp:669-711) for scalar
 * systems on a structured n0 x n1 x n2 grid of rows (row id =
 * i0 + n0*(i1 + n1*i2)): no column indices; vals[s*nrows + i] couples row i
 * with column i + off_s, off_s = d0 + n0*d1 + n0*n1*d2 for (d2, d1, d0) in
 * lexicographic order over {-1,0,1}^3 (ascending offsets, s = 13 the
 * diagonal), zero where there is no coupling.  Any 1x1 matrix whose columns
 * lie within those offsets of their row is representable (the trilinear hex
 * Poisson matrix of smh_assemble_hex_poisson is).  Products sum in ascending
 * column order with each product rounded, so smh_stencil27_mv equals the
 * blocked-CSR smh_bsrmv of the same matrix bit for bit. */
typedef struct {
  uint32_t n0, n1, n2;
  int32_t precision;     /* SMH_F64, SMH_F32, SMH_F32X2, SMH_F32D          */
  const void *vals;      /* device [27*nrows] (double for SMH_F64, else float) */
  const void *vals_lo;   /* SMH_F32X2: device float[27*nrows], else NULL    */
  uint32_t flags;        /* SMH_STENCIL_PERIODIC, or 0                       */
} smh_stencil27;
/* Neighbour coordinates wrap around (the rows of a grid edge set's points,
 * whose edges are periodic: graph.h:121-168, system-gemv-tests.cpp:60-337).
 * smh_stencil27_mv only; the solver rejects it. */
#define SMH_STENCIL_PERIODIC 1u
/* scatter a 1x1 blocked-CSR matrix (any precision; SMH_F32X2 copies both
 * halves) into stencil storage; SMH_ERR_INVALID if a column lies outside its
 * row's 27 offsets. */
int smh_stencil27_from_bsr(smh_ctx *ctx, const smh_bsr *A, uint32_t n0, uint32_t n1, uint32_t n2,
                           void *vals, void *vals_lo);
/* the float[nnz] learned correction of smh_gnn_precond (pass it as the
 * values of an SMH_F32 1x1 matrix with A's pattern) as IEEE half stencil
 * slots, out = device uint16[27*n] (4-byte aligned): the opts.learned_corr
 * of smh_cg_solve_stencil27 */
int smh_stencil27_correction(smh_ctx *ctx, const smh_bsr *C, uint32_t n0, uint32_t n1, uint32_t n2,
                             void *out);
/* y = A*x (vector types as smh_bsrmv) */
int smh_stencil27_mv(smh_ctx *ctx, const smh_stencil27 *A, const void *x, void *y);
/* smh_cg_solve over stencil storage: same options, state machine and
 * certification; preconditioners none, Jacobi (block Jacobi = Jacobi for
 * 1x1) and learned (opts.learned float[n], opts.learned_corr FP16 [27*n]
 * from smh_stencil27_correction) */
int smh_cg_solve_stencil27(smh_ctx *ctx, const smh_stencil27 *A, const double *b, double *x,
                           const smh_cg_opts *opts, smh_cg_stats *stats);

/* ---- batched CG: many independent systems in one launch ------------------
 * The systems are the diagonal blocks of one matrix: system m owns rows
 * [rows[m], rows[m+1]) (host array of nbatch+1 entries from 0 to nrows) and
 * no block couples two systems (checked).  Each system runs its own CG
 * (own alpha, beta and stopping test on its own ||b_m||, as if solved alone
 * by smh_cg_solve) in lockstep; opts->check_every iterations are captured
 * once as a hipGraph and replayed, the replays going out ahead of the host,
 * which watches their progress in pinned memory; the SpMV skips the row
 * tiles of systems that have stopped.  The SpMV's workgroup count is
 * autotuned on first use of a pattern.  stats: host array of nbatch
 * entries (may be NULL); seconds (may be NULL): wall time of the solve. */
typedef struct {
  int32_t iters;
  int32_t status;        /* SMH_OK / SMH_ERR_BREAKDOWN / SMH_ERR_NOT_CONVERGED */
  double true_rel;       /* ||b_m - A_m x_m|| / ||b_m|| in FP64 at exit     */
} smh_cg_batch_stats;
int smh_cg_solve_batched(smh_ctx *ctx, const smh_bsr *A, uint32_t nbatch, const uint32_t *rows,
                         const double *b, double *x, const smh_cg_opts *opts,
                         smh_cg_batch_stats *stats, double *seconds);

/* One implicit springs step of many independent lattices (the systems of
 * smh_cg_solve_batched) as captured graphs: the first graph holds the
 * assembly (smh_assemble_springs into vals / vals_lo / rhs on idx's
 * pattern; SMH_F32X2 when opts->mixed, else SMH_F64), the solver's set-up
 * and the first check_every iterations, the second check_every more
 * iterations and is replayed until every system has stopped.  The host
 * learns of that from pinned memory the replays write (no copies or stream
 * synchronisation between replays).  x: initial guess and solution. */
int smh_springs_solve_batched(smh_ctx *ctx, const smh_index *idx, const smh_springs_params *p,
                              const smh_springs_set *s, void *vals, void *vals_lo, double *rhs,
                              uint32_t nbatch, const uint32_t *rows, double *x,
                              const smh_cg_opts *opts, smh_cg_batch_stats *stats,
                              double *seconds);

/* ---- learned block-Jacobi correction (no counterpart in the reference) ----
 * A message-passing network over the matrix graph (3 layers, width 32, BF16
 * weights and MFMA inputs, FP32 accumulation) predicts a symmetric correction
 * C to the block-Jacobi preconditioner on A's own sparsity pattern, in the
 * block-Jacobi-normalised space Ahat = Linv A Linv^T.  C is exactly
 * symmetric; each block is soft-saturated (|C_k|_F < 1) but not capped per
 * row.  A spectral guard keeps I + C positive definite: a 32-step Lanczos
 * estimate of lambda_min of the stored C, less 5% of the estimated spectral
 * width, is held at or above -kappa (kappa in (0, 1)) by a global scale of C,
 * so lambda_min(I + C) >= 1 - kappa and the preconditioner
 * Linv^T (I + C) Linv is symmetric positive definite.
 * Training minimises the stochastic Frobenius loss
 *   L = sum (V - (I + C) Ahat V)^2 / (K * ndof)
 * over K Gaussian probe vectors V, with Adam, entirely in the library's own
 * forward and backward kernels.  A must be 3x3 or 1x1 blocks with a
 * symmetric pattern whose rows hold their diagonal block. */
typedef struct smh_gnn smh_gnn;
#define SMH_GNN_NPARAMS 18192
int smh_gnn_create(smh_ctx *ctx, uint64_t seed, double kappa, smh_gnn **out);
int smh_gnn_destroy(smh_ctx *ctx, smh_gnn *g);
/* flat float32 parameters: W_ev[16x32] b_ev W_ee[16x32] b_ee, per layer
 * W_m[96x32] b_m W_u[64x32] b_u, then W_d[96x16] b_d (row-major, in x out) */
int smh_gnn_get_params(smh_ctx *ctx, smh_gnn *g, float *host);
int smh_gnn_set_params(smh_ctx *ctx, smh_gnn *g, const float *host);
/* the preconditioner of A (3x3 or 1x1 blocks): linv = device
 * float[nrows*bdim*bdim], corr = device bytes [nnz*12] (3x3, FP8) or
 * [nnz*4] (1x1, float) in the layout of opts.learned_corr (pass both as
 * opts.learned / opts.learned_corr) */
int smh_gnn_precond(smh_ctx *ctx, smh_gnn *g, const smh_bsr *A, float *linv, uint8_t *corr);
/* the spectral guard of the last smh_gnn_precond: the Lanczos estimate of
 * lambda_min and the spectral width of the network's C (before scaling), the
 * scale applied to C (1 when none was needed) and the Lanczos steps taken;
 * any pointer may be NULL */
int smh_gnn_precond_info(smh_gnn *g, double *lambda_min, double *spectral_width, double *scale,
                         int32_t *steps);
/* loss and gradient on A for nprobes probe vectors: probes = device
 * float[bdim*nrows*nprobes] (row-major [dof][probe]) or NULL to draw them
 * from seed; grads (host float[SMH_GNN_NPARAMS]) may be NULL */
int smh_gnn_loss_grad(smh_ctx *ctx, smh_gnn *g, const smh_bsr *A, int32_t nprobes,
                      const float *probes, uint64_t seed, double *loss, float *grads);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on A with fresh probes;
 * loss (host) may be NULL, which avoids a synchronisation */
int smh_gnn_train_step(smh_ctx *ctx, smh_gnn *g, const smh_bsr *A, int32_t nprobes,
                       uint64_t seed, double lr, double *loss);

/* ---- compiled Simit programs ------------------------------------------------
 * The front-end (Program::loadFile) turns each map's element function into
 * HIP C++.  These entry points compile it at run time and launch it: the
 * counterpart of the reference GPU backend's PTX generation,
 * cuModuleLoadDataEx and cuLaunchKernel (GPUFunction::init,
 * gpu_function.cpp:622-760, :930).
 *
 * smh_hiprtc_compile: HIP source -> code object for arch ("gfx950" when
 * NULL), with the library's flags (-O3, no FMA contraction, no packed FP32).
 * Needs no device.  *image is malloc'ed (free with smh_image_free); log gets
 * the compiler's messages (errors on failure). */
int smh_hiprtc_compile(const char *source, const char *name, const char *arch, void **image,
                       size_t *bytes, char *log, size_t log_len);
void smh_image_free(void *image);
typedef struct smh_module smh_module;
int smh_module_load(smh_ctx *ctx, const void *image, size_t bytes, smh_module **out);
/* launches kernel over nthreads threads (256 per workgroup) on the context's
 * stream; args is the kernel's argument struct, args_bytes long */
int smh_module_launch(smh_ctx *ctx, smh_module *m, const char *kernel, uint64_t nthreads,
                      const void *args, size_t args_bytes);
int smh_module_destroy(smh_ctx *ctx, smh_module *m);
/* A map's output tensor from its staged writes.  The element kernel stages
 * every write in its own slot; entry i of out (bs components) takes the
 * writes src[ptr[i]..ptr[i+1]) in the order the reference performs them
 * (elements ascending, then the element function's execution order).
 * `map f to E reduce +` (taken NULL): out = ((0 + w0) + w1) + ..., the
 * compound writes of lower_maps.cpp:37-49 into the zeroed global tensor.
 * A map without a reduction (taken = per-slot flags of the writes executed):
 * plain writes, so out takes the last write taken and an entry no element
 * wrote keeps its value. */
int smh_stage_reduce(smh_ctx *ctx, uint64_t nout, uint32_t bs, const uint64_t *ptr,
                     const uint32_t *src, const double *stage, const uint8_t *taken,
                     double *out);
/* The elementwise statements of an exported Simit function over system
 * vectors (and matrix values), one rounding per operation, no contraction
 * (the reference's scalar loops, lower_indexexprs.cpp:669-711):
 *   LINCOMB out = a*x + b*y (y NULL: a*x)   DIVS out = x / a
 *   EMUL out = x .* y   EDIV out = x ./ y   FILL out = a
 *   ADDS out = x + a    SDIV out = a / x
 * out may alias x or y. */
#define SMH_VEC_LINCOMB 0
#define SMH_VEC_DIVS 1
#define SMH_VEC_EMUL 2
#define SMH_VEC_EDIV 3
#define SMH_VEC_FILL 4
#define SMH_VEC_ADDS 5
#define SMH_VEC_SDIV 6
int smh_vec_op(smh_ctx *ctx, int32_t op, uint64_t n, double a, const double *x, double b,
               const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
