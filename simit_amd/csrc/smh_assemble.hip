// Deterministic, atomic-free assembly of `map f to E reduce +` into the
// blocked-CSR pattern of the vev index.
//
// Reference semantics (src/inline.cpp:821-888, src/lower/lower_maps.cpp:
// 35-120): zero the result, then for every element e in ascending order run
// the element function and add each written block into A[loc(i,j)].  Each
// stored value is therefore  0 + c(e1) + c(e2) + ...  over the elements that
// touch that block, in ascending element order.
//
// MI355X design ("pull"): one 64-lane wavefront owns 64 consecutive stored
// blocks, one lane per block (i, j).  The wave walks the ascending element
// lists (Link::ve) of its rows in chunks of 64 (row, element) pairs:
//   A. every lane takes one pair, gathers the element and computes what the
//      row's blocks need (geometry, scaled gradients ...) into a compact
//      per-pair record in LDS;
//   B. every block lane walks its row's pairs in list order and, for each
//      element that also holds j, forms block (i, j) from the record and adds
//      it.
// Each block is summed by exactly one lane in exactly the CPU order, in FP64,
// with no fused multiply-add: bit-identical to the CPU loop for SMH_F64.  The
// 64 results are staged through LDS so the stores are coalesced.  No atomics,
// so repeated runs are bitwise identical.
#include <algorithm>
#include <cstdlib>

#include "smh_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kWave = 64;

// ---- element programs --------------------------------------------------------
// Each op gives: State (registers, phase A) loaded from the element; Pair
// (LDS, one per staged pair) made for the row's endpoint slot a; block()
// forms the row's block with endpoint slot b from a Pair; extra() and rhs()
// the diagonal extra term and the right-hand-side contribution of the row.

// test/input/program/isprings.sim:19-69
struct SpringArgs {
  smh_springs_params p;
  const int32_t *ep;
  const double *x, *v, *m, *k, *l0;
  const uint8_t *fixed;
};
struct SpringOp {
  static constexpr int CARD = 2, BD = 3;
  static constexpr bool kHasRhs = true, kExtra = false;
  using Args = SpringArgs;
  struct State {
    int32_t vid[2];
    int32_t free_[2];
    double fe0[3];
    double kb[9];
  };
  struct Pair {
    int32_t vid[2];
    int32_t a, fa, fb;
    double kb[9];
    double fe0[3];
  };
  __device__ static void load(const Args &A, uint64_t e, State &st) {
    const int32_t a = A.ep[2 * e], b = A.ep[2 * e + 1];
    st.vid[0] = a;
    st.vid[1] = b;
    st.free_[0] = !A.fixed[a];
    st.free_[1] = !A.fixed[b];
    const double h = A.p.h, stiffness = A.k[e], l0 = A.l0[e];
    double dx[3];
    for (int q = 0; q < 3; ++q) dx[q] = A.x[3 * (int64_t)b + q] - A.x[3 * (int64_t)a + q];
    const double dxtdx = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double l = sqrt(dxtdx);
    const double s = h * stiffness / (l0 * l0) * (l - l0);
    for (int q = 0; q < 3; ++q) st.fe0[q] = s * dx[q] / l;
    const double c = h * h * stiffness / (l0 * l0 * l * l);
    const double r = (l - l0) / l;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double o = dx[i] * dx[j];
        const double d = (i == j ? dxtdx : 0.0) - o;
        st.kb[3 * i + j] = c * (o + r * d);
      }
  }
  __device__ static void pair(const Args &, const State &st, int a, Pair &P) {
    P.vid[0] = st.vid[0];
    P.vid[1] = st.vid[1];
    P.a = a;
    P.fa = a == 0 ? st.free_[0] : st.free_[1];
    P.fb = a == 0 ? st.free_[1] : st.free_[0];
    for (int q = 0; q < 9; ++q) P.kb[q] = st.kb[q];
    for (int q = 0; q < 3; ++q) P.fe0[q] = st.fe0[q];
  }
  // K(p(a),p(b)) written iff free_a and (a == b or free_b) (isprings.sim:58-68)
  __device__ static bool block(const Args &, const Pair &P, int b, double *out) {
    if (!P.fa || (P.a != b && !P.fb)) return false;
    for (int q = 0; q < 9; ++q) out[q] = P.a == b ? P.kb[q] : -P.kb[q];
    return true;
  }
  __device__ static bool extra(const Args &, const Pair &, double *) { return false; }
  __device__ static void rhs(const Args &, const Pair &P, double *f) {
    if (!P.fa) return;
    for (int q = 0; q < 3; ++q) f[q] += P.a == 0 ? P.fe0[q] : -P.fe0[q];
  }
  // MDK = MD + K and f = fmg + fe, with MD/fmg from compute_mass_damping
  // (isprings.sim:19-36), for the lane that owns row i's diagonal block.
  __device__ static void finish_diag(const Args &A, uint32_t i, double *blk, double *f) {
    double md[9], fmg[3] = {0.0, 0.0, 0.0};
    const double mass = A.m[i];
    if (A.fixed[i]) {
      for (int q = 0; q < 9; ++q)
        md[q] = (A.p.fixed_identity && (q % 4 != 0)) ? 0.0 : A.p.fixed_fill;
    } else {
      const double dm = mass + A.p.h * A.p.viscous;
      for (int q = 0; q < 9; ++q) md[q] = (q % 4 == 0) ? dm : 0.0;
      for (int q = 0; q < 3; ++q)
        fmg[q] = mass * A.v[3 * (int64_t)i + q] + A.p.h * mass * A.p.grav[q];
    }
    for (int q = 0; q < 9; ++q) blk[q] = md[q] + blk[q];
    for (int q = 0; q < 3; ++q) f[q] = fmg[q] + f[q];
  }
  __device__ static void isolated_rhs(const Args &A, uint32_t i, double *f) {
    f[0] = f[1] = f[2] = 0.0;
    if (A.fixed[i]) return;
    const double mass = A.m[i];
    for (int q = 0; q < 3; ++q)
      f[q] = mass * A.v[3 * (int64_t)i + q] + A.p.h * mass * A.p.grav[q] + 0.0;
  }
  __device__ static void finish_block(const Args &, uint32_t, uint32_t, double *) {}
};

// apps/fem/fem_linear.sim:95-194
struct TetArgs {
  smh_tet_params p;
  const int32_t *ep;
  const double *x, *mu, *lam;
  const int32_t *c;
};
__device__ __forceinline__ double det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
__device__ __forceinline__ void inv3(const double *a, double *inv) {
  const double cof00 = a[4] * a[8] - a[5] * a[7];
  const double cof01 = -a[3] * a[8] + a[5] * a[6];
  const double cof02 = a[3] * a[7] - a[4] * a[6];
  const double cof10 = -a[1] * a[8] + a[2] * a[7];
  const double cof11 = a[0] * a[8] - a[2] * a[6];
  const double cof12 = -a[0] * a[7] + a[1] * a[6];
  const double cof20 = a[1] * a[5] - a[2] * a[4];
  const double cof21 = -a[0] * a[5] + a[2] * a[3];
  const double cof22 = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * cof00 + a[1] * cof01 + a[2] * cof02;
  d = 1.0 / d;
  inv[0] = cof00 * d; inv[1] = cof10 * d; inv[2] = cof20 * d;
  inv[3] = cof01 * d; inv[4] = cof11 * d; inv[5] = cof21 * d;
  inv[6] = cof02 * d; inv[7] = cof12 * d; inv[8] = cof22 * d;
}
// element a of a 4-slot array (stride apart) by selects, for a runtime a
template <typename T>
__device__ __forceinline__ T pick4(const T *v, int a, int stride = 1) {
  const T x01 = a == 0 ? v[0] : v[stride];
  const T x23 = a == 2 ? v[2 * stride] : v[3 * stride];
  return a < 2 ? x01 : x23;
}
struct TetOp {
  static constexpr int CARD = 4, BD = 3;
  static constexpr bool kHasRhs = true, kExtra = true;
  using Args = TetArgs;
  struct State {
    int32_t vid[4];
    int32_t free_[4];
    double dN[12];
    double W, mu, lam;
  };
  // u = (s*W*lam) g_a and w = (s*W*mu) g_a for the row's slot a, and all four
  // shape gradients g_b
  struct Pair {
    int32_t vid[4];
    uint32_t freemask;
    int32_t fa;
    double u[3], w[3];
    double g[12];
    double W;
  };
  __device__ static void load(const Args &A, uint64_t e, State &st) {
    double xs[12];
    for (int a = 0; a < 4; ++a) {
      const int32_t v = A.ep[4 * e + a];
      st.vid[a] = v;
      st.free_[a] = A.c[v] <= 0;
      for (int q = 0; q < 3; ++q) xs[3 * a + q] = A.x[3 * (int64_t)v + q];
    }
    double M[9], B[9];
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) M[3 * jj + ii] = xs[3 * ii + jj] - xs[9 + jj];
    inv3(M, B);
    st.W = -(1.0 / 6.0) * det3(M);
    for (int di = 0; di < 3; ++di) st.dN[9 + di] = 0.0;
    for (int vi = 0; vi < 3; ++vi)
      for (int di = 0; di < 3; ++di) {
        st.dN[3 * vi + di] = B[3 * vi + di];
        st.dN[9 + di] = st.dN[9 + di] - B[3 * vi + di];
      }
    st.mu = A.mu[e];
    st.lam = A.lam[e];
  }
  __device__ static void pair(const Args &A, const State &st, int a, Pair &P) {
    uint32_t fm = 0;
    for (int t = 0; t < 4; ++t) {
      P.vid[t] = st.vid[t];
      fm |= (st.free_[t] ? 1u : 0u) << t;
    }
    P.freemask = fm;
    P.fa = (fm >> a) & 1u;
    const double sw = A.p.stiff_scale * st.W, cl = sw * st.lam, cm = sw * st.mu;
    for (int i = 0; i < 3; ++i) {
      const double ga = pick4(st.dN + i, a, 3);
      P.u[i] = cl * ga;
      P.w[i] = cm * ga;
    }
    for (int q = 0; q < 12; ++q) P.g[q] = st.dN[q];
    P.W = st.W;
  }
  // blk[i][j] = u_i gb_j + w_j gb_i + [i==j] (w . gb)  (s * W * B'EB, closed form)
  __device__ static bool block(const Args &, const Pair &P, int b, double *out) {
    if (!P.fa || !((P.freemask >> b) & 1u)) return false;
    const double *gb = P.g + 3 * b;
    const double wg = P.w[0] * gb[0] + P.w[1] * gb[1] + P.w[2] * gb[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = P.u[i] * gb[j] + P.w[j] * gb[i];
        if (i == j) t = t + wg;
        out[3 * i + j] = t;
      }
    return true;
  }
  // lumped mass after the element's stiffness writes (fem_linear.sim:168-173)
  __device__ static bool extra(const Args &A, const Pair &P, double *out) {
    if (A.p.rho == 0.0) return false;
    const double mm = 0.25 * A.p.rho * P.W;
    for (int q = 0; q < 9; ++q) out[q] = (q % 4 == 0) ? mm : 0.0;
    return true;
  }
  __device__ static void rhs(const Args &A, const Pair &P, double *f) {
    if (!P.fa) return;
    const double wq = 0.25 * P.W;
    for (int q = 0; q < 3; ++q) f[q] += wq * A.p.load[q];
  }
  __device__ static void finish_diag(const Args &A, uint32_t i, double *blk, double *) {
    if (A.p.fixed_diag == 0.0 || A.c[i] <= 0) return;
    for (int q = 0; q < 9; ++q) blk[q] = ((q % 4 == 0) ? A.p.fixed_diag : 0.0) + blk[q];
  }
  __device__ static void isolated_rhs(const Args &, uint32_t, double *f) {
    f[0] = f[1] = f[2] = 0.0;
  }
  __device__ static void finish_block(const Args &, uint32_t, uint32_t, double *) {}
};

// scalar trilinear hex Poisson, corner a = dx + 2dy + 4dz
struct HexArgs {
  smh_hex_params p;
  const int32_t *ep;
  const int32_t *dir;
  const double *kappa;
};
struct HexOp {
  static constexpr int CARD = 8, BD = 1;
  static constexpr bool kHasRhs = true, kExtra = false;
  using Args = HexArgs;
  struct State {
    int32_t vid[8];
    uint32_t freemask;
    double ke;
  };
  struct Pair {
    int32_t vid[8];
    uint32_t freemask;
    int32_t a;
    double ke;
  };
  __device__ static void load(const Args &A, uint64_t e, State &st) {
    st.freemask = 0;
    for (int a = 0; a < 8; ++a) {
      const int32_t v = A.ep[8 * e + a];
      st.vid[a] = v;
      st.freemask |= (A.dir[v] <= 0 ? 1u : 0u) << a;
    }
    st.ke = A.kappa[e] * A.p.h;
  }
  __device__ static void pair(const Args &, const State &st, int a, Pair &P) {
    for (int t = 0; t < 8; ++t) P.vid[t] = st.vid[t];
    P.freemask = st.freemask;
    P.a = a;
    P.ke = st.ke;
  }
  __device__ static bool block(const Args &, const Pair &P, int b, double *out) {
    if (!((P.freemask >> P.a) & 1u) || !((P.freemask >> b) & 1u)) return false;
    const int d = __popc((unsigned)(P.a ^ b));
    const double kr = d == 0 ? (1.0 / 3.0) : (d == 1 ? 0.0 : -(1.0 / 12.0));
    out[0] = P.ke * kr;
    return true;
  }
  __device__ static bool extra(const Args &, const Pair &, double *) { return false; }
  __device__ static void rhs(const Args &A, const Pair &P, double *f) {
    if (!((P.freemask >> P.a) & 1u)) return;
    f[0] += A.p.source * (A.p.h * A.p.h * A.p.h) * 0.125;
  }
  __device__ static void finish_diag(const Args &A, uint32_t i, double *blk, double *) {
    if (A.dir[i] <= 0) return;
    blk[0] = A.p.fixed_diag + blk[0];
  }
  __device__ static void isolated_rhs(const Args &, uint32_t, double *f) { f[0] = 0.0; }
  __device__ static void finish_block(const Args &, uint32_t, uint32_t, double *) {}
};

// one per-edge block written into the four endpoint blocks with signs, then
// A = alpha*I + beta*A (test/input/program/cg.sim, assembly/matrix_vv.sim,
// system/gemv_blocked.sim)
struct EdgeArgs {
  const int32_t *ep;
  const double *w;
  double sign[4];
  double alpha, beta;
  int32_t affine;
};
template <int BDIM>
struct EdgeOp {
  static constexpr int CARD = 2, BD = BDIM;
  static constexpr bool kHasRhs = false, kExtra = false;
  using Args = EdgeArgs;
  struct State {
    int32_t vid[2];
    double w[BDIM * BDIM];
  };
  struct Pair {
    int32_t vid[2];
    int32_t a, pad;
    double w[BDIM * BDIM];
  };
  __device__ static void load(const Args &A, uint64_t e, State &st) {
    st.vid[0] = A.ep[2 * e];
    st.vid[1] = A.ep[2 * e + 1];
    for (int q = 0; q < BDIM * BDIM; ++q) st.w[q] = A.w[e * BDIM * BDIM + q];
  }
  __device__ static void pair(const Args &, const State &st, int a, Pair &P) {
    P.vid[0] = st.vid[0];
    P.vid[1] = st.vid[1];
    P.a = a;
    for (int q = 0; q < BDIM * BDIM; ++q) P.w[q] = st.w[q];
  }
  __device__ static bool block(const Args &A, const Pair &P, int b, double *out) {
    const double s = A.sign[2 * P.a + b];
    for (int q = 0; q < BDIM * BDIM; ++q) out[q] = s * P.w[q];
    return true;
  }
  __device__ static bool extra(const Args &, const Pair &, double *) { return false; }
  __device__ static void rhs(const Args &, const Pair &, double *) {}
  __device__ static void finish_diag(const Args &, uint32_t, double *, double *) {}
  __device__ static void isolated_rhs(const Args &, uint32_t, double *f) {
    for (int q = 0; q < BDIM; ++q) f[q] = 0.0;
  }
  __device__ static void finish_block(const Args &A, uint32_t row, uint32_t col, double *blk) {
    if (!A.affine) return;
    for (int i = 0; i < BDIM; ++i)
      for (int j = 0; j < BDIM; ++j)
        blk[i * BDIM + j] = ((row == col && i == j) ? A.alpha : 0.0) + A.beta * blk[i * BDIM + j];
  }
};

// Neo-Hookean tets, test/input/program/femTet.sim: main's two maps, A = map
// compute_stiffness(h) (:74-124) and f = map compute_force(h) (:37-72).  The
// arithmetic repeats the oracle's operation for operation (FP64, no fused
// multiply-add), log included, so SMH_F64 results are bit-identical.
struct FemTetArgs {
  smh_femtet_params p;
  const int32_t *ep;
  const double *x, *v, *u, *l, *W, *B;
  const int32_t *c;
};
// log x = k ln 2 + 2 atanh(s), x = m 2^k with m in [sqrt(1/2), sqrt(2)),
// s = (m - 1)/(m + 1): the series 2 s (1 + s^2/3 + ... + s^24/25)
__device__ double fem_log(double x) {
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
#pragma unroll
  for (int n = 11; n >= 0; --n) t = t * s2 + 1.0 / (double)(2 * n + 1);
  return (double)k * 0.69314718055994530942 + 2.0 * s * t;
}
// C = A*B and C = A*B' (k ascending)
__device__ __forceinline__ void fem_mm(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
__device__ __forceinline__ void fem_mmt(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] =
          A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
struct FemTetOp {
  static constexpr int CARD = 4, BD = 3;
  static constexpr bool kHasRhs = true, kExtra = true;
  using Args = FemTetArgs;
  struct State {
    int32_t vid[4];
    uint32_t freemask;
    double xs[12];
    double B[9];
    double u, l, W;
  };
  // the row's four blocks (0 where a pair is not free-free), the element's
  // mass term, and its writes to the row's force, in program order
  struct Pair {
    int32_t vid[4];
    uint32_t freemask;
    int32_t fa, nf, pad;
    double blk[4][9];
    double mq;
    double fw[4][3];
  };
  __device__ static void load(const Args &A, uint64_t e, State &st) {
    st.freemask = 0;
    for (int a = 0; a < 4; ++a) {
      const int32_t v = A.ep[4 * e + a];
      st.vid[a] = v;
      st.freemask |= (A.c[v] <= 0 ? 1u : 0u) << a;
      for (int q = 0; q < 3; ++q) st.xs[3 * a + q] = A.x[3 * (int64_t)v + q];
    }
    for (int q = 0; q < 9; ++q) st.B[q] = A.B[9 * e + q];
    st.u = A.u[e];
    st.l = A.l[e];
    st.W = A.W[e];
  }
  __device__ static void pair(const Args &A, const State &st, int a, Pair &P) {
    for (int t = 0; t < 4; ++t) P.vid[t] = st.vid[t];
    P.freemask = st.freemask;
    P.fa = (st.freemask >> a) & 1u;
    const double *B = st.B;
    const double u = st.u, l = st.l, W = st.W;
    // F = Ds*B, Ds(jj,ii) = x_ii(jj) - x_3(jj)
    double Ds[9], F[9], Fi[9], FiT[9];
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) Ds[3 * jj + ii] = st.xs[3 * ii + jj] - st.xs[9 + jj];
    fem_mm(Ds, B, F);
    const double JJ = fem_log(det3(F));
    inv3(F, Fi);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) FiT[3 * i + j] = Fi[3 * j + i];
    const double nW = -W;
    // compute_force: H = -W * PK1(F) * B'; f(v(a)) gets h*H(:,a) (a < 3) or
    // -h*H(:,ii) for ii = 0..2 (a = 3), then h*fg + m*v(a)
    {
      const double lJ = l * JJ;
      double Ps[9], H[9];
      for (int q = 0; q < 9; ++q) Ps[q] = nW * (u * (F[q] - FiT[q]) + lJ * FiT[q]);
      fem_mmt(Ps, B, H);
      const double h = A.p.h;
      const double mq = 0.25 * A.p.rho * W;
      P.mq = mq;
      int nf = 0;
      if (a < 3) {
        for (int q = 0; q < 3; ++q) P.fw[0][q] = h * H[3 * q + a];
        nf = 1;
      } else {
        for (int ii = 0; ii < 3; ++ii)
          for (int q = 0; q < 3; ++q) P.fw[ii][q] = -h * H[3 * q + ii];
        nf = 3;
      }
      const int64_t va = st.vid[a];
      for (int q = 0; q < 3; ++q) P.fw[nf][q] = h * (mq * A.p.grav[q]) + mq * A.v[3 * va + q];
      P.nf = nf + 1;
    }
    // compute_stiffness: block (a, b) = -(h*h) Kb(a,:,:) for the dF
    // directions (b, kk)
    const double hh = A.p.h * A.p.h;
    for (int b = 0; b < 4; ++b) {
      if (!P.fa || !((st.freemask >> b) & 1u)) {
        for (int q = 0; q < 9; ++q) P.blk[b][q] = 0.0;
        continue;
      }
      double dFrow[3];
      if (b < 3) {
        for (int ll = 0; ll < 3; ++ll) dFrow[ll] = B[3 * b + ll];
      } else {
        for (int ii = 0; ii < 3; ++ii) dFrow[ii] = -(B[ii] + B[3 + ii] + B[6 + ii]);
      }
      for (int kk = 0; kk < 3; ++kk) {
        double dF[9], FidF[9], T1[9], T2[9], dPs[9], dH[9];
        for (int q = 0; q < 9; ++q) dF[q] = 0.0;
        for (int ll = 0; ll < 3; ++ll) dF[3 * kk + ll] = dFrow[ll];
        fem_mm(Fi, dF, FidF);
        const double s1 = u - l * JJ;
        for (int q = 0; q < 9; ++q) T1[q] = s1 * FiT[q];
        fem_mmt(T1, FidF, T2);
        const double s2 = l * (FidF[0] + FidF[4] + FidF[8]);
        for (int q = 0; q < 9; ++q) dPs[q] = nW * (u * dF[q] + T2[q] + s2 * FiT[q]);
        fem_mmt(dPs, B, dH);
        for (int ll = 0; ll < 3; ++ll) {
          const double kb = a < 3 ? dH[3 * ll + a]
                                  : -(dH[3 * ll + 0] + dH[3 * ll + 1] + dH[3 * ll + 2]);
          P.blk[b][3 * ll + kk] = -(hh * kb);
        }
      }
    }
  }
  // K(v(jj), v(row)) is written iff both are free (femTet.sim:111-117)
  __device__ static bool block(const Args &, const Pair &P, int b, double *out) {
    if (!P.fa || !((P.freemask >> b) & 1u)) return false;
    for (int q = 0; q < 9; ++q) out[q] = P.blk[b][q];
    return true;
  }
  // the lumped mass, on every diagonal after the stiffness (:119-123)
  __device__ static bool extra(const Args &, const Pair &P, double *out) {
    for (int q = 0; q < 9; ++q) out[q] = P.mq * ((q % 4 == 0) ? 1.0 : 0.0);
    return true;
  }
  __device__ static void rhs(const Args &, const Pair &P, double *f) {
    if (!P.fa) return;
    for (int k = 0; k < P.nf; ++k)
      for (int q = 0; q < 3; ++q) f[q] += P.fw[k][q];
  }
  __device__ static void finish_diag(const Args &, uint32_t, double *, double *) {}
  __device__ static void isolated_rhs(const Args &, uint32_t, double *f) {
    f[0] = f[1] = f[2] = 0.0;
  }
  __device__ static void finish_block(const Args &, uint32_t, uint32_t, double *) {}
};

// initializeTet (femTet.sim:126-148): per tet B = 1.0*inv(M), W = -(1/6)
// det(M); then m of each vertex from the last (highest) tet holding it
__global__ void k_femtet_geom(uint64_t nt, const int32_t *__restrict__ ep,
                              const double *__restrict__ x, double *__restrict__ W,
                              double *__restrict__ Bout) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double xs[12], M[9], Mi[9];
  for (int a = 0; a < 4; ++a) {
    const int32_t v = ep[4 * t + a];
    for (int q = 0; q < 3; ++q) xs[3 * a + q] = x[3 * (int64_t)v + q];
  }
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj) M[3 * jj + ii] = xs[3 * ii + jj] - xs[9 + jj];
  inv3(M, Mi);
  for (int q = 0; q < 9; ++q) Bout[9 * t + q] = 1.0 * Mi[q];
  W[t] = -(1.0 / 6.0) * det3(M);
}
__global__ void k_femtet_mass(uint32_t nv, double rho, const uint32_t *__restrict__ ve_ptr,
                              const uint32_t *__restrict__ ve_elem, const double *__restrict__ W,
                              double *__restrict__ m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv || ve_ptr[i] == ve_ptr[i + 1]) return;
  m[i] = 0.25 * rho * W[ve_elem[ve_ptr[i + 1] - 1]];
}

// ---- the pull kernel --------------------------------------------------------
template <int PREC>
__device__ __forceinline__ void store_val(void *vals, void *lo, uint64_t i, double v) {
  if (PREC == SMH_F64) {
    ((double *)vals)[i] = v;
  } else if (PREC == SMH_F32) {
    ((float *)vals)[i] = (float)v;
  } else {
    const float h = (float)v;
    ((float *)vals)[i] = h;
    ((float *)lo)[i] = (float)(v - (double)h);
  }
}

template <class Op, int PREC>
__global__ void __launch_bounds__(kWave)
k_assemble(typename Op::Args A, uint32_t nrows, uint64_t nnz,
           const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ colidx,
           const uint32_t *__restrict__ wave_row,
           const uint32_t *__restrict__ ve_ptr, const uint32_t *__restrict__ ve_elem,
           void *__restrict__ vals, void *__restrict__ vals_lo, double *__restrict__ rhs,
           uint64_t nwaves) {
  constexpr int BD = Op::BD, BB = BD * BD, CARD = Op::CARD;
  using Pair = typename Op::Pair;
  // per staged pair: its record and flags (bit 0: the row's slot was found);
  // after the last chunk the record area stages the coalesced output stores
  constexpr size_t kPairBytes = sizeof(Pair) * kWave;
  constexpr size_t kOutBytes = sizeof(double) * kWave * BB;
  __shared__ __align__(16) unsigned char smem[kPairBytes > kOutBytes ? kPairBytes : kOutBytes];
  Pair *PP = reinterpret_cast<Pair *>(smem);
  __shared__ uint32_t M[kWave];
  __shared__ uint32_t R[kWave + 1];    // rowptr window
  __shared__ uint32_t VP[kWave + 2];   // ve_ptr window of the wave's rows

  const int lane = threadIdx.x;
  // Persistent grid (a multiple of 8 workgroups).  Workgroups are dealt to
  // the 8 XCDs round-robin; XCD x works through the x-th eighth of the waves,
  // so neighbouring waves, which gather the same vertices and elements, share
  // that XCD's L2.
  const uint32_t xcd = blockIdx.x & 7u, per = gridDim.x >> 3;
  const uint64_t w1 = nwaves * (xcd + 1) / 8;
  for (uint64_t w = nwaves * xcd / 8 + (blockIdx.x >> 3); w < w1; w += per) {
  const uint64_t b0 = w * kWave;
  const uint64_t bk = b0 + lane;
  const bool valid = bk < nnz;
  const uint64_t blast = min(b0 + kWave, nnz) - 1;

  // row of b0 (last r with rowptr[r] <= b0), precomputed with the index
  const uint32_t rfirst = wave_row[w];
  R[lane] = rowptr[min(rfirst + (uint32_t)lane, nrows)];
  if (lane == 0) R[kWave] = rowptr[min(rfirst + (uint32_t)kWave, nrows)];
  __syncthreads();
  uint32_t row;
  if ((uint64_t)R[kWave] > blast) {
    // the wave's rows lie in rfirst .. rfirst+63: last l with R[l] <= bk
    int lo = 0, hi = kWave + 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((uint64_t)R[mid] <= bk) lo = mid; else hi = mid;
    }
    row = rfirst + (uint32_t)lo;
    if (row >= nrows) row = nrows - 1;
  } else {
    uint32_t lo = rfirst, hi = nrows;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)rowptr[mid] <= bk) lo = mid; else hi = mid;
    }
    row = lo;
  }
  uint32_t rlast = valid ? row : rfirst;
  for (int off = 32; off > 0; off >>= 1) rlast = max(rlast, (uint32_t)__shfl_xor((int)rlast, off, 64));
  const uint32_t col = valid ? colidx[bk] : 0xffffffffu;
  const bool is_diag = valid && col == row;
  const uint32_t nwr = rlast - rfirst + 1;  // rows of the wave
  const bool vp_lds = nwr <= (uint32_t)kWave;
  if (vp_lds)
    for (uint32_t l = lane; l <= nwr; l += kWave) VP[l] = ve_ptr[rfirst + l];
  __syncthreads();

  double acc[BB];
  for (int q = 0; q < BB; ++q) acc[q] = 0.0;
  double f[BD];
  for (int q = 0; q < BD; ++q) f[q] = 0.0;

  const uint32_t P0 = vp_lds ? VP[0] : ve_ptr[rfirst];
  const uint32_t P1 = vp_lds ? VP[nwr] : ve_ptr[rlast + 1];
  const uint32_t myb = valid ? ve_ptr[row] : 0, mye = valid ? ve_ptr[row + 1] : 0;
  for (uint32_t c0 = P0; c0 < P1; c0 += kWave) {
    // phase A: lane l builds the record of pair p = c0 + l
    const uint32_t p = c0 + lane;
    uint32_t ok = 0;
    if (p < P1) {
      const uint32_t e = ve_elem[p];
      // the row this pair belongs to: last r with ve_ptr[r] <= p
      uint32_t pr;
      if (vp_lds) {
        int lo = 0, hi = (int)nwr;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (VP[mid] <= p) lo = mid; else hi = mid;
        }
        pr = rfirst + (uint32_t)lo;
      } else {
        uint32_t lo = rfirst, hi = rlast + 1;
        while (hi - lo > 1) {
          const uint32_t mid = lo + (hi - lo) / 2;
          if (ve_ptr[mid] <= p) lo = mid; else hi = mid;
        }
        pr = lo;
      }
      const uint32_t prb = vp_lds ? VP[pr - rfirst] : ve_ptr[pr];
      // an element holding the row k times is listed k times in a row: the
      // occ-th copy stands for the occ-th endpoint slot equal to the row
      int occ = 0;
      while (p - occ > prb && ve_elem[p - occ - 1] == e) ++occ;
      typename Op::State st;
      Op::load(A, e, st);
      int a = -1, seen = 0;
      for (int t = 0; t < CARD; ++t)
        if ((uint32_t)st.vid[t] == pr && a < 0) {
          if (seen == occ) a = t;
          ++seen;
        }
      if (a >= 0) {
        Op::pair(A, st, a, PP[lane]);
        ok = 1;
      }
    }
    M[lane] = ok;
    __syncthreads();
    // phase B: every block lane adds its contributions in list order
    if (valid) {
      const uint32_t lo = max(myb, c0), hi = min(mye, c0 + kWave);
      for (uint32_t q = lo; q < hi; ++q) {
        const uint32_t l = q - c0;
        if (!M[l]) continue;
        const Pair &P = PP[l];
        double blk[BB];
        // slots of the element that hold this lane's column, ascending (the
        // element function's b order); normally exactly one, so the wave
        // forms one block per pair instead of one per slot
        uint32_t match = 0;
        for (int b = 0; b < CARD; ++b) match |= ((uint32_t)P.vid[b] == col ? 1u : 0u) << b;
        while (match) {
          const int b = __builtin_ctz(match);
          match &= match - 1;
          if (Op::block(A, P, b, blk))
            for (int k = 0; k < BB; ++k) acc[k] += blk[k];
        }
        if (is_diag) {
          if (Op::kExtra && Op::extra(A, P, blk))
            for (int k = 0; k < BB; ++k) acc[k] += blk[k];
          Op::rhs(A, P, f);  // the element's writes to the row's f, in order
        }
      }
    }
    __syncthreads();
  }
  if (valid) Op::finish_block(A, row, col, acc);
  if (is_diag) {
    Op::finish_diag(A, row, acc, f);
    if (rhs)
      for (int q = 0; q < BD; ++q) rhs[(uint64_t)row * BD + q] = f[q];
  }
  // coalesced store through LDS
  double *outbuf = reinterpret_cast<double *>(smem);
  for (int q = 0; q < BB; ++q) outbuf[lane * BB + q] = acc[q];
  __syncthreads();
  const uint64_t nval = (min(b0 + kWave, nnz) - b0) * BB;
  for (int q = 0; q < BB; ++q) {
    const uint64_t i = (uint64_t)q * kWave + lane;
    if (i < nval) store_val<PREC>(vals, vals_lo, b0 * BB + i, outbuf[i]);
  }
  __syncthreads();  // LDS is reused by the next wave of blocks
  }
}

// right-hand side of vertices that belong to no element (empty rows)
template <class Op>
__global__ void k_isolated(typename Op::Args A, uint32_t nrows, const uint32_t *__restrict__ rowptr,
                           double *__restrict__ rhs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!rhs || i >= nrows || rowptr[i] != rowptr[i + 1]) return;
  double f[Op::BD];
  Op::isolated_rhs(A, i, f);
  for (int q = 0; q < Op::BD; ++q) rhs[(uint64_t)i * Op::BD + q] = f[q];
}

// Resident workgroups per CU for the persistent assembly grid: the
// occupancy limit of the instantiation (registers, LDS), or
// SMH_ASSEMBLY_WGS_PER_CU when set (for tuning).
template <class Op, int P>
unsigned assembly_wgs_per_cu() {
  static const unsigned n = [] {
    if (const char *e = getenv("SMH_ASSEMBLY_WGS_PER_CU")) {
      const int v = atoi(e);
      if (v > 0) return (unsigned)v;
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_assemble<Op, P>, kWave, 0) !=
            hipSuccess ||
        occ < 1) {
      (void)hipGetLastError();
      occ = 4;
    }
    return (unsigned)occ;
  }();
  return n;
}

template <class Op>
int launch_assembly(smh_ctx *ctx, const smh_index *idx, const int32_t *ep, const typename Op::Args &A,
                    int32_t prec, void *vals, void *vals_lo, double *rhs, const char *name) {
  using namespace smh;
  SMH_CHECK(ctx, idx && vals && (rhs || !Op::kHasRhs), SMH_ERR_INVALID,
            "%s: null index, vals or rhs", name);
  SMH_CHECK(ctx, prec == SMH_F64 || prec == SMH_F32 || prec == SMH_F32X2, SMH_ERR_INVALID,
            "%s: bad precision %d", name, prec);
  SMH_CHECK(ctx, prec != SMH_F32X2 || vals_lo, SMH_ERR_INVALID, "%s: SMH_F32X2 needs vals_lo",
            name);
  int set = -1;
  for (uint32_t s = 0; s < idx->nsets; ++s)
    if (idx->ep[s] == ep && idx->card[s] == (uint32_t)Op::CARD) set = (int)s;
  SMH_CHECK(ctx, set >= 0, SMH_ERR_INVALID,
            "%s: the element set's endpoints are not one of the index's edge sets", name);
  SMH_CHECK(ctx, idx->nnz * Op::BD * Op::BD < 0xffffffffull * 16, SMH_ERR_UNSUPPORTED,
            "%s: too many values", name);
  const uint32_t nrows = idx->nrows;
  if (nrows == 0) return SMH_OK;
  const uint64_t nwaves = (idx->nnz + kWave - 1) / kWave;
  if (nwaves) {
    // grid: a multiple of 8 (one share per XCD), no more than fills the chip
    auto grid = [&](unsigned per_cu) {
      const uint64_t full = (uint64_t)ctx->num_cu * per_cu;
      const uint64_t g = std::min<uint64_t>(full, nwaves);
      return (unsigned)((g + 7) / 8 * 8);
    };
#define SMH_ASM(P)                                                                          \
  hipLaunchKernelGGL((k_assemble<Op, P>), dim3(grid(assembly_wgs_per_cu<Op, P>())),         \
                     dim3(kWave), 0, ctx->stream, A, nrows, idx->nnz, idx->rowptr,          \
                     idx->colidx, idx->wave_row, idx->ve_ptr[set], idx->ve_elem[set], vals, \
                     vals_lo, rhs, nwaves)
    if (prec == SMH_F64) SMH_ASM(SMH_F64);
    else if (prec == SMH_F32) SMH_ASM(SMH_F32);
    else SMH_ASM(SMH_F32X2);
#undef SMH_ASM
  }
  hipLaunchKernelGGL(k_isolated<Op>, dim3(grid_for(nrows, 256)), dim3(256), 0, ctx->stream, A,
                     nrows, idx->rowptr, rhs);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

}  // namespace

extern "C" int smh_assemble_springs(smh_ctx *ctx, const smh_index *idx, const smh_springs_params *p,
                                    const smh_springs_set *s, int32_t prec, void *vals,
                                    void *vals_lo, double *rhs) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, p && s, SMH_ERR_INVALID, "smh_assemble_springs: null params or set");
  SMH_CHECK(ctx, idx && s->npoints == idx->nrows, SMH_ERR_INVALID,
            "smh_assemble_springs: point count does not match the index rows");
  SMH_CHECK(ctx, s->x && s->v && s->m && s->fixed && (s->nsprings == 0 || (s->k && s->l0)),
            SMH_ERR_INVALID, "smh_assemble_springs: null field");
  SpringArgs A{*p, s->endpoints, s->x, s->v, s->m, s->k, s->l0, s->fixed};
  return launch_assembly<SpringOp>(ctx, idx, s->endpoints, A, prec, vals, vals_lo, rhs,
                                   "smh_assemble_springs");
}

extern "C" int smh_assemble_tet_linear(smh_ctx *ctx, const smh_index *idx, const smh_tet_params *p,
                                       const smh_tet_set *s, int32_t prec, void *vals,
                                       void *vals_lo, double *rhs) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, p && s, SMH_ERR_INVALID, "smh_assemble_tet_linear: null params or set");
  SMH_CHECK(ctx, idx && s->nverts == idx->nrows, SMH_ERR_INVALID,
            "smh_assemble_tet_linear: vertex count does not match the index rows");
  SMH_CHECK(ctx, s->x && s->c && (s->ntets == 0 || (s->mu && s->lam)), SMH_ERR_INVALID,
            "smh_assemble_tet_linear: null field");
  TetArgs A{*p, s->endpoints, s->x, s->mu, s->lam, s->c};
  return launch_assembly<TetOp>(ctx, idx, s->endpoints, A, prec, vals, vals_lo, rhs,
                                "smh_assemble_tet_linear");
}

extern "C" int smh_assemble_hex_poisson(smh_ctx *ctx, const smh_index *idx, const smh_hex_params *p,
                                        const smh_hex_set *s, int32_t prec, void *vals,
                                        void *vals_lo, double *rhs) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, p && s, SMH_ERR_INVALID, "smh_assemble_hex_poisson: null params or set");
  SMH_CHECK(ctx, idx && s->nverts == idx->nrows, SMH_ERR_INVALID,
            "smh_assemble_hex_poisson: vertex count does not match the index rows");
  SMH_CHECK(ctx, s->dirichlet && (s->nhex == 0 || s->kappa), SMH_ERR_INVALID,
            "smh_assemble_hex_poisson: null field");
  HexArgs A{*p, s->endpoints, s->dirichlet, s->kappa};
  return launch_assembly<HexOp>(ctx, idx, s->endpoints, A, prec, vals, vals_lo, rhs,
                                "smh_assemble_hex_poisson");
}

extern "C" int smh_assemble_edge_blocks(smh_ctx *ctx, const smh_index *idx, const smh_edge_blocks *e,
                                        int32_t prec, void *vals, void *vals_lo) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, e && idx && e->nverts == idx->nrows, SMH_ERR_INVALID,
            "smh_assemble_edge_blocks: null argument or vertex count mismatch");
  SMH_CHECK(ctx, e->nedges == 0 || e->w, SMH_ERR_INVALID, "smh_assemble_edge_blocks: null w");
  EdgeArgs A{e->endpoints, e->w, {e->sign[0], e->sign[1], e->sign[2], e->sign[3]}, e->alpha,
             e->beta, !(e->alpha == 0.0 && e->beta == 1.0)};
  switch (e->bdim) {
    case 1:
      return launch_assembly<EdgeOp<1>>(ctx, idx, e->endpoints, A, prec, vals, vals_lo, nullptr,
                                        "smh_assemble_edge_blocks");
    case 2:
      return launch_assembly<EdgeOp<2>>(ctx, idx, e->endpoints, A, prec, vals, vals_lo, nullptr,
                                        "smh_assemble_edge_blocks");
    case 3:
      return launch_assembly<EdgeOp<3>>(ctx, idx, e->endpoints, A, prec, vals, vals_lo, nullptr,
                                        "smh_assemble_edge_blocks");
  }
  return smh::fail(ctx, SMH_ERR_UNSUPPORTED, "smh_assemble_edge_blocks: bdim %u (1..3)", e->bdim);
}

extern "C" int smh_assemble_tet_neohookean(smh_ctx *ctx, const smh_index *idx,
                                           const smh_femtet_params *p, const smh_femtet_set *s,
                                           int32_t prec, void *vals, void *vals_lo, double *rhs) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, p && s, SMH_ERR_INVALID, "smh_assemble_tet_neohookean: null params or set");
  SMH_CHECK(ctx, idx && s->nverts == idx->nrows, SMH_ERR_INVALID,
            "smh_assemble_tet_neohookean: vertex count does not match the index rows");
  SMH_CHECK(ctx, s->x && s->v && s->c &&
                     (s->ntets == 0 || (s->u && s->l && s->W && s->B)),
            SMH_ERR_INVALID, "smh_assemble_tet_neohookean: null field");
  FemTetArgs A{*p, s->endpoints, s->x, s->v, s->u, s->l, s->W, s->B, s->c};
  return launch_assembly<FemTetOp>(ctx, idx, s->endpoints, A, prec, vals, vals_lo, rhs,
                                   "smh_assemble_tet_neohookean");
}

extern "C" int smh_femtet_precompute(smh_ctx *ctx, const smh_index *idx,
                                     const smh_femtet_params *p, const smh_femtet_set *s,
                                     double *W, double *B, double *m) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, p && s && idx && W && B && m && s->x, SMH_ERR_INVALID,
            "smh_femtet_precompute: null argument");
  SMH_CHECK(ctx, s->nverts == idx->nrows, SMH_ERR_INVALID,
            "smh_femtet_precompute: vertex count does not match the index rows");
  int set = -1;
  for (uint32_t q = 0; q < idx->nsets; ++q)
    if (idx->ep[q] == s->endpoints && idx->card[q] == 4u) set = (int)q;
  SMH_CHECK(ctx, set >= 0, SMH_ERR_INVALID,
            "smh_femtet_precompute: the tets' endpoints are not one of the index's edge sets");
  if (s->ntets)
    hipLaunchKernelGGL(k_femtet_geom, dim3(smh::grid_for(s->ntets, 256)), dim3(256), 0,
                       ctx->stream, s->ntets, s->endpoints, s->x, W, B);
  if (s->nverts)
    hipLaunchKernelGGL(k_femtet_mass, dim3(smh::grid_for(s->nverts, 256)), dim3(256), 0,
                       ctx->stream, s->nverts, p->rho, idx->ve_ptr[set], idx->ve_elem[set],
                       (const double *)W, m);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}
