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
#include <type_traits>
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
// The *_entry() forms give one value q of a block (q = r*BD + s) or one
// component r of the right-hand side with exactly the arithmetic of the full
// forms, for the lanes that sum a diagonal block entry by entry.

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
  __device__ static bool block_entry(const Args &, const Pair &P, int b, int q, double &v) {
    if (!P.fa || (P.a != b && !P.fb)) return false;
    v = P.a == b ? P.kb[q] : -P.kb[q];
    return true;
  }
  __device__ static bool extra_entry(const Args &, const Pair &, int, double &) { return false; }
  __device__ static void rhs_entry(const Args &, const Pair &P, int r, double &f) {
    if (!P.fa) return;
    f += P.a == 0 ? P.fe0[r] : -P.fe0[r];
  }
  __device__ static void finish_diag_entry(const Args &A, uint32_t i, int q, double &v) {
    double md;
    if (A.fixed[i]) md = (A.p.fixed_identity && (q % 4 != 0)) ? 0.0 : A.p.fixed_fill;
    else md = (q % 4 == 0) ? A.m[i] + A.p.h * A.p.viscous : 0.0;
    v = md + v;
  }
  __device__ static void finish_rhs_entry(const Args &A, uint32_t i, int r, double &f) {
    double fmg = 0.0;
    if (!A.fixed[i]) {
      const double mass = A.m[i];
      fmg = mass * A.v[3 * (int64_t)i + r] + A.p.h * mass * A.p.grav[r];
    }
    f = fmg + f;
  }
  __device__ static void finish_block_entry(const Args &, uint32_t, uint32_t, int, double &) {}
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
  // the four shape gradients, the row's own g_a first and then the other
  // slots ascending; the row's slot a and the coefficients cl = s*W*lam,
  // cm = s*W*mu of u = cl g_a and w = cm g_a (144 bytes)
  static constexpr uint32_t kFa = 16u;
  struct Pair {
    int32_t vid[4];
    // freemask (bits 0-3), fa (bit 4), a (bits 8-9), freemask in stored
    // order (bits 12-15: the own slot, then the others ascending)
    uint32_t bits;
    int32_t pad;
    double cl, cm, W;
    double g[12];
    __device__ uint32_t freemask() const { return bits & 15u; }
    __device__ bool fa() const { return bits & kFa; }
    __device__ int a() const { return (int)((bits >> 8) & 3u); }
    // where slot b's gradient is stored
    __device__ int pos(int b) const { return b == a() ? 0 : (b < a() ? b + 1 : b); }
    __device__ double u(int i) const { return cl * g[i]; }
    __device__ double w(int i) const { return cm * g[i]; }
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
    uint32_t fmp = (fm >> a) & 1u;
    for (int t = 0; t < 4; ++t)
      if (t != a) fmp |= ((fm >> t) & 1u) << (t < a ? t + 1 : t);
    P.bits = fm | (((fm >> a) & 1u) << 4) | ((uint32_t)a << 8) | (fmp << 12);
    const double sw = A.p.stiff_scale * st.W;
    P.cl = sw * st.lam;
    P.cm = sw * st.mu;
    P.W = st.W;
    for (int i = 0; i < 3; ++i) P.g[i] = pick4(st.dN + i, a, 3);
    for (int t = 0; t < 4; ++t) {
      if (t == a) continue;
      const int o = 3 * (t < a ? t + 1 : t);
      for (int i = 0; i < 3; ++i) P.g[o + i] = st.dN[3 * t + i];
    }
  }
  // blk[i][j] = u_i gb_j + w_j gb_i + [i==j] (w . gb)  (s * W * B'EB, closed form)
  __device__ static bool block(const Args &, const Pair &P, int b, double *out) {
    if (!P.fa() || !((P.freemask() >> b) & 1u)) return false;
    const double *gb = P.g + 3 * P.pos(b);
    double u[3], w[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = P.u(i);
      w[i] = P.w(i);
    }
    const double wg = w[0] * gb[0] + w[1] * gb[1] + w[2] * gb[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = u[i] * gb[j] + w[j] * gb[i];
        if (i == j) t = t + wg;
        out[3 * i + j] = t;
      }
    return true;
  }
  // block() for the slot of rank r among the slots other than the row's
  // (stored at position r + 1): the fast kernel's plan names r, so the
  // gradients are read at offsets known before the record is
  __device__ static bool block_rank(const Args &, const Pair &P, int r, double *out) {
    const uint32_t bits = P.bits;
    const double cl = P.cl, cm = P.cm;
    const double g0 = P.g[0], g1 = P.g[1], g2 = P.g[2];
    const double *gb = P.g + 3 * (r + 1);
    const double b0 = gb[0], b1 = gb[1], b2 = gb[2];
    if (!(bits & kFa) || !((bits >> (13 + r)) & 1u)) return false;
    const double u[3] = {cl * g0, cl * g1, cl * g2}, w[3] = {cm * g0, cm * g1, cm * g2};
    const double gbv[3] = {b0, b1, b2};
    const double wg = w[0] * b0 + w[1] * b1 + w[2] * b2;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = u[i] * gbv[j] + w[j] * gbv[i];
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
    if (!P.fa()) return;
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
  __device__ static bool block_entry(const Args &, const Pair &P, int b, int q, double &v) {
    if (!P.fa() || !((P.freemask() >> b) & 1u)) return false;
    const double *gb = P.g + 3 * P.pos(b);
    const int i = q / 3, j = q - 3 * (q / 3);
    double t = P.u(i) * gb[j] + P.w(j) * gb[i];
    if (i == j) t = t + (P.w(0) * gb[0] + P.w(1) * gb[1] + P.w(2) * gb[2]);
    v = t;
    return true;
  }
  __device__ static bool extra_entry(const Args &A, const Pair &P, int q, double &v) {
    if (A.p.rho == 0.0) return false;
    v = (q % 4 == 0) ? 0.25 * A.p.rho * P.W : 0.0;
    return true;
  }
  __device__ static void rhs_entry(const Args &A, const Pair &P, int r, double &f) {
    if (!P.fa()) return;
    // (selects keep the kernel argument a uniform load for a per-lane r)
    const double L = r == 0 ? A.p.load[0] : (r == 1 ? A.p.load[1] : A.p.load[2]);
    f += 0.25 * P.W * L;
  }
  __device__ static void finish_diag_entry(const Args &A, uint32_t i, int q, double &v) {
    if (A.p.fixed_diag == 0.0 || A.c[i] <= 0) return;
    v = ((q % 4 == 0) ? A.p.fixed_diag : 0.0) + v;
  }
  __device__ static void finish_rhs_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_block_entry(const Args &, uint32_t, uint32_t, int, double &) {}
  // After the off-diagonal blocks are formed, the pair's record is turned
  // in place into its writes to the row's diagonal: g[q] = block(a, a)[q]
  // (0 when not written), g[9 + r] = the rhs term of component r, cl = the
  // lumped mass (added to q = 0, 4, 8 when rho != 0).  Same arithmetic as
  // block()/extra()/rhs(); adding the 0 of an unwritten term leaves a sum
  // unchanged (a sum that starts at +0 is never -0).
  __device__ static void to_diag(const Args &A, Pair &P) {
    const double cl = P.cl, cm = P.cm, W = P.W;
    const bool fa = P.bits & kFa;
    const double g0 = P.g[0], g1 = P.g[1], g2 = P.g[2];
    const double g[3] = {g0, g1, g2};
    const double wg = (cm * g0) * g0 + (cm * g1) * g1 + (cm * g2) * g2;
    double D[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = (cl * g[i]) * g[j] + (cm * g[j]) * g[i];
        if (i == j) t = t + wg;
        D[3 * i + j] = fa ? t : 0.0;
      }
    const double wq = 0.25 * W;
    for (int q = 0; q < 9; ++q) P.g[q] = D[q];
    for (int r = 0; r < 3; ++r) P.g[9 + r] = fa ? wq * A.p.load[r] : 0.0;
    P.cl = 0.25 * A.p.rho * W;
  }
  // entry q of the diagonal (and rhs component q/4 for q = 0, 4, 8) from a
  // record made by to_diag, as its loads and its adds: the fast kernel
  // issues several records' loads before their adds (the same adds, in the
  // same order)
  struct DiagTerm {
    double v, c, f;
  };
  __device__ static DiagTerm diag_load(const Pair &P, int q) {
    DiagTerm t;
    t.v = P.g[q];
    t.c = q % 4 == 0 ? P.cl : 0.0;
    t.f = q % 4 == 0 ? P.g[9 + q / 4] : 0.0;
    return t;
  }
  __device__ static void diag_add(const Args &A, const DiagTerm &t, int q, double &acc,
                                  double &f) {
    acc += t.v;
    if (q % 4 == 0) {
      if (A.p.rho != 0.0) acc += t.c;
      f += t.f;
    }
  }
  __device__ static void diag_sum(const Args &A, const Pair &P, int q, double &acc, double &f) {
    diag_add(A, diag_load(P, q), q, acc, f);
  }
};

// ops that turn a record into its diagonal terms (to_diag / diag_sum)
template <class Op, class = void>
struct HasToDiag : std::false_type {};
template <class Op>
struct HasToDiag<Op, std::void_t<decltype(&Op::to_diag)>> : std::true_type {};
// ops whose diag_sum comes as diag_load + diag_add
template <class Op, class = void>
struct HasDiagLoad : std::false_type {};
template <class Op>
struct HasDiagLoad<Op, std::void_t<decltype(&Op::diag_load)>> : std::true_type {};
// ops whose off-diagonal blocks are formed from the plan's slot ranks
template <class Op, class = void>
struct HasBlockRank : std::false_type {};
template <class Op>
struct HasBlockRank<Op, std::void_t<decltype(&Op::block_rank)>> : std::true_type {};

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
  __device__ static bool block_entry(const Args &A, const Pair &P, int b, int, double &v) {
    return block(A, P, b, &v);
  }
  __device__ static bool extra_entry(const Args &, const Pair &, int, double &) { return false; }
  __device__ static void rhs_entry(const Args &A, const Pair &P, int, double &f) { rhs(A, P, &f); }
  __device__ static void finish_diag_entry(const Args &A, uint32_t i, int, double &v) {
    if (A.dir[i] <= 0) return;
    v = A.p.fixed_diag + v;
  }
  __device__ static void finish_rhs_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_block_entry(const Args &, uint32_t, uint32_t, int, double &) {}
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
  // A.sign[k] by selects (a uniform kernel-argument load for a per-lane k)
  __device__ static double sign_of(const Args &A, int k) {
    const double s01 = k == 0 ? A.sign[0] : A.sign[1];
    const double s23 = k == 2 ? A.sign[2] : A.sign[3];
    return k < 2 ? s01 : s23;
  }
  __device__ static bool block(const Args &A, const Pair &P, int b, double *out) {
    const double s = sign_of(A, 2 * P.a + b);
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
  __device__ static bool block_entry(const Args &A, const Pair &P, int b, int q, double &v) {
    v = sign_of(A, 2 * P.a + b) * P.w[q];
    return true;
  }
  __device__ static bool extra_entry(const Args &, const Pair &, int, double &) { return false; }
  __device__ static void rhs_entry(const Args &, const Pair &, int, double &) {}
  __device__ static void finish_diag_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_rhs_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_block_entry(const Args &A, uint32_t row, uint32_t col, int q,
                                            double &v) {
    if (!A.affine) return;
    const int i = q / BDIM, j = q - BDIM * (q / BDIM);
    v = ((row == col && i == j) ? A.alpha : 0.0) + A.beta * v;
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
  __device__ static bool block_entry(const Args &, const Pair &P, int b, int q, double &v) {
    if (!P.fa || !((P.freemask >> b) & 1u)) return false;
    v = P.blk[b][q];
    return true;
  }
  __device__ static bool extra_entry(const Args &, const Pair &P, int q, double &v) {
    v = P.mq * ((q % 4 == 0) ? 1.0 : 0.0);
    return true;
  }
  __device__ static void rhs_entry(const Args &, const Pair &P, int r, double &f) {
    if (!P.fa) return;
    for (int k = 0; k < P.nf; ++k) f += P.fw[k][r];
  }
  __device__ static void finish_diag_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_rhs_entry(const Args &, uint32_t, int, double &) {}
  __device__ static void finish_block_entry(const Args &, uint32_t, uint32_t, int, double &) {}
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
// A wave the fast kernel cannot take: marked by the plan (nd == 255: its
// rows do not fit the 64-row window, a row has more pairs than mask bits, an
// element holds a row twice), or its diagonal blocks have more than 128
// entries to sum.
__device__ __forceinline__ bool generic_wave(uint32_t nd, int bb) {
  return nd == 255u || nd * (uint32_t)bb > 2u * kWave;
}
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
           uint64_t nwaves, const uint8_t *__restrict__ wnd) {
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
  if (wnd && !generic_wave(wnd[w], BB)) continue;  // the fast kernel's wave
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

// ---- the per-pattern plan and the fast kernel ----------------------------------
// Row i's pairs are ve_elem[ve_ptr[i] .. ve_ptr[i+1]), elements ascending (an
// element holding i k times is listed k times in a row).  Per pair: the
// endpoint slot it stands for (the occ-th slot equal to i), with bit 7 set
// when the element repeats a vertex and bit 6 when one of the row's blocks
// has more contributions than the rank list holds (their waves take the
// generic kernel).  Per off-diagonal block (i, j): bit k set when the
// element of the row's k-th pair holds j, and (rank != null, elements of at
// most 4 slots) for its contributions in bit order 2 bits each: the rank of
// j's slot b among the element's slots other than the row's slot a
// (b - (b > a)).  One thread per row, which owns the row's blocks.
template <typename MT>
__global__ void k_asm_plan(uint32_t nrows, uint32_t card, const int32_t *__restrict__ ep,
                           const uint32_t *__restrict__ rowptr,
                           const uint32_t *__restrict__ colidx,
                           const uint32_t *__restrict__ ve_ptr,
                           const uint32_t *__restrict__ ve_elem, uint8_t *__restrict__ slot,
                           MT *__restrict__ mask, uint32_t *__restrict__ rank) {
  constexpr uint32_t kBits = 8 * sizeof(MT);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const uint32_t c0 = rowptr[i], c1 = rowptr[i + 1];
  for (uint32_t q = c0; q < c1; ++q) mask[q] = 0;
  const uint32_t rb = ve_ptr[i], re = ve_ptr[i + 1];
  uint32_t prev = 0xffffffffu;
  uint32_t occ = 0;
  for (uint32_t p = rb; p < re; ++p) {
    const uint32_t e = ve_elem[p];
    occ = e == prev ? occ + 1 : 0;
    prev = e;
    const int32_t *v = ep + (uint64_t)e * card;
    uint32_t a = 255, seen = 0;
    bool repeat = false;
    for (uint32_t t = 0; t < card; ++t) {
      if ((uint32_t)v[t] == i) {
        if (seen == occ) a = t;
        ++seen;
      }
      for (uint32_t t2 = 0; t2 < t; ++t2) repeat = repeat || v[t2] == v[t];
    }
    slot[p] = (uint8_t)(repeat ? (a | 0x80u) : a);
    const uint32_t k = p - rb;
    if (k >= kBits) continue;
    for (uint32_t t = 0; t < card; ++t) {
      const uint32_t j = (uint32_t)v[t];
      if (j == i) continue;
      uint32_t lo = c0, hi = c1;  // first position with colidx >= j
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (colidx[mid] < j) lo = mid + 1; else hi = mid;
      }
      if (lo < c1 && colidx[lo] == j) mask[lo] |= (MT)1 << k;
    }
  }
  if (!rank) return;
  bool over = false;
  for (uint32_t q = c0; q < c1; ++q) {
    const uint32_t j = colidx[q];
    MT m = mask[q];
    uint32_t rk = 0, n = 0;
    while (m) {
      const uint32_t k = sizeof(MT) == 8 ? (uint32_t)__builtin_ctzll((unsigned long long)m)
                                         : (uint32_t)__builtin_ctz((unsigned)m);
      m &= m - (MT)1;
      const uint32_t a = slot[rb + k] & 0x3fu;
      const int32_t *v = ep + (uint64_t)ve_elem[rb + k] * card;
      uint32_t b = 0;
      while (b < card && (uint32_t)v[b] != j) ++b;
      if (n < 16) rk |= ((b > a ? b - 1 : b) & 3u) << (2 * n);
      ++n;
    }
    rank[q] = rk;
    over = over || n > 16;
  }
  if (over)
    for (uint32_t p = rb; p < re; ++p) slot[p] = (uint8_t)(slot[p] | 0x40u);
}

__global__ void k_asm_maxdeg(uint32_t nrows, const uint32_t *__restrict__ ve_ptr,
                             uint32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrows) atomicMax(out, ve_ptr[i + 1] - ve_ptr[i]);
}

// Per wave of 64 blocks: the number of diagonal blocks it holds, or 255 when
// its rows span more than the 64-row window, a row has more pairs than the
// mask has bits, or a pair has no slot or a repeated one; and the number of
// LDS chunks its pairs take.  cnt[0..2]: generic waves for 1x1, 2x2 and 3x3
// blocks.
__global__ void k_asm_waves(uint64_t nwaves, uint64_t nnz, uint32_t nrows, uint32_t bits,
                            uint32_t card, const uint32_t *__restrict__ rowptr,
                            const uint32_t *__restrict__ colidx,
                            const uint32_t *__restrict__ wave_row,
                            const uint32_t *__restrict__ ve_ptr,
                            const uint8_t *__restrict__ slot, uint8_t *__restrict__ wnd,
                            uint8_t *__restrict__ wnc, unsigned long long *__restrict__ cnt) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwaves) return;
  const uint64_t b0 = w * kWave, blast = min(b0 + kWave, nnz) - 1;
  const uint32_t rfirst = wave_row[w];
  uint32_t lo = rfirst, hi = nrows;  // last row with rowptr <= blast
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)rowptr[mid] <= blast) lo = mid; else hi = mid;
  }
  const uint32_t rlast = lo;
  uint32_t nd = 0;
  bool slow = rlast - rfirst + 1 > (uint32_t)kWave;
  for (uint32_t r = rfirst; !slow && r <= rlast; ++r) {
    if (ve_ptr[r + 1] - ve_ptr[r] > bits) slow = true;
    // a pair without its slot (an index that does not match the endpoints),
    // or of an element that holds the row more than once
    for (uint32_t p = ve_ptr[r]; p < ve_ptr[r + 1]; ++p)
      if (slot[p] >= card) slow = true;
    uint32_t a = rowptr[r], b = rowptr[r + 1];
    while (a < b) {
      const uint32_t mid = a + (b - a) / 2;
      if (colidx[mid] < r) a = mid + 1; else b = mid;
    }
    if (a < rowptr[r + 1] && colidx[a] == r && a >= b0 && a <= blast) ++nd;
  }
  wnd[w] = slow ? 255 : (uint8_t)nd;
  // chunks: the fewest n such that, with each row's list cut into n parts
  // ([c*deg/n, (c+1)*deg/n) for part c), every chunk c (part c of each row)
  // holds at most 64 pairs
  uint32_t n = 1;
  if (!slow) {
    const uint32_t T = ve_ptr[rlast + 1] - ve_ptr[rfirst];
    n = T > 0 ? (T + kWave - 1) / kWave : 1;
    for (;; ++n) {
      uint32_t worst = 0;
      for (uint32_t c = 0; c < n; ++c) {
        uint32_t sz = 0;
        for (uint32_t r = rfirst; r <= rlast; ++r) {
          const uint32_t d = ve_ptr[r + 1] - ve_ptr[r];
          sz += (c + 1) * d / n - c * d / n;
        }
        worst = max(worst, sz);
      }
      if (worst <= (uint32_t)kWave || n >= 255) break;
    }
  }
  wnc[w] = (uint8_t)n;
  for (int c = 0; c < 3; ++c)
    if (generic_wave(slow ? 255u : nd, (c + 1) * (c + 1))) atomicAdd(cnt + c, 1ull);
}

template <typename MT>
__device__ __forceinline__ int ctz_mask(MT m) {
  return sizeof(MT) == 8 ? __builtin_ctzll((unsigned long long)m) : __builtin_ctz((unsigned)m);
}

// The fast kernel: the same waves of 64 blocks and the same sums as the
// generic kernel, with the plan in place of searches.
//   A. lane l builds the record of pair c0 + l from its element and slot;
//   D. the diagonal blocks are summed one entry per lane (a lane for each of
//      the BB entries of each diagonal in the wave, two rounds of 64): every
//      pair of the row adds to every entry of the diagonal, so these sums are
//      the long ones (a row's pair count) and need all the lanes;
//   B. an off-diagonal block lane adds only the pairs its mask names (at most
//      a few, the elements holding both vertices).
// Ops with to_diag() run D after B: each pair lane first turns its record
// into its diagonal terms (full lanes), so D is a load and an add per pair.
// Waves with an element that holds a row twice take the generic kernel.
// Every value is summed by one lane in the generic kernel's (the CPU loop's)
// order with the same arithmetic, so the results are bit-identical.
// resident waves per SIMD the fast kernel is compiled for (registers capped
// to fit them): linear tets hold no more than 128 without spilling
template <class Op> struct FastWaves { static constexpr int value = 1; };
template <> struct FastWaves<TetOp> { static constexpr int value = 4; };

template <class Op, int PREC, typename MT>
__global__ void __launch_bounds__(kWave, FastWaves<Op>::value)
k_assemble_fast(typename Op::Args A, uint32_t nrows, uint64_t nnz,
                const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ colidx,
                const uint32_t *__restrict__ wave_row, const uint32_t *__restrict__ ve_ptr,
                const uint32_t *__restrict__ ve_elem, const uint8_t *__restrict__ slot,
                const MT *__restrict__ mask, const uint32_t *__restrict__ rank,
                const uint8_t *__restrict__ wnd, const uint8_t *__restrict__ wnc,
                void *__restrict__ vals,
                void *__restrict__ vals_lo, double *__restrict__ rhs, uint64_t nwaves) {
  constexpr int BD = Op::BD, BB = BD * BD, CARD = Op::CARD;
  constexpr int64_t kBits = 8 * (int64_t)sizeof(MT);
  using Pair = typename Op::Pair;
  constexpr size_t kPairBytes = sizeof(Pair) * kWave;
  constexpr size_t kOutBytes = sizeof(double) * kWave * BB;
  __shared__ __align__(16) unsigned char smem[kPairBytes > kOutBytes ? kPairBytes : kOutBytes];
  Pair *PP = reinterpret_cast<Pair *>(smem);
  __shared__ uint32_t R[kWave + 1];   // rowptr of the wave's rows
  __shared__ uint32_t VP[kWave + 1];  // ve_ptr of the wave's rows
  __shared__ uint8_t DR[kWave];       // local row of the wave's d-th diagonal block
  __shared__ uint8_t DL[kWave];       // and its lane
  __shared__ uint8_t CW[kWave + 1];   // per row: its first slot in the current chunk
  __shared__ uint8_t CS[kWave];       // per row: the rank of that pair in the row's list

  const int lane = threadIdx.x;
  // One workgroup per wave of blocks (a grid of 8 * ceil(nwaves / 8)).
  // Workgroups are dealt to the 8 XCDs round-robin; XCD x takes the x-th
  // eighth of the waves, so neighbouring waves, which gather the same
  // vertices and elements, share that XCD's L2.
  const uint32_t xcd = blockIdx.x & 7u;
  const uint64_t w = nwaves * xcd / 8 + (blockIdx.x >> 3);
  if (w >= nwaves * (xcd + 1) / 8 || generic_wave(wnd[w], BB)) return;
  {
    const uint64_t b0 = w * kWave, bk = b0 + lane;
    const bool valid = bk < nnz;
    const uint32_t rfirst = wave_row[w];
    R[lane] = rowptr[min(rfirst + (uint32_t)lane, nrows)];
    VP[lane] = ve_ptr[min(rfirst + (uint32_t)lane, nrows)];
    if (lane == 0) {
      R[kWave] = rowptr[min(rfirst + (uint32_t)kWave, nrows)];
      VP[kWave] = ve_ptr[min(rfirst + (uint32_t)kWave, nrows)];
    }
    const uint32_t col = valid ? colidx[bk] : 0xffffffffu;
    const MT bm = valid ? mask[bk] : (MT)0;
    uint32_t rk = HasBlockRank<Op>::value && valid ? rank[bk] : 0u;
    __syncthreads();
    // local row: last l with R[l] <= bk (the plan put the wave's rows in the window)
    int lr = 0;
    {
      int hi = kWave + 1;
      while (hi - lr > 1) {
        const int mid = (lr + hi) >> 1;
        if ((uint64_t)R[mid] <= bk) lr = mid; else hi = mid;
      }
    }
    if (!valid) lr = 0;
    const uint32_t row = rfirst + (uint32_t)lr;
    const bool is_diag = valid && col == row;
    int lrl = lr;
    for (int off = 32; off > 0; off >>= 1) lrl = max(lrl, __shfl_xor(lrl, off, 64));
    const unsigned long long dball = __ballot(is_diag);
    const int nd = __popcll(dball);
    if (is_diag) {
      const int rank = __popcll(dball & ((1ull << lane) - 1ull));
      DR[rank] = (uint8_t)lr;
      DL[rank] = (uint8_t)lane;
    }
    __syncthreads();
    // diagonal-entry roles: t = g*64 + lane is entry t % BB of diagonal t / BB
    bool dact[2];
    int dq[2], dl[2];
    uint32_t drow[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int t = g * kWave + lane, d = t / BB;
      dact[g] = d < nd;
      dq[g] = t - d * BB;
      dl[g] = dact[g] ? DR[d] : 0;
      drow[g] = rfirst + (uint32_t)dl[g];
    }

    // The wave's pairs go through LDS in nch chunks of at most 64: chunk c
    // holds part c of every row's list (ranks [c*deg/nch, (c+1)*deg/nch)),
    // rows in order.  A row's pairs stay in list order across the chunks, and
    // every row has about deg/nch of its pairs in each chunk, which bounds the
    // per-row loops below.
    const int nwr = lrl + 1;
    const uint32_t nch = wnc[w];
    const uint32_t rdeg = lane < nwr ? VP[lane + 1] - VP[lane] : 0u;  // lane l: row l
    double acc[BB];
    for (int q = 0; q < BB; ++q) acc[q] = 0.0;
    double a2[2] = {0.0, 0.0}, f2[2] = {0.0, 0.0};
    uint32_t rs = 0;  // lane l < nwr: first rank of row l in this chunk
    for (uint32_t c = 0; c < nch; ++c) {
      const uint32_t re = (c + 1) * rdeg / nch, part = re - rs;
      uint32_t incl = part;  // inclusive prefix over rows -> chunk slots
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      if (lane < nwr) {
        CW[lane] = (uint8_t)(incl - part);
        CS[lane] = (uint8_t)rs;
      }
      if (lane == nwr - 1) CW[nwr] = (uint8_t)incl;
      rs = re;
      __syncthreads();
      const uint32_t csz = CW[nwr];
      if ((uint32_t)lane < csz) {
        int l = 0, hi = nwr;  // the row of slot lane: last l with CW[l] <= lane
        while (hi - l > 1) {
          const int mid = (l + hi) >> 1;
          if (CW[mid] <= (uint32_t)lane) l = mid; else hi = mid;
        }
        const uint32_t p = VP[l] + CS[l] + ((uint32_t)lane - CW[l]);
        typename Op::State st;
        Op::load(A, ve_elem[p], st);
        Op::pair(A, st, slot[p], PP[lane]);
      }
      __syncthreads();
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        if (HasToDiag<Op>::value || !dact[g]) continue;  // (summed after phase B)
        const int q = dq[g];
        const bool rq = Op::kHasRhs && q % (BD + 1) == 0;  // also rhs component q/(BD+1)
        const uint32_t lo = CW[dl[g]], hi = CW[dl[g] + 1];
        for (uint32_t k = lo; k < hi; ++k) {
          const Pair &P = PP[k];
          double v;
          // the slots holding the row (one, unless the element repeats it)
          uint32_t match = 0;
          for (int b = 0; b < CARD; ++b) match |= ((uint32_t)P.vid[b] == drow[g] ? 1u : 0u) << b;
          while (match) {
            const int b = __builtin_ctz(match);
            match &= match - 1;
            if (Op::block_entry(A, P, b, q, v)) a2[g] += v;
          }
          if (Op::kExtra && Op::extra_entry(A, P, q, v)) a2[g] += v;
          if (rq) Op::rhs_entry(A, P, q / (BD + 1), f2[g]);
        }
      }
      if (valid && !is_diag) {
        // the mask bits (row ranks) of this chunk: [k0, k1), at slots CW[lr] + k - k0
        const uint32_t base = CW[lr], k0 = CS[lr], k1 = k0 + (CW[lr + 1] - base);
        MT m = bm;
        if (k1 < (uint32_t)kBits) m &= ((MT)1 << k1) - (MT)1;
        m = k0 >= (uint32_t)kBits ? (MT)0 : (MT)(m & (~(MT)0 << k0));
        while (m) {
          const int k = ctz_mask(m);
          m &= m - (MT)1;
          const Pair &P = PP[base + (uint32_t)k - k0];
          if constexpr (HasBlockRank<Op>::value) {
            // contributions come in mask order, each with the next 2 bits
            double blk[BB];
            const int r = (int)(rk & 3u);
            rk >>= 2;
            if (Op::block_rank(A, P, r, blk))
              for (int q = 0; q < BB; ++q) acc[q] += blk[q];
            continue;
          }
          uint32_t match = 0;
          for (int b = 0; b < CARD; ++b) match |= ((uint32_t)P.vid[b] == col ? 1u : 0u) << b;
          while (match) {
            const int b = __builtin_ctz(match);
            match &= match - 1;
            double blk[BB];
            if (Op::block(A, P, b, blk))
              for (int q = 0; q < BB; ++q) acc[q] += blk[q];
          }
        }
      }
      __syncthreads();
      if constexpr (HasToDiag<Op>::value) {
        // each record becomes its pair's diagonal terms, then the diagonal
        // entries are summed from them in list order
        if ((uint32_t)lane < csz) Op::to_diag(A, PP[lane]);
        __syncthreads();
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          if (!dact[g]) continue;
          const uint32_t lo = CW[dl[g]], hi = CW[dl[g] + 1];
          if constexpr (HasDiagLoad<Op>::value) {
            // four records' loads in flight, then their adds in list order
            const int q = dq[g];
            uint32_t k = lo;
            for (; k + 4 <= hi; k += 4) {
              typename Op::DiagTerm t[4];
#pragma unroll
              for (int j = 0; j < 4; ++j) t[j] = Op::diag_load(PP[k + j], q);
#pragma unroll
              for (int j = 0; j < 4; ++j) Op::diag_add(A, t[j], q, a2[g], f2[g]);
            }
            for (; k < hi; ++k) Op::diag_add(A, Op::diag_load(PP[k], q), q, a2[g], f2[g]);
          } else {
            for (uint32_t k = lo; k < hi; ++k) Op::diag_sum(A, PP[k], dq[g], a2[g], f2[g]);
          }
        }
        __syncthreads();
      }
    }
    double *outbuf = reinterpret_cast<double *>(smem);
    if (valid && !is_diag) {
      Op::finish_block(A, row, col, acc);
      for (int q = 0; q < BB; ++q) outbuf[lane * BB + q] = acc[q];
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      if (!dact[g]) continue;
      const int q = dq[g], d = (g * kWave + lane) / BB;
      double v = a2[g];
      Op::finish_block_entry(A, drow[g], drow[g], q, v);
      Op::finish_diag_entry(A, drow[g], q, v);
      outbuf[DL[d] * BB + q] = v;
      if (Op::kHasRhs && rhs && q % (BD + 1) == 0) {
        double f = f2[g];
        Op::finish_rhs_entry(A, drow[g], q / (BD + 1), f);
        rhs[(uint64_t)drow[g] * BD + q / (BD + 1)] = f;
      }
    }
    __syncthreads();
    const uint64_t nval = (min(b0 + kWave, nnz) - b0) * BB;
    for (int q = 0; q < BB; ++q) {
      const uint64_t i = (uint64_t)q * kWave + lane;
      if (i < nval) store_val<PREC>(vals, vals_lo, b0 * BB + i, outbuf[i]);
    }
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

// The set's plan (slots, masks, wave classes), built on first use and kept
// with the index; one-time work that synchronizes the stream.
int asm_plan(smh_ctx *ctx, const smh_index *idx, int set, smh_index::AsmPlan **out) {
  using namespace smh;
  if (idx->asm_plan.size() < idx->nsets) idx->asm_plan.resize(idx->nsets);
  smh_index::AsmPlan &pl = idx->asm_plan[set];
  *out = &pl;
  if (pl.bits) return SMH_OK;
  const uint32_t nrows = idx->nrows;
  const uint64_t nnz = idx->nnz, ninc = idx->n_inc[set];
  const uint64_t nwaves = (nnz + kWave - 1) / kWave;
  const uint32_t card = idx->card[set];
  unsigned long long *dcnt = nullptr;
  SMH_CHECK(ctx, dmalloc(ctx, (void **)&dcnt, 4 * sizeof(unsigned long long)) == SMH_OK,
            SMH_ERR_HIP, "assembly plan: out of device memory");
  unsigned long long hcnt[4] = {0, 0, 0, 0};
  int rc = SMH_OK;
  auto fail_free = [&](int code) {
    dfree(dcnt);
    dfree(pl.slot);
    dfree(pl.mask);
    dfree(pl.wnd);
    dfree(pl.rank);
    dfree(pl.wnc);
    pl.wnc = nullptr;
    pl.slot = nullptr;
    pl.mask = nullptr;
    pl.wnd = nullptr;
    pl.rank = nullptr;
    return code;
  };
  SMH_HIP(ctx, hipMemsetAsync(dcnt, 0, 4 * sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_asm_maxdeg, dim3(grid_for(nrows, 256)), dim3(256), 0, ctx->stream, nrows,
                     (const uint32_t *)idx->ve_ptr[set], (uint32_t *)(dcnt + 3));
  if (hipMemcpyAsync(hcnt, dcnt, sizeof(hcnt), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail_free(fail(ctx, SMH_ERR_HIP, "assembly plan: device copy failed"));
  const uint32_t maxdeg = (uint32_t)(hcnt[3] & 0xffffffffu);
  const int bits = maxdeg <= 32 ? 32 : 64;
  if ((rc = dmalloc(ctx, (void **)&pl.slot, ninc ? ninc : 1)) != SMH_OK ||
      (rc = dmalloc(ctx, &pl.mask, (bits / 8) * (nnz ? nnz : 1))) != SMH_OK ||
      (rc = dmalloc(ctx, (void **)&pl.wnd, nwaves ? nwaves : 1)) != SMH_OK ||
      (rc = dmalloc(ctx, (void **)&pl.wnc, nwaves ? nwaves : 1)) != SMH_OK ||
      (card <= 4 && (rc = dmalloc(ctx, (void **)&pl.rank, 4 * (nnz ? nnz : 1))) != SMH_OK))
    return fail_free(rc);
  if (nrows) {
    if (bits == 32)
      hipLaunchKernelGGL(k_asm_plan<uint32_t>, dim3(grid_for(nrows, 256)), dim3(256), 0,
                         ctx->stream, nrows, card, idx->ep[set], (const uint32_t *)idx->rowptr,
                         (const uint32_t *)idx->colidx, (const uint32_t *)idx->ve_ptr[set],
                         (const uint32_t *)idx->ve_elem[set], pl.slot, (uint32_t *)pl.mask,
                         pl.rank);
    else
      hipLaunchKernelGGL(k_asm_plan<uint64_t>, dim3(grid_for(nrows, 256)), dim3(256), 0,
                         ctx->stream, nrows, card, idx->ep[set], (const uint32_t *)idx->rowptr,
                         (const uint32_t *)idx->colidx, (const uint32_t *)idx->ve_ptr[set],
                         (const uint32_t *)idx->ve_elem[set], pl.slot, (uint64_t *)pl.mask,
                         pl.rank);
  }
  if (nwaves)
    hipLaunchKernelGGL(k_asm_waves, dim3(grid_for(nwaves, 256)), dim3(256), 0, ctx->stream,
                       nwaves, nnz, nrows, (uint32_t)bits, card, (const uint32_t *)idx->rowptr,
                       (const uint32_t *)idx->colidx, (const uint32_t *)idx->wave_row,
                       (const uint32_t *)idx->ve_ptr[set], (const uint8_t *)pl.slot, pl.wnd,
                       pl.wnc, dcnt);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(hcnt, dcnt, sizeof(hcnt), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail_free(fail(ctx, SMH_ERR_HIP, "assembly plan: kernels failed"));
  dfree(dcnt);
  for (int c = 0; c < 3; ++c) pl.nslow[c] = hcnt[c];
  pl.bits = bits;
  return SMH_OK;
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
    // SMH_ASSEMBLY_GENERIC=1: every wave through the generic kernel (tests)
    // (a plan not built before a stream capture is not built inside it)
    const char *gen = getenv("SMH_ASSEMBLY_GENERIC");
    smh_index::AsmPlan *pl = nullptr;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess) {
      (void)hipGetLastError();
      cap = hipStreamCaptureStatusNone;
    }
    const bool built = (uint32_t)set < idx->asm_plan.size() && idx->asm_plan[set].bits;
    if (!(gen && atoi(gen) == 1) && (built || cap == hipStreamCaptureStatusNone)) {
      const int rc = asm_plan(ctx, idx, set, &pl);
      if (rc != SMH_OK) return rc;
    }
    constexpr int BB = Op::BD * Op::BD;
    const uint64_t nslow = pl ? pl->nslow[Op::BD - 1] : nwaves;
    // grid: a multiple of 8 (one share per XCD), no more than fills the chip
    auto grid = [&](unsigned per_cu, uint64_t n) {
      const uint64_t full = (uint64_t)ctx->num_cu * per_cu;
      const uint64_t g = std::min<uint64_t>(full, n);
      return (unsigned)((g + 7) / 8 * 8);
    };
    static_assert(BB == 1 || BB == 4 || BB == 9, "1x1, 2x2 or 3x3 blocks");
    if (pl && nslow < nwaves) {
#define SMH_FAST(P, MT)                                                                      \
  hipLaunchKernelGGL((k_assemble_fast<Op, P, MT>), dim3((unsigned)((nwaves + 7) / 8 * 8)),   \
                     dim3(kWave), 0, ctx->stream, A, nrows, idx->nnz, idx->rowptr,           \
                     idx->colidx, idx->wave_row, idx->ve_ptr[set], idx->ve_elem[set],        \
                     pl->slot, (const MT *)pl->mask, pl->rank, pl->wnd, pl->wnc, vals,        \
                     vals_lo, rhs, nwaves)
#define SMH_FAST_P(MT)                            \
  do {                                            \
    if (prec == SMH_F64) SMH_FAST(SMH_F64, MT);   \
    else if (prec == SMH_F32) SMH_FAST(SMH_F32, MT); \
    else SMH_FAST(SMH_F32X2, MT);                 \
  } while (0)
      if (pl->bits == 32) SMH_FAST_P(uint32_t);
      else SMH_FAST_P(uint64_t);
#undef SMH_FAST_P
#undef SMH_FAST
    }
    if (nslow) {
      const uint8_t *wnd = pl ? pl->wnd : nullptr;
#define SMH_ASM(P)                                                                          \
  hipLaunchKernelGGL((k_assemble<Op, P>), dim3(grid(assembly_wgs_per_cu<Op, P>(), nslow)),  \
                     dim3(kWave), 0, ctx->stream, A, nrows, idx->nnz, idx->rowptr,          \
                     idx->colidx, idx->wave_row, idx->ve_ptr[set], idx->ve_elem[set], vals, \
                     vals_lo, rhs, nwaves, wnd)
      if (prec == SMH_F64) SMH_ASM(SMH_F64);
      else if (prec == SMH_F32) SMH_ASM(SMH_F32);
      else SMH_ASM(SMH_F32X2);
#undef SMH_ASM
    }
  }
  hipLaunchKernelGGL(k_isolated<Op>, dim3(grid_for(nrows, 256)), dim3(256), 0, ctx->stream, A,
                     nrows, idx->rowptr, rhs);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}

}  // namespace

namespace smh {
int prepare_assembly(smh_ctx *ctx, const smh_index *idx, const int32_t *ep, uint32_t card) {
  for (uint32_t s = 0; s < idx->nsets; ++s)
    if (idx->ep[s] == ep && idx->card[s] == card) {
      smh_index::AsmPlan *pl = nullptr;
      return idx->nnz ? asm_plan(ctx, idx, (int)s, &pl) : SMH_OK;
    }
  return SMH_OK;
}
}  // namespace smh

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
