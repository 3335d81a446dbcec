// Program tests through the C++ simit API on the HIP backend, written the
// way the reference's gtest program tests are (test/program-isprings.cpp,
// test/program-cg-test.cpp): build sets, load the program, bind, runSafe,
// compare against the reference's asserted values.
//
// Usage: simit_program_tests [name ...]; exits non-zero on any failure.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "simit/graph.h"
#include "simit/mesh.h"
#include "simit/program.h"
#include "simit/reorder.h"
#include "simit_hip.h"

using namespace simit;

static int failures = 0;

#define EXPECT_NEAR(want, got, tol)                                                        \
  do {                                                                                     \
    const double w_ = (want), g_ = (got);                                                  \
    if (!(std::fabs(w_ - g_) <= (tol))) {                                                  \
      std::printf("  FAIL %s:%d: %s = %.17g, expected %.17g\n", __FILE__, __LINE__, #got, \
                  g_, w_);                                                                 \
      ++failures;                                                                          \
    }                                                                                      \
  } while (0)
#define EXPECT_TRUE(c)                                                        \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);              \
      ++failures;                                                             \
    }                                                                         \
  } while (0)

// The programs run either as this backend's fixed pipelines (by name) or,
// when g_programs is set, parsed from Simit source in tests/programs/ (this
// backend's own restatements of the reference programs, same arithmetic).
static bool g_programs = false;
static std::string programFile(const std::string &program) {
  if (program == "isprings") return "tests/programs/springs.sim";
  if (program == "cg") return "tests/programs/cg_points.sim";
  if (program == "femTet") return "tests/programs/neohookean.sim";
  return "";
}
static Function loadFn(const std::string &program, const std::string &function) {
  Program p;
  const std::string file = g_programs ? programFile(program) : "";
  const int rc = file.empty() ? p.loadBuiltin(program) : p.loadFile(file);
  if (rc != 0) {
    std::printf("  FAIL: cannot load %s: %s\n", (file.empty() ? program : file).c_str(),
                p.getDiagnostics().c_str());
    ++failures;
    return Function();
  }
  return p.compile(function);
}
static Function load(const std::string &program) { return loadFn(program, "main"); }

// test/program-isprings.cpp:91-175
static void isprings() {
  const double l0_ = 1.0, stiffness = 100.0, density = 1000.0, radius = 0.01;
  const double PI_ = 3.14159265358979, zfloor = 0.4;
  Set points;
  FieldRef<double, 3> x = points.addField<double, 3>("x");
  FieldRef<double, 3> v = points.addField<double, 3>("v");
  FieldRef<double> m = points.addField<double>("m");
  FieldRef<bool> fixed = points.addField<bool>("fixed");
  Set springs(points, points);
  FieldRef<double> k = springs.addField<double>("k");
  FieldRef<double> l0 = springs.addField<double>("l0");
  ElementRef p0 = points.add(), p1 = points.add(), p2 = points.add();
  x(p0) = {0.0, 0.0, 0.0};
  x(p1) = {0.0, 0.0, 1.0};
  x(p2) = {0.0, 0.0, 2.0};
  for (ElementRef p : points) {
    v(p) = {0.0, 0.0, 0.0};
    m(p) = PI_ * radius * radius * l0_ * density;
    fixed(p) = x(p)(2) < zfloor;
  }
  ElementRef s0 = springs.add(p0, p1), s1 = springs.add(p1, p2);
  l0(s0) = l0_;
  l0(s1) = l0_;
  k(s0) = stiffness;
  k(s1) = stiffness;
  Function func = load("isprings");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  for (int i = 0; i < 10; ++i) func.runSafe();
  const double tol = 1e-13;
  EXPECT_NEAR(0.0, v(p1)(0), tol);
  EXPECT_NEAR(-0.54482994491692449, v(p1)(2), tol);
  EXPECT_NEAR(-0.8480293605755741, v(p2)(2), tol);
  EXPECT_NEAR(0.0, x(p0)(2), tol);
  EXPECT_NEAR(0.96180188527075183, x(p1)(2), tol);
  EXPECT_NEAR(1.9500008677096612, x(p2)(2), tol);
}

// test/program-cg-test.cpp:10-44
static void cg() {
  Set points;
  FieldRef<double> b = points.addField<double>("b");
  FieldRef<double> c = points.addField<double>("c");
  FieldRef<int> id = points.addField<int>("id");
  ElementRef p0 = points.add(), p1 = points.add(), p2 = points.add();
  b.set(p0, 1.0);
  b.set(p1, 2.0);
  b.set(p2, 3.0);
  (void)id;
  Set springs(points, points);
  FieldRef<double> a = springs.addField<double>("a");
  ElementRef s0 = springs.add(p0, p1), s1 = springs.add(p1, p2);
  a.set(s0, 4.0);
  a.set(s1, 5.0);
  Function func = load("cg");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  EXPECT_NEAR(0.95883777239709455653, c.get(p0), 1e-13);
  EXPECT_NEAR(1.98789346246973352983, c.get(p1), 1e-13);
  EXPECT_NEAR(3.05326876513317202466, c.get(p2), 1e-13);
}

// the 20^3 lattice step: converges to a 1e-8 relative FP64 residual
static void springs_lattice() {
  Set points;
  FieldRef<double, 3> x = points.addField<double, 3>("x");
  FieldRef<double, 3> v = points.addField<double, 3>("v");
  FieldRef<double> m = points.addField<double>("m");
  FieldRef<bool> fixed = points.addField<bool>("fixed");
  Set springs(points, points);
  FieldRef<double> k = springs.addField<double>("k");
  FieldRef<double> l0 = springs.addField<double>("l0");
  Box box = createBox(&points, &springs, 20, 20, 20);
  EXPECT_TRUE(points.getSize() == 8000 && springs.getSize() == 22800);
  unsigned seed = 12345;
  auto jit = [&]() {
    seed = seed * 1103515245u + 12345u;
    return ((seed >> 8) & 0xffff) / 65535.0 * 0.2 - 0.1;
  };
  for (unsigned i = 0; i < 20; ++i)
    for (unsigned j = 0; j < 20; ++j)
      for (unsigned l = 0; l < 20; ++l) {
        ElementRef p = box(i, j, l);
        x(p) = {i + jit(), j + jit(), l + jit()};
        v(p) = {0.0, 0.0, 0.0};
        m(p) = 1.0;
        fixed(p) = l == 0;
      }
  for (ElementRef s : springs) {
    k(s) = 1e3;
    l0(s) = 1.0;
  }
  Function func = load("springs_implicit");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  SolveStats st = func.lastSolve();
  std::printf("  springs_implicit: %d iterations, relative residual %.3e\n", st.iters, st.true_rel);
  EXPECT_TRUE(st.status == 0);
  EXPECT_TRUE(st.true_rel <= 1e-8);
  EXPECT_NEAR(0.0, v(box(3, 4, 0))(2), 0.0);  // fixed layer stays put
  EXPECT_TRUE(v(box(10, 10, 10))(2) < 0.0);   // the rest falls
}

// The reference's tetgen bunny (apps/data/tet-bunny, a test fixture here)
// loaded with MeshVol, its edges as springs at rest length, one implicit
// step: an irregular mesh through the whole C++ pipeline.
static void mesh_bunny() {
  MeshVol mesh;
  const std::string dir = "tests/golden/tet-bunny/";
  EXPECT_TRUE(mesh.loadTet(dir + "bunny.1.node", dir + "bunny.1.ele") == 0);
  EXPECT_TRUE(mesh.loadTetEdge(dir + "bunny.1.edge") == 0);
  EXPECT_TRUE(mesh.v.size() == 3607 && mesh.e.size() == 12687 && mesh.edges.size() == 19233);
  if (mesh.v.size() != 3607 || mesh.edges.size() != 19233) return;
  EXPECT_NEAR(-0.89565700000000004, mesh.v[0][0], 0.0);
  EXPECT_TRUE(mesh.e[0].size() == 4 && mesh.e[0][0] == 2342 && mesh.e[0][3] == 3169);
  EXPECT_TRUE(mesh.edges[0][0] == 3169 && mesh.edges[0][1] == 2323);
  EXPECT_TRUE(MeshVol().loadTet(dir + "missing.node", dir + "bunny.1.ele") == -1);
  Set points;
  FieldRef<double, 3> x = points.addField<double, 3>("x");
  FieldRef<double, 3> v = points.addField<double, 3>("v");
  FieldRef<double> m = points.addField<double>("m");
  FieldRef<bool> fixed = points.addField<bool>("fixed");
  Set springs(points, points);
  FieldRef<double> k = springs.addField<double>("k");
  FieldRef<double> l0 = springs.addField<double>("l0");
  double zmin = 1e300, zmax = -1e300;
  for (const auto &p : mesh.v) {
    zmin = std::fmin(zmin, p[2]);
    zmax = std::fmax(zmax, p[2]);
  }
  std::vector<ElementRef> pe;
  for (const auto &p : mesh.v) {
    ElementRef q = points.add();
    x(q) = {p[0], p[1], p[2]};
    v(q) = {0.0, 0.0, 0.0};
    m(q) = 1.0;
    fixed(q) = p[2] <= zmin + 0.05 * (zmax - zmin);
    pe.push_back(q);
  }
  for (const auto &ed : mesh.edges) {
    ElementRef s = springs.add(pe[ed[0]], pe[ed[1]]);
    const auto &a = mesh.v[ed[0]], &b = mesh.v[ed[1]];
    k(s) = 1e3;
    l0(s) = std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                      (a[2] - b[2]) * (a[2] - b[2]));
  }
  Function func = load("springs_implicit");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  SolveStats st = func.lastSolve();
  std::printf("  bunny springs_implicit: %d iterations, relative residual %.3e\n", st.iters,
              st.true_rel);
  EXPECT_TRUE(st.status == 0);
  EXPECT_TRUE(st.true_rel <= 1e-8);
  EXPECT_TRUE(v(pe[0])(2) < 0.0 || fixed(pe[0]));  // free points fall
}

// ---- reordering (host only) ---------------------------------------------------------------
// test/reorder-tests.cpp:177-300 (reorderInt, reorderDouble): three vertices
// x = (i,i,i), a = (3+i) x 4; edges (1,2) with 20s and (0,1) with 10s;
// vertices moved by {2,0,1}, edges by {1,0}
template <typename T>
static void reorder_values() {
  Set verts;
  Set edges(verts, verts);
  FieldRef<T, 3> x = verts.addField<T, 3>("x");
  FieldRef<T, 4> a = verts.addField<T, 4>("a");
  FieldRef<T, 5> tx = edges.addField<T, 5>("x");
  std::vector<ElementRef> vr;
  for (int i = 0; i < 3; ++i) {
    vr.push_back(verts.add());
    x(vr.back()) = {(T)i, (T)i, (T)i};
    a(vr.back()) = {(T)(3 + i), (T)(3 + i), (T)(3 + i), (T)(3 + i)};
  }
  ElementRef e0 = edges.add(vr[1], vr[2]);
  tx(e0) = {(T)20, (T)20, (T)20, (T)20, (T)20};
  ElementRef e1 = edges.add(vr[0], vr[1]);
  tx(e1) = {(T)10, (T)10, (T)10, (T)10, (T)10};
  std::vector<int> vo{2, 0, 1}, eo{1, 0};
  reorderVertexSet(edges, verts, vo);
  reorderEdgeSet(edges, eo);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      EXPECT_NEAR((double)i, (double)x(ElementRef(vo[i]))(j), 0.0);
      EXPECT_TRUE(x(ElementRef(i))(j) != (T)i);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) EXPECT_NEAR(3.0 + i, (double)a(ElementRef(vo[i]))(j), 0.0);
  const double want[2] = {20, 10};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 5; ++j) EXPECT_NEAR(want[i], (double)tx(ElementRef(eo[i]))(j), 0.0);
  // the endpoints followed: old edge 0 = (1,2) -> (vo[1], vo[2]) at position eo[0]
  EXPECT_TRUE(edges.getEndpoint(ElementRef(eo[0]), 0).getIdent() == vo[1]);
  EXPECT_TRUE(edges.getEndpoint(ElementRef(eo[0]), 1).getIdent() == vo[2]);
  EXPECT_TRUE(edges.getEndpoint(ElementRef(eo[1]), 0).getIdent() == vo[0]);
  EXPECT_TRUE(edges.getEndpoint(ElementRef(eo[1]), 1).getIdent() == vo[1]);
  bool threw = false;
  try {
    std::vector<int> bad{0, 0, 1};
    reorderVertexSet(edges, verts, bad);
  } catch (const SimitException &) {
    threw = true;
  }
  EXPECT_TRUE(threw);
}
static void reorder_int() { reorder_values<int>(); }
static void reorder_double() { reorder_values<double>(); }

// the Hilbert curve of order 3: a bijection onto [0, 512) whose consecutive
// indices are lattice neighbours
static void hilbert_curve() {
  const int bits = 3, n = 1 << bits;
  std::vector<int> at(n * n * n, -1);
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < n; ++z) {
        const uint64_t h = hilbertIndex3(x, y, z, bits);
        EXPECT_TRUE(h < (uint64_t)(n * n * n));
        if (h >= (uint64_t)(n * n * n)) return;
        EXPECT_TRUE(at[h] < 0);
        at[h] = (x * n + y) * n + z;
      }
  EXPECT_TRUE(hilbertIndex3(0, 0, 0, bits) == 0);
  for (int h = 1; h < n * n * n; ++h) {
    const int p = at[h - 1], q = at[h];
    const int d = std::abs(p / (n * n) - q / (n * n)) + std::abs(p / n % n - q / n % n) +
                  std::abs(p % n - q % n);
    EXPECT_TRUE(d == 1);
    if (d != 1) return;
  }
}

struct BunnySets {
  MeshVol mesh;
  Set points;
  Set springs{points, points};
  std::vector<ElementRef> pe;
};
static bool bunny_sets(BunnySets &b) {
  const std::string dir = "tests/golden/tet-bunny/";
  if (b.mesh.loadTet(dir + "bunny.1.node", dir + "bunny.1.ele") != 0 ||
      b.mesh.loadTetEdge(dir + "bunny.1.edge") != 0)
    return false;
  FieldRef<double, 3> x = b.points.addField<double, 3>("x");
  FieldRef<double, 3> v = b.points.addField<double, 3>("v");
  FieldRef<double> m = b.points.addField<double>("m");
  FieldRef<bool> fixed = b.points.addField<bool>("fixed");
  FieldRef<double> k = b.springs.addField<double>("k");
  FieldRef<double> l0 = b.springs.addField<double>("l0");
  double zmin = 1e300, zmax = -1e300;
  for (const auto &p : b.mesh.v) {
    zmin = std::fmin(zmin, p[2]);
    zmax = std::fmax(zmax, p[2]);
  }
  for (const auto &p : b.mesh.v) {
    ElementRef q = b.points.add();
    x(q) = {p[0], p[1], p[2]};
    v(q) = {0.0, 0.0, 0.0};
    m(q) = 1.0;
    fixed(q) = p[2] <= zmin + 0.05 * (zmax - zmin);
    b.pe.push_back(q);
  }
  for (const auto &ed : b.mesh.edges) {
    ElementRef s = b.springs.add(b.pe[ed[0]], b.pe[ed[1]]);
    const auto &p = b.mesh.v[ed[0]], &q = b.mesh.v[ed[1]];
    k(s) = 1e3;
    l0(s) = std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                      (p[2] - q[2]) * (p[2] - q[2]));
  }
  b.points.setSpatialField("x");
  return true;
}
static double mean_span(const Set &edges) {
  double s = 0;
  const int *ep = edges.getEndpointsData();
  for (int e = 0; e < edges.getSize(); ++e) s += std::abs(ep[2 * e] - ep[2 * e + 1]);
  return s / edges.getSize();
}

// Hilbert reordering of the bunny: valid permutations, every field and
// endpoint moved with its element, and shorter index spans along the edges
static void reorder_bunny_host() {
  BunnySets b;
  EXPECT_TRUE(bunny_sets(b));
  if (b.points.getSize() != 3607) return;
  const double span0 = mean_span(b.springs);
  std::vector<double> x0, l00;
  FieldRef<double, 3> x = b.points.getField<double, 3>("x");
  FieldRef<double> l0 = b.springs.getField<double>("l0");
  for (ElementRef p : b.points)
    for (int d = 0; d < 3; ++d) x0.push_back(x(p)(d));
  for (ElementRef s : b.springs) l00.push_back(l0(s));
  std::vector<int> eo, vo;
  reorder(b.springs, b.points, eo, vo);
  const double span1 = mean_span(b.springs);
  std::printf("  mean |i - j| over the edges: %.1f before, %.1f after\n", span0, span1);
  EXPECT_TRUE(span1 < 0.5 * span0);
  for (int i = 0; i < b.points.getSize(); ++i)
    for (int d = 0; d < 3; ++d) EXPECT_NEAR(x0[3 * i + d], x(ElementRef(vo[i]))(d), 0.0);
  for (int e = 0; e < b.springs.getSize(); ++e) {
    EXPECT_NEAR(l00[e], l0(ElementRef(eo[e])), 0.0);
    const int a = b.springs.getEndpoint(ElementRef(eo[e]), 0).getIdent();
    const int c = b.springs.getEndpoint(ElementRef(eo[e]), 1).getIdent();
    EXPECT_TRUE(a == vo[b.mesh.edges[e][0]] && c == vo[b.mesh.edges[e][1]]);
  }
}

// host-only part of the loader checks
static void mesh_load() {
  MeshVol mesh;
  const std::string dir = "tests/golden/tet-bunny/";
  EXPECT_TRUE(mesh.loadTet(dir + "bunny.1.node", dir + "bunny.1.ele") == 0);
  EXPECT_TRUE(mesh.loadTetEdge(dir + "bunny.1.edge") == 0);
  EXPECT_TRUE(mesh.v.size() == 3607 && mesh.e.size() == 12687 && mesh.edges.size() == 19233);
  EXPECT_TRUE(MeshVol().loadTet(dir + "missing.node", dir + "bunny.1.ele") == -1);
}

// the springs step on the bunny, unreordered and reordered: the same motion
static void reorder_bunny_springs() {
  BunnySets a, b;
  EXPECT_TRUE(bunny_sets(a) && bunny_sets(b));
  std::vector<int> eo, vo;
  reorder(b.springs, b.points, eo, vo);
  Function fa = load("springs_implicit"), fb = load("springs_implicit");
  if (!fa.defined() || !fb.defined()) return;
  fa.bind("points", &a.points);
  fa.bind("springs", &a.springs);
  fb.bind("points", &b.points);
  fb.bind("springs", &b.springs);
  fa.runSafe();
  fb.runSafe();
  std::printf("  iterations: %d unreordered, %d reordered\n", fa.lastSolve().iters,
              fb.lastSolve().iters);
  FieldRef<double, 3> va = a.points.getField<double, 3>("v"), vb = b.points.getField<double, 3>("v");
  double scale = 0, diff = 0;
  for (int i = 0; i < a.points.getSize(); ++i)
    for (int d = 0; d < 3; ++d) {
      scale = std::fmax(scale, std::fabs(va(ElementRef(i))(d)));
      diff = std::fmax(diff, std::fabs(va(ElementRef(i))(d) - vb(ElementRef(vo[i]))(d)));
    }
  std::printf("  max |v - v_reordered| = %.3e (max |v| %.3e)\n", diff, scale);
  EXPECT_TRUE(scale > 0 && diff <= 1e-6 * scale);
}

// test/function-tests.cpp:96-139: an extern int matrix bound as CSR
// TensorData and negated by the function; then the same CSR with double
// values through y = A x, and a small SPD CSR through x = cg(A, b)
static void extern_csr() {
  Set V;
  V.add();
  V.add();
  V.add();
  // 1.0 2.0 0.0
  // 0.0 3.0 4.0
  // 0.0 0.0 0.0
  int A_rowPtr[4] = {0, 2, 4, 4};
  int A_colInd[4] = {0, 1, 1, 2};
  int A_vals[4] = {1, 2, 3, 4};
  {
    Function f = load("extern_neg");
    if (!f.defined()) return;
    f.bind("V", &V);
    TensorData data(A_rowPtr, A_colInd, A_vals, 4, 4);
    f.bind("A", data);
    f.runSafe();
    EXPECT_TRUE(A_vals[0] == -1 && A_vals[1] == -2 && A_vals[2] == -3 && A_vals[3] == -4);
    std::ostringstream os;
    f.print(os);
    f.printMachine(os);
    std::printf("%s", os.str().c_str());
    EXPECT_TRUE(os.str().find("extern A") != std::string::npos);
    EXPECT_TRUE(os.str().find("gfx950") != std::string::npos);
  }
  {
    double vals[4] = {1.0, 2.0, 3.0, 4.0}, x[3] = {1.0, 2.0, 3.0}, y[3] = {-7, -7, -7};
    Function f = load("extern_spmv");
    if (!f.defined()) return;
    f.bind("V", &V);
    TensorData data(A_rowPtr, A_colInd, vals, 4, 4);
    f.bind("A", data);
    f.bind("x", x);
    f.bind("y", y);
    f.runSafe();
    EXPECT_NEAR(5.0, y[0], 0.0);
    EXPECT_NEAR(18.0, y[1], 0.0);
    EXPECT_NEAR(0.0, y[2], 0.0);
    // a rowLen that does not match the bound set is rejected
    int badRow[3] = {0, 2, 4};
    TensorData bad(badRow, A_colInd, vals, 3, 4);
    f.bind("A", bad);
    bool threw = false;
    try {
      f.runSafe();
    } catch (const SimitException &) {
      threw = true;
    }
    EXPECT_TRUE(threw);
  }
  {
    // [4 1 0; 1 3 1; 0 1 2] x = [1 2 3]
    int rp[4] = {0, 2, 5, 7}, ci[7] = {0, 1, 0, 1, 2, 1, 2};
    double vals[7] = {4, 1, 1, 3, 1, 1, 2}, b[3] = {1, 2, 3}, x[3] = {0, 0, 0};
    Function f = load("extern_cg");
    if (!f.defined()) return;
    f.bind("V", &V);
    TensorData data(rp, ci, vals, 4, 7);
    f.bind("A", data);
    f.bind("b", b);
    f.bind("x", x);
    f.runSafe();
    // exact solution (2/9, 1/9, 13/9)
    EXPECT_NEAR(2.0 / 9.0, x[0], 1e-7);
    EXPECT_NEAR(1.0 / 9.0, x[1], 1e-7);
    EXPECT_NEAR(13.0 / 9.0, x[2], 1e-7);
  }
}

// test/program-fem-test.cpp:14-106: the bar2k mesh, initializeTet, ten
// steps of femTet.sim's main; and test/reorder-tests.cpp:403-441, the same
// after a Hilbert reordering of the vertices and tets
static FieldRef<double, 3> initializeFem(const MeshVol &mv, Set &verts, Set &tets,
                                         std::vector<ElementRef> &vertRefs) {
  FieldRef<double, 3> x = verts.addField<double, 3>("x");
  FieldRef<double, 3> v = verts.addField<double, 3>("v");
  FieldRef<double, 3> fe = verts.addField<double, 3>("fe");
  FieldRef<int> c = verts.addField<int>("c");
  FieldRef<double> m = verts.addField<double>("m");
  FieldRef<double> u = tets.addField<double>("u");
  FieldRef<double> l = tets.addField<double>("l");
  tets.addField<double>("W");
  tets.addField<double, 3, 3>("B");
  const double E = 5e3, nu = 0.45;
  const double uval = 0.5 * E / nu, lval = E * nu / ((1 + nu) * (1 - 2 * nu));
  const double eps = 0.0001;
  for (size_t ii = 0; ii < mv.v.size(); ++ii) {
    vertRefs.push_back(verts.add());
    ElementRef p = vertRefs.back();
    x.set(p, {mv.v[ii][0], mv.v[ii][1], mv.v[ii][2]});
    v.set(p, {0.1, 0.0, 0.1});
    c.set(p, 0);
    fe.set(p, {0.0, 0.0, 0.0});
    m.set(p, 0.0);
    if (mv.v[ii][1] < eps) {
      c.set(p, 1);
      v.set(p, {0.0, 0.0, 0.0});
    }
  }
  for (size_t ii = 0; ii < mv.e.size(); ++ii) {
    ElementRef t = tets.add(vertRefs[mv.e[ii][0]], vertRefs[mv.e[ii][1]],
                            vertRefs[mv.e[ii][2]], vertRefs[mv.e[ii][3]]);
    u.set(t, uval);
    l.set(t, lval);
  }
  return x;
}
static void runFem(Set &verts, Set &tets, int nSteps) {
  Function pre = loadFn("femTet", "initializeTet");
  if (!pre.defined()) return;
  pre.bind("verts", &verts);
  pre.bind("tets", &tets);
  pre.init();
  pre.runSafe();
  Function step = loadFn("femTet", "main");
  if (!step.defined()) return;
  step.bind("verts", &verts);
  step.bind("tets", &tets);
  step.init();
  for (int i = 0; i < nSteps; ++i) {
    step.runSafe();
    std::printf("  step %d: cg iterations %d\n", i, step.lastSolve().iters);
  }
}
static void checkFem(FieldRef<double, 3> &x, ElementRef one, ElementRef two, ElementRef three) {
  EXPECT_NEAR(0.010771915616785779, x(one)(0), 1e-13);
  EXPECT_NEAR(0.058853573999788439, x(one)(1), 1e-13);
  EXPECT_NEAR(0.030899457015375883, x(one)(2), 1e-13);
  EXPECT_NEAR(0.0028221631202928516, x(two)(0), 1e-13);
  EXPECT_NEAR(0.017969982607667911, x(two)(1), 1e-13);
  EXPECT_NEAR(0.012885386063393013, x(two)(2), 1e-13);
  EXPECT_NEAR(0.02411959295647129, x(three)(0), 1e-13);
  EXPECT_NEAR(0.052036155669135678, x(three)(1), 1e-13);
  EXPECT_NEAR(0.030173075240629205, x(three)(2), 1e-13);
}
static void fem_tet() {
  MeshVol mv;
  EXPECT_TRUE(mv.loadTet("tests/golden/fem-bar2k/bar2k.node",
                         "tests/golden/fem-bar2k/bar2k.ele") == 0);
  Set verts;
  Set tets(verts, verts, verts, verts);
  std::vector<ElementRef> vertRefs;
  FieldRef<double, 3> x = initializeFem(mv, verts, tets, vertRefs);
  runFem(verts, tets, 10);
  checkFem(x, vertRefs[100], vertRefs[200], vertRefs[300]);
}
static void reorder_fem_tet() {
  MeshVol mv;
  EXPECT_TRUE(mv.loadTet("tests/golden/fem-bar2k/bar2k.node",
                         "tests/golden/fem-bar2k/bar2k.ele") == 0);
  Set verts;
  Set tets(verts, verts, verts, verts);
  std::vector<ElementRef> vertRefs;
  FieldRef<double, 3> x = initializeFem(mv, verts, tets, vertRefs);
  std::vector<int> vertexOrdering, edgeOrdering;
  verts.setSpatialField("x");
  reorder(tets, verts, edgeOrdering, vertexOrdering);
  int moved = 0;
  for (size_t i = 0; i < vertexOrdering.size(); ++i) moved += vertexOrdering[i] != (int)i;
  std::printf("  %d of %zu vertices moved\n", moved, vertexOrdering.size());
  EXPECT_TRUE(moved > 0);
  runFem(verts, tets, 10);
  checkFem(x, vertRefs[vertexOrdering[100]], vertRefs[vertexOrdering[200]],
           vertRefs[vertexOrdering[300]]);
}

// extern_file DIR: the CSR of DIR/{rowptr,colidx}.i32, DIR/vals.f64 (3x3
// blocks) and DIR/b.f64, bound as TensorData: y = A b written to DIR/y.f64
// and x = cg(A, b) to DIR/x.f64 (the GPU test checks both with the oracle)
static std::vector<char> slurp(const std::string &path) {
  std::FILE *fp = std::fopen(path.c_str(), "rb");
  std::vector<char> out;
  if (!fp) return out;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), fp)) > 0) out.insert(out.end(), buf, buf + n);
  std::fclose(fp);
  return out;
}
static void spit(const std::string &path, const void *p, size_t bytes) {
  std::FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) return;
  std::fwrite(p, 1, bytes, fp);
  std::fclose(fp);
}
static int extern_file(const std::string &dir) {
  std::vector<char> rpb = slurp(dir + "/rowptr.i32"), cib = slurp(dir + "/colidx.i32"),
                    vb = slurp(dir + "/vals.f64"), bb = slurp(dir + "/b.f64");
  if (rpb.size() < 8 || vb.size() != cib.size() / 4 * 72 || bb.empty()) {
    std::printf("  FAIL: bad input files in %s\n", dir.c_str());
    return 1;
  }
  const int nrows = (int)(rpb.size() / 4) - 1, nnz = (int)(cib.size() / 4);
  Set V;
  for (int i = 0; i < nrows; ++i) V.add();
  const int *rp = (const int *)rpb.data(), *ci = (const int *)cib.data();
  double *vals = (double *)vb.data(), *b = (double *)bb.data();
  std::vector<double> y(3 * (size_t)nrows, 0.0), x(3 * (size_t)nrows, 0.0);
  TensorData data(rp, ci, vals, nrows + 1, nnz);
  Function mv = load("extern_spmv_3x3");
  mv.bind("V", &V);
  mv.bind("A", data);
  mv.bind("x", b);
  mv.bind("y", y.data());
  mv.runSafe();
  spit(dir + "/y.f64", y.data(), 8 * y.size());
  Function cg = load("extern_cg_3x3");
  cg.bind("V", &V);
  cg.bind("A", data);
  cg.bind("b", b);
  cg.bind("x", x.data());
  cg.runSafe();
  spit(dir + "/x.f64", x.data(), 8 * x.size());
  SolveStats st = cg.lastSolve();
  std::printf("extern_file: %d rows, %d blocks; cg iters %d status %d true_rel %.3e\n", nrows, nnz,
              st.iters, st.status, st.true_rel);
  return failures ? 1 : 0;
}

// test/system-gemv-tests.cpp, run from this backend's own programs in
// tests/programs/: c = A b on an edge set (:12-58), on a periodic 1-D grid
// edge set (:60-110) and on a periodic 3 x 2 grid (:171-251); the inputs are
// kept and the outputs are exact.
static Function loadProgram(const std::string &file) {
  Program p;
  if (p.loadFile(file) != 0) {
    std::printf("  FAIL: cannot load %s: %s\n", file.c_str(), p.getDiagnostics().c_str());
    ++failures;
    return Function();
  }
  return p.compile("main");
}
static void sim_gemv() {
  Set points;
  FieldRef<double> b = points.addField<double>("b");
  FieldRef<double> c = points.addField<double>("c");
  ElementRef p0 = points.add(), p1 = points.add(), p2 = points.add();
  b.set(p0, 1.0);
  b.set(p1, 2.0);
  b.set(p2, 3.0);
  c.set(p0, 42.0);
  c.set(p2, 42.0);
  Set springs(points, points);
  FieldRef<double> a = springs.addField<double>("a");
  a.set(springs.add(p0, p1), 1.0);
  a.set(springs.add(p1, p2), 2.0);
  Function func = loadProgram("tests/programs/gemv.sim");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  EXPECT_NEAR(1.0, b.get(p0), 0.0);
  EXPECT_NEAR(2.0, b.get(p1), 0.0);
  EXPECT_NEAR(3.0, b.get(p2), 0.0);
  EXPECT_NEAR(3.0, c.get(p0), 0.0);
  EXPECT_NEAR(13.0, c.get(p1), 0.0);
  EXPECT_NEAR(10.0, c.get(p2), 0.0);
}
static void sim_gemv_grid() {
  Set points;
  FieldRef<double> b = points.addField<double>("b");
  FieldRef<double> c = points.addField<double>("c");
  Set springs(points, {3});
  FieldRef<double> a = springs.addField<double>("a");
  ElementRef p0 = springs.getGridPoint({0}), p1 = springs.getGridPoint({1}),
             p2 = springs.getGridPoint({2});
  b.set(p0, 1.0);
  b.set(p1, 2.0);
  b.set(p2, 3.0);
  c.set(p0, 42.0);
  c.set(p2, 42.0);
  a.set(springs.getGridEdge({0}, 0), 1.0);
  a.set(springs.getGridEdge({1}, 0), 2.0);
  a.set(springs.getGridEdge({2}, 0), 0.0);
  Function func = loadProgram("tests/programs/grid_gemv.sim");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  EXPECT_NEAR(1.0, b.get(p0), 0.0);
  EXPECT_NEAR(2.0, b.get(p1), 0.0);
  EXPECT_NEAR(3.0, b.get(p2), 0.0);
  EXPECT_NEAR(3.0, c.get(p0), 0.0);
  EXPECT_NEAR(13.0, c.get(p1), 0.0);
  EXPECT_NEAR(10.0, c.get(p2), 0.0);
}
static void sim_gemv_grid_2d() {
  Set points;
  FieldRef<double> b = points.addField<double>("b");
  FieldRef<double> c = points.addField<double>("c");
  Set springs(points, {3, 2});
  FieldRef<double> a = springs.addField<double>("a");
  std::vector<ElementRef> p;
  for (int y = 0; y < 2; ++y)
    for (int x = 0; x < 3; ++x) p.push_back(springs.getGridPoint({x, y}));
  for (int i = 0; i < 6; ++i) b.set(p[i], 1.0 + i);
  c.set(p[0], 42.0);
  c.set(p[5], 42.0);
  // weights 1..6 on the x-direction links, 7..12 on the y-direction links
  for (int dir = 0; dir < 2; ++dir)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 3; ++x) a.set(springs.getGridEdge({x, y}, dir), 1.0 + 6 * dir + 3 * y + x);
  Function func = loadProgram("tests/programs/grid_gemv_2d.sim");
  if (!func.defined()) return;
  func.bind("points", &points);
  func.bind("springs", &springs);
  func.runSafe();
  const double want[6] = {100.0, 146.0, 211.0, 181.0, 224.0, 304.0};
  for (int i = 0; i < 6; ++i) {
    EXPECT_NEAR(1.0 + i, b.get(p[i]), 0.0);
    EXPECT_NEAR(want[i], c.get(p[i]), 0.0);
  }
}
// the grid edge set API on the host (graph.h:121-240)
static void grid_set() {
  Set points;
  points.addField<double>("b");
  Set grid(points, {3, 2});
  EXPECT_TRUE(points.getSize() == 6 && grid.getSize() == 12 && grid.getCardinality() == 2);
  EXPECT_TRUE(grid.getKind() == Set::Grid && points.getKind() == Set::Unstructured);
  EXPECT_TRUE(grid.getDimensions() == std::vector<int>({3, 2}));
  EXPECT_TRUE(grid.getGridPoint({2, 1}).getIdent() == 5);
  EXPECT_TRUE(grid.getGridEdge({1, 1}, 1).getIdent() == 9);
  EXPECT_TRUE(grid.getGridPointCoords(ElementRef(4)) == std::vector<int>({1, 1}));
  // edge p*2 + k joins p to its periodic neighbour along k
  EXPECT_TRUE(grid.getEndpoint(grid.getGridEdge({2, 0}, 0), 0).getIdent() == 2);
  EXPECT_TRUE(grid.getEndpoint(grid.getGridEdge({2, 0}, 0), 1).getIdent() == 0);
  EXPECT_TRUE(grid.getEndpoint(grid.getGridEdge({1, 1}, 1), 1).getIdent() == 1);
  bool threw = false;
  try {
    grid.getGridPoint({3, 1});
  } catch (const SimitException &) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  threw = false;
  try {
    Set again(points, {2});  // points is no longer empty
  } catch (const SimitException &) {
    threw = true;
  }
  EXPECT_TRUE(threw);
  threw = false;
  try {
    std::vector<int> perm(12);
    for (int i = 0; i < 12; ++i) perm[i] = 11 - i;
    grid.permuteElements(perm);
  } catch (const SimitException &) {
    threw = true;
  }
  EXPECT_TRUE(threw);
}

// sim_fem_ops PROGRAM DIR: the bar2k set-up, PROGRAM's initializeTet and
// operators (a parsed program that assembles main's force and stiffness and
// writes y = A v); the fields the GPU test checks against the oracle are
// written to DIR/<field>.f64
static int sim_fem_ops(const std::string &program, const std::string &dir) {
  MeshVol mv;
  if (mv.loadTet("tests/golden/fem-bar2k/bar2k.node", "tests/golden/fem-bar2k/bar2k.ele") != 0) {
    std::printf("  FAIL: cannot read bar2k\n");
    return 1;
  }
  Set verts;
  Set tets(verts, verts, verts, verts);
  std::vector<ElementRef> vertRefs;
  initializeFem(mv, verts, tets, vertRefs);
  verts.addField<double, 3>("y");
  verts.addField<double, 3>("f");
  Program p;
  if (p.loadFile(program) != 0) {
    std::printf("  FAIL: %s: %s\n", program.c_str(), p.getDiagnostics().c_str());
    return 1;
  }
  for (const char *fn : {"initializeTet", "operators"}) {
    Function f = p.compile(fn);
    f.bind("verts", &verts);
    f.bind("tets", &tets);
    f.runSafe();
  }
  for (const char *n : {"m", "f", "y", "v"})
    spit(dir + "/" + n + ".f64", verts.field(n)->data(), verts.field(n)->bytes.size());
  for (const char *n : {"W", "B"})
    spit(dir + "/" + n + ".f64", tets.field(n)->data(), tets.field(n)->bytes.size());
  std::printf("sim_fem_ops: %zu vertices, %zu tets\n", (size_t)verts.getSize(),
              (size_t)tets.getSize());
  return 0;
}

int main(int argc, char **argv) {
  init("hip", 8);
  if (argc >= 3 && std::string(argv[1]) == "compile") {
    // compile FILE [FUNCTION...]: parse, check and generate each exported
    // function (or the listed ones) and compile the kernels for gfx950 (no
    // device needed)
    Program p;
    if (p.loadFile(argv[2]) != 0) {
      std::printf("%s: %s", argv[2], p.getDiagnostics().c_str());
      return 1;
    }
    std::vector<std::string> fns(argv + 3, argv + argc);
    if (fns.empty()) fns = p.getExportedFunctionNames();
    int bad = 0;
    for (auto &f : fns) {
      try {
        Function fn = p.compile(f);
        const std::string src = p.getGeneratedSource(f);
        void *img = nullptr;
        size_t bytes = 0;
        std::vector<char> log(1 << 16);
        const int rc = smh_hiprtc_compile(src.c_str(), f.c_str(), "gfx950", &img, &bytes,
                                          log.data(), log.size());
        smh_image_free(img);
        std::printf("%s: %s: %zu bytes of generated HIP, code object %zu bytes%s\n", argv[2],
                    f.c_str(), src.size(), bytes, rc ? " (COMPILE FAILED)" : "");
        if (rc) {
          std::printf("%s\n", log.data());
          ++bad;
        }
      } catch (const std::exception &e) {
        std::printf("%s: %s: not compiled: %s\n", argv[2], f.c_str(), e.what());
        ++bad;
      }
    }
    return bad ? 1 : 0;
  }
  if (argc == 4 && std::string(argv[1]) == "sim_fem_ops") {
    try {
      return sim_fem_ops(argv[2], argv[3]);
    } catch (const std::exception &e) {
      std::printf("  FAIL: exception %s\n", e.what());
      return 1;
    }
  }
  if (argc == 3 && std::string(argv[1]) == "extern_file") {
    try {
      return extern_file(argv[2]);
    } catch (const std::exception &e) {
      std::printf("  FAIL: exception %s\n", e.what());
      return 1;
    }
  }
  std::vector<std::pair<std::string, std::function<void()>>> tests = {
      {"isprings", isprings}, {"cg", cg}, {"springs_lattice", springs_lattice},
      {"mesh_bunny", mesh_bunny}, {"reorder_bunny_springs", reorder_bunny_springs},
      {"reorder_int", reorder_int}, {"reorder_double", reorder_double},
      {"hilbert_curve", hilbert_curve}, {"reorder_bunny_host", reorder_bunny_host},
      {"mesh_load", mesh_load}, {"extern_csr", extern_csr}, {"fem_tet", fem_tet},
      {"reorder_fem_tet", reorder_fem_tet},
      // the same programs and asserted values, parsed from Simit source
      {"sim_isprings", [] { g_programs = true; isprings(); g_programs = false; }},
      {"sim_cg", [] { g_programs = true; cg(); g_programs = false; }},
      {"sim_fem_tet", [] { g_programs = true; fem_tet(); g_programs = false; }},
      {"sim_reorder_fem_tet", [] { g_programs = true; reorder_fem_tet(); g_programs = false; }},
      {"sim_gemv", sim_gemv}, {"sim_gemv_grid", sim_gemv_grid},
      {"sim_gemv_grid_2d", sim_gemv_grid_2d}, {"grid_set", grid_set}};
  for (auto &t : tests) {
    bool run = argc <= 1;
    for (int i = 1; i < argc; ++i) run |= t.first == argv[i];
    if (!run) continue;
    const int before = failures;
    std::printf("[ RUN  ] %s\n", t.first.c_str());
    try {
      t.second();
    } catch (const std::exception &e) {
      std::printf("  FAIL: exception %s\n", e.what());
      ++failures;
    }
    std::printf("[ %s ] %s\n", failures == before ? " OK " : "FAIL", t.first.c_str());
  }
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
