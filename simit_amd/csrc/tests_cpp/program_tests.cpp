// Program tests through the C++ simit API on the HIP backend, written the
// way the reference's gtest program tests are (test/program-isprings.cpp,
// test/program-cg-test.cpp): build sets, load the program, bind, runSafe,
// compare against the reference's asserted values.
//
// Usage: simit_program_tests [name ...]; exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "simit/graph.h"
#include "simit/mesh.h"
#include "simit/program.h"

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

static Function load(const std::string &program) {
  Program p;
  if (p.loadFile(program + ".sim") != 0) {
    std::printf("  FAIL: cannot load %s\n", program.c_str());
    ++failures;
    return Function();
  }
  return p.compile("main");
}

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

int main(int argc, char **argv) {
  init("hip", 8);
  std::vector<std::pair<std::string, std::function<void()>>> tests = {
      {"isprings", isprings}, {"cg", cg}, {"springs_lattice", springs_lattice},
      {"mesh_bunny", mesh_bunny}};
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
