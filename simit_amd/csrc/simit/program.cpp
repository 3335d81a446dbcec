// HIP backend functions for the hot-path Simit programs (see program.h).
#include "program.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <mutex>

#include <dlfcn.h>

#include "simit_hip.h"
#include "hip_function.h"
#include "frontend.h"
#include "sim_function.h"

namespace simit {

namespace {
std::string g_backend = "hip";

}  // namespace

namespace hipimpl {
// one context per process (device 0), as the reference GPU backend uses one
// CUDA context on device 0 (gpu_function.cpp:115)
smh_ctx *device_ctx() {
  static smh_ctx *ctx = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    if (smh_ctx_create(0, &ctx) != SMH_OK) ctx = nullptr;
  });
  if (!ctx) throw SimitException("no HIP device available for the hip backend");
  return ctx;
}

void check(int rc, const char *what) {
  if (rc != SMH_OK) {
    smh_ctx *c = device_ctx();
    throw SimitException(std::string(what) + ": " + smh_ctx_last_error(c));
  }
}

}  // namespace hipimpl

namespace {
using namespace hipimpl;

// ---- implicit springs (test/input/program/isprings.sim:71-103) ------------------
class SpringsFunction : public HipFunction {
 public:
  SpringsFunction(smh_springs_params p, smh_cg_opts o)
      : HipFunction({{"points", 0,
                      {{"x", FieldType::Float64, 3, true},
                       {"v", FieldType::Float64, 3, true},
                       {"m", FieldType::Float64, 1, false},
                       {"fixed", FieldType::Bool, 1, false}}},
                     {"springs", 2,
                      {{"k", FieldType::Float64, 1, false}, {"l0", FieldType::Float64, 1, false}}}},
                    {}, {},
                    "smh_assemble_springs (k_assemble<SpringOp>, k_isolated) -> smh_cg_solve "
                    "(tile SpMV / update / reductions per iteration) -> smh_axpby (x += h v)"),
        params(p),
        cg(o) {}

 protected:
  void setup(smh_ctx *) override {
    buildIndex("points", "springs");
    vals = (double *)alloc(sizeof(double) * 9 * (view.nnz ? view.nnz : 1));
    rhs = (double *)alloc(sizeof(double) * 3 * (size_t)set("points")->getSize() + 8);
  }
  void runDevice(smh_ctx *ctx) override {
    Set *pts = set("points");
    smh_springs_set s{(uint32_t)pts->getSize(), (uint64_t)set("springs")->getSize(),
                      endpoints.at("springs"), (const double *)dev("points.x"),
                      (const double *)dev("points.v"), (const double *)dev("points.m"),
                      (const uint8_t *)dev("points.fixed"), (const double *)dev("springs.k"),
                      (const double *)dev("springs.l0")};
    check(smh_assemble_springs(ctx, index, &params, &s, SMH_F64, vals, nullptr, rhs), "assembly");
    smh_bsr A{view.nrows, 3, view.nnz, view.rowptr, view.colidx, SMH_F64, vals, nullptr};
    // v = cg(MDK, f) from v; points.x = points.x + h * points.v
    solve(A, rhs, (double *)dev("points.v"), cg);
    check(smh_axpby(ctx, 3 * (uint64_t)pts->getSize(), params.h, (const double *)dev("points.v"),
                    1.0, (double *)dev("points.x")),
          "x update");
  }

 private:
  smh_springs_params params;
  smh_cg_opts cg;
  double *vals = nullptr, *rhs = nullptr;
};

// ---- test/input/program/cg.sim -------------------------------------------------------
class CgProgramFunction : public HipFunction {
 public:
  CgProgramFunction()
      : HipFunction({{"points", 0,
                      {{"b", FieldType::Float64, 1, false}, {"c", FieldType::Float64, 1, true}}},
                     {"springs", 2, {{"a", FieldType::Float64, 1, false}}}},
                    {}, {},
                    "smh_assemble_edge_blocks (A = I - 0.01 K) -> smh_cg_solve (5 iterations, "
                    "norm(r) > tol)") {}

 protected:
  void setup(smh_ctx *) override {
    buildIndex("points", "springs");
    vals = (double *)alloc(sizeof(double) * (view.nnz ? view.nnz : 1));
  }
  void runDevice(smh_ctx *ctx) override {
    // A = I - 0.01 * (map f to springs reduce +)  (cg.sim:55-58)
    smh_edge_blocks e{view.nrows, (uint64_t)set("springs")->getSize(), 1, endpoints.at("springs"),
                      (const double *)dev("springs.a"), {1.0, -1.0, -1.0, 1.0}, 1.0, -0.01};
    check(smh_assemble_edge_blocks(ctx, index, &e, SMH_F64, vals, nullptr), "assembly");
    smh_bsr A{view.nrows, 1, view.nnz, view.rowptr, view.colidx, SMH_F64, vals, nullptr};
    double *c = (double *)dev("points.c");
    check(smh_memset(ctx, c, 0, sizeof(double) * view.nrows), "x0");  // xguess = 0
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 5;
    o.tol = 1e-6;
    o.rel = 0;
    o.norm_form = 1;
    solve(A, (const double *)dev("points.b"), c, o);
  }

 private:
  double *vals = nullptr;
};

// ---- linear-elastic tets, statics ---------------------------------------------------
class FemStaticFunction : public HipFunction {
 public:
  FemStaticFunction()
      : HipFunction({{"verts", 0,
                      {{"x", FieldType::Float64, 3, false},
                       {"c", FieldType::Int32, 1, false},
                       {"v", FieldType::Float64, 3, true}}},
                     {"tets", 4,
                      {{"u", FieldType::Float64, 1, false}, {"l", FieldType::Float64, 1, false}}}},
                    {}, {},
                    "smh_assemble_tet_linear (FP32 hi/lo) -> smh_cg_solve (mixed block-Jacobi "
                    "PCG, FP64 reliable updates)") {}

 protected:
  void setup(smh_ctx *) override {
    buildIndex("verts", "tets");
    const size_t nv = (size_t)(view.nnz ? view.nnz : 1) * 9;
    hi = (float *)alloc(sizeof(float) * nv);
    lo = (float *)alloc(sizeof(float) * nv);
    rhs = (double *)alloc(sizeof(double) * 3 * (size_t)set("verts")->getSize() + 8);
  }
  void runDevice(smh_ctx *ctx) override {
    smh_tet_params p{1.0, 0.0, {scalar("load_x", 0.0), scalar("load_y", 0.0),
                                scalar("load_z", -9.81e3)}, 1.0};
    smh_tet_set s{(uint32_t)set("verts")->getSize(), (uint64_t)set("tets")->getSize(),
                  endpoints.at("tets"), (const double *)dev("verts.x"),
                  (const int32_t *)dev("verts.c"), (const double *)dev("tets.u"),
                  (const double *)dev("tets.l")};
    check(smh_assemble_tet_linear(ctx, index, &p, &s, SMH_F32X2, hi, lo, rhs), "assembly");
    smh_bsr A{view.nrows, 3, view.nnz, view.rowptr, view.colidx, SMH_F32X2, hi, lo};
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 20000;
    o.tol = 1e-8;
    o.precond = SMH_PRECOND_BLOCK_JACOBI;
    o.mixed = 1;
    solve(A, rhs, (double *)dev("verts.v"), o);
  }

 private:
  float *hi = nullptr, *lo = nullptr;
  double *rhs = nullptr;
};

// ---- scalar trilinear hex Poisson ---------------------------------------------------------
class ThermalHexFunction : public HipFunction {
 public:
  ThermalHexFunction()
      : HipFunction({{"nodes", 0,
                      {{"dirichlet", FieldType::Int32, 1, false},
                       {"T", FieldType::Float64, 1, true}}},
                     {"hexes", 8, {{"k", FieldType::Float64, 1, false}}}},
                    {}, {},
                    "smh_assemble_hex_poisson (FP32 hi/lo) -> smh_cg_solve (mixed Jacobi PCG)") {}

 protected:
  void setup(smh_ctx *) override {
    buildIndex("nodes", "hexes");
    const size_t nv = view.nnz ? view.nnz : 1;
    hi = (float *)alloc(sizeof(float) * nv);
    lo = (float *)alloc(sizeof(float) * nv);
    rhs = (double *)alloc(sizeof(double) * (size_t)set("nodes")->getSize() + 8);
  }
  void runDevice(smh_ctx *ctx) override {
    smh_hex_params p{scalar("h", 1.0), scalar("source", 1.0), 1.0};
    smh_hex_set s{(uint32_t)set("nodes")->getSize(), (uint64_t)set("hexes")->getSize(),
                  endpoints.at("hexes"), (const int32_t *)dev("nodes.dirichlet"),
                  (const double *)dev("hexes.k")};
    check(smh_assemble_hex_poisson(ctx, index, &p, &s, SMH_F32X2, hi, lo, rhs), "assembly");
    smh_bsr A{view.nrows, 1, view.nnz, view.rowptr, view.colidx, SMH_F32X2, hi, lo};
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 20000;
    o.tol = 1e-8;
    o.precond = SMH_PRECOND_JACOBI;
    o.mixed = 1;
    solve(A, rhs, (double *)dev("nodes.T"), o);
  }

 private:
  float *hi = nullptr, *lo = nullptr;
  double *rhs = nullptr;
};

// ---- test/input/program/femTet.sim: neo-Hookean tets ------------------------------------
// initializeTet: verts.m = map precomputeTetMat to tets (Tet.B, Tet.W, Vert.m)
class FemTetInitFunction : public HipFunction {
 public:
  FemTetInitFunction()
      : HipFunction({{"verts", 0,
                      {{"x", FieldType::Float64, 3, false}, {"m", FieldType::Float64, 1, true}}},
                     {"tets", 4,
                      {{"W", FieldType::Float64, 1, true}, {"B", FieldType::Float64, 9, true}}}},
                    {}, {}, "smh_femtet_precompute (k_femtet_geom, k_femtet_mass)") {}

 protected:
  void setup(smh_ctx *) override { buildIndex("verts", "tets"); }
  void runDevice(smh_ctx *ctx) override {
    smh_femtet_params p{0.01, 1000.0, {0.0, -10.0, 0.0}};
    smh_femtet_set s{(uint32_t)set("verts")->getSize(), (uint64_t)set("tets")->getSize(),
                     endpoints.at("tets"), (const double *)dev("verts.x"), nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr};
    check(smh_femtet_precompute(ctx, index, &p, &s, (double *)dev("tets.W"),
                                (double *)dev("tets.B"), (double *)dev("verts.m")),
          "precompute");
  }
};

// main: A = map compute_stiffness(h), f = map compute_force(h), the inline
// CG from v (r'r > 1e-6, 100 iterations), verts.v = x, verts.x = h*x + x
class FemTetStepFunction : public HipFunction {
 public:
  FemTetStepFunction()
      : HipFunction({{"verts", 0,
                      {{"x", FieldType::Float64, 3, true},
                       {"v", FieldType::Float64, 3, true},
                       {"c", FieldType::Int32, 1, false}}},
                     {"tets", 4,
                      {{"u", FieldType::Float64, 1, false},
                       {"l", FieldType::Float64, 1, false},
                       {"W", FieldType::Float64, 1, false},
                       {"B", FieldType::Float64, 9, false}}}},
                    {}, {},
                    "smh_assemble_tet_neohookean (k_assemble<FemTetOp>) -> smh_cg_solve "
                    "(FP64, absolute r'r) -> smh_axpby (x += h v)") {}

 protected:
  void setup(smh_ctx *) override {
    buildIndex("verts", "tets");
    vals = (double *)alloc(sizeof(double) * 9 * (view.nnz ? view.nnz : 1));
    rhs = (double *)alloc(sizeof(double) * 3 * (size_t)set("verts")->getSize() + 8);
  }
  void runDevice(smh_ctx *ctx) override {
    const double h = 0.01;
    smh_femtet_params p{h, 1000.0, {0.0, -10.0, 0.0}};
    Set *vs = set("verts");
    smh_femtet_set s{(uint32_t)vs->getSize(), (uint64_t)set("tets")->getSize(),
                     endpoints.at("tets"), (const double *)dev("verts.x"),
                     (const double *)dev("verts.v"), (const int32_t *)dev("verts.c"),
                     (const double *)dev("tets.u"), (const double *)dev("tets.l"),
                     (const double *)dev("tets.W"), (const double *)dev("tets.B")};
    check(smh_assemble_tet_neohookean(ctx, index, &p, &s, SMH_F64, vals, nullptr, rhs),
          "assembly");
    smh_bsr A{view.nrows, 3, view.nnz, view.rowptr, view.colidx, SMH_F64, vals, nullptr};
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 100;
    o.tol = 1e-6;
    o.rel = 0;
    solve(A, rhs, (double *)dev("verts.v"), o);
    check(smh_axpby(ctx, 3 * (uint64_t)vs->getSize(), h, (const double *)dev("verts.v"), 1.0,
                    (double *)dev("verts.x")),
          "x update");
  }

 private:
  double *vals = nullptr, *rhs = nullptr;
};

// ---- extern sparse matrices bound as CSR TensorData -------------------------------------
// A = -A on an extern int matrix: the reference's extern-CSR function test
// (test/function-tests.cpp:96-139, an IndexExpr -A(i,j) over extern A).
class ExternNegFunction : public HipFunction {
 public:
  ExternNegFunction()
      : HipFunction({{"V", 0, {}}}, {{"A", "V", 1, ComponentType::Int, true}}, {},
                    "smh_axpby (A = -A over the bound CSR values)") {}

 protected:
  void setup(smh_ctx *) override {}
  void runDevice(smh_ctx *ctx) override {
    smh_bsr A = bsrOf("A");
    check(smh_axpby(ctx, A.nnz, 0.0, (const double *)A.vals, -1.0, (double *)A.vals),
          "negation");
  }
};

// y = A * x over an extern CSR matrix of b x b blocks (the blocked row loop
// of lower_indexexprs.cpp:630-667, in its ascending order)
class ExternSpmvFunction : public HipFunction {
 public:
  explicit ExternSpmvFunction(int b)
      : HipFunction({{"V", 0, {}}}, {{"A", "V", b, ComponentType::Double, false}},
                    {{"x", "V", b, ComponentType::Double, false},
                     {"y", "V", b, ComponentType::Double, true}},
                    "smh_bsrmv (exact order) y = A x") {}

 protected:
  void setup(smh_ctx *) override {}
  void runDevice(smh_ctx *ctx) override {
    smh_bsr A = bsrOf("A");
    check(smh_bsrmv(ctx, &A, devVector("x"), devVector("y"), 1), "SpMV");
  }
};

// x = cg(A, b) over an extern CSR matrix: relative 1e-8, (block) Jacobi,
// FP64; x holds the initial guess
class ExternCgFunction : public HipFunction {
 public:
  explicit ExternCgFunction(int b)
      : HipFunction({{"V", 0, {}}}, {{"A", "V", b, ComponentType::Double, false}},
                    {{"b", "V", b, ComponentType::Double, false},
                     {"x", "V", b, ComponentType::Double, true}},
                    "smh_cg_solve (FP64 (block-)Jacobi PCG to a relative 1e-8)"),
        bdim(b) {}

 protected:
  void setup(smh_ctx *) override {}
  void runDevice(smh_ctx *) override {
    smh_bsr A = bsrOf("A");
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 20000;
    o.tol = 1e-8;
    o.rel = 1;
    o.precond = bdim == 3 ? SMH_PRECOND_BLOCK_JACOBI : SMH_PRECOND_JACOBI;
    solve(A, devVector("b"), devVector("x"), o);
  }

 private:
  int bdim;
};

}  // namespace

void hipimpl::HipFunction::hipInfo(std::ostream &os) {
  char buf[512] = {0};
  if (smh_ctx_describe(device_ctx(), buf, sizeof(buf)) == SMH_OK) os << buf << "\n";
  Dl_info info{};
  if (dladdr((const void *)&smh_cg_solve, &info) && info.dli_fname)
    os << "code object: " << info.dli_fname << " (hand-written HIP kernels for gfx950)\n";
}

namespace {

const std::vector<std::string> kPrograms = {"isprings", "springs_implicit", "cg", "fem_static",
                                            "thermal_hex", "femTet", "extern_neg", "extern_spmv",
                                            "extern_spmv_3x3", "extern_cg", "extern_cg_3x3"};

std::shared_ptr<backend::Function> makeFunction(const std::string &program,
                                                const std::string &function) {
  if (program == "femTet") {
    if (function == "initializeTet") return std::make_shared<FemTetInitFunction>();
    return std::make_shared<FemTetStepFunction>();
  }
  if (program == "isprings") {
    // test/input/program/isprings.sim: h = 1e-2, viscous = 1e-1, gravity
    // -9.81 z, MD(p,p) = 1.0 for fixed points, CG tol 1e-5 on r'r, 25 iterations
    smh_springs_params p{1e-2, 1e-1, {0.0, 0.0, -9.81}, 1.0, 0};
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 25;
    o.tol = 1e-5;
    o.rel = 0;
    return std::make_shared<SpringsFunction>(p, o);
  }
  if (program == "springs_implicit") {
    smh_springs_params p{1e-2, 0.0, {0.0, 0.0, -9.8}, 1.0, 1};
    smh_cg_opts o;
    smh_cg_default_opts(&o);
    o.maxiters = 500;
    o.tol = 1e-8;
    o.rel = 1;
    return std::make_shared<SpringsFunction>(p, o);
  }
  if (program == "cg") return std::make_shared<CgProgramFunction>();
  if (program == "fem_static") return std::make_shared<FemStaticFunction>();
  if (program == "thermal_hex") return std::make_shared<ThermalHexFunction>();
  if (program == "extern_neg") return std::make_shared<ExternNegFunction>();
  if (program == "extern_spmv") return std::make_shared<ExternSpmvFunction>(1);
  if (program == "extern_spmv_3x3") return std::make_shared<ExternSpmvFunction>(3);
  if (program == "extern_cg") return std::make_shared<ExternCgFunction>(1);
  if (program == "extern_cg_3x3") return std::make_shared<ExternCgFunction>(3);
  return nullptr;
}

}  // namespace

// ---- public API -------------------------------------------------------------------------
void init(const Settings &settings) {
  if (settings.backend != "hip" && settings.backend != "gpu")
    throw SimitException("Invalid backend: " + settings.backend + " (this build runs \"hip\")");
  if (settings.floatSize != 4 && settings.floatSize != 8)
    throw SimitException("Invalid float bytes: " + std::to_string(settings.floatSize));
  g_backend = "hip";
}
void init(const std::string &backend, int floatSize) {
  Settings s;
  s.backend = backend;
  s.floatSize = floatSize;
  init(s);
}
const std::string &currentBackend() { return g_backend; }

void Function::bind(const std::string &name, Set *set) {
  if (!impl) throw SimitException("bind on an undefined function");
  impl->bind(name, set);
}
void Function::bind(const std::string &name, void *data) {
  if (!impl) throw SimitException("bind on an undefined function");
  impl->bind(name, data);
}
void Function::bind(const std::string &name, const TensorType &ttype, void *data) {
  if (!impl) throw SimitException("bind on an undefined function");
  (void)ttype;  // dense externs are widened to double on the device whatever their type
  impl->bind(name, data);
}
void Function::bind(const std::string &name, TensorData &data) {
  if (!impl) throw SimitException("bind on an undefined function");
  impl->bind(name, data);
}
void Function::print(std::ostream &os) const {
  if (impl) impl->print(os);
}
void Function::printMachine(std::ostream &os) const {
  if (impl) impl->printMachine(os);
}
std::ostream &operator<<(std::ostream &os, const Function &f) {
  f.print(os);
  return os;
}
bool backend::Function::hasBindable(const std::string &name) const {
  for (auto &b : bindables)
    if (b.name == name) return true;
  return false;
}
size_t TensorType::getSize() const {
  size_t n = 1;
  for (int d : dimensions) n *= (size_t)d;
  for (int d : blockDimensions) n *= (size_t)d;
  return n;
}
void Function::init() {
  if (!impl) throw SimitException("init on an undefined function");
  funcPtr = impl->init();
}
// function.cpp:121-129: init if needed, unmap (push), run, map (pull)
void Function::runSafe() {
  if (!impl) throw SimitException("runSafe on an undefined function");
  if (!impl->isInitialized()) init();
  impl->unmapArgs(true);
  funcPtr();
  impl->mapArgs();
}
void Function::mapArgs() {
  if (impl) impl->mapArgs();
}
void Function::unmapArgs(bool updated) {
  if (impl) impl->unmapArgs(updated);
}
void Function::setSolverOptions(const SolverOptions &o) {
  if (auto *c = dynamic_cast<hip::SolverControl *>(impl.get())) c->setSolverOptions(o);
}
SolveStats Function::lastSolve() const {
  if (auto *c = dynamic_cast<const hip::SolverControl *>(impl.get())) return c->lastSolve();
  return SolveStats{};
}

int Program::loadFile(const std::string &path) {
  std::ifstream in(path);
  if (!in.good()) {
    diagnostics = "Could not load file: " + path;
    return 2;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  return loadString(ss.str());
}
int Program::loadString(const std::string &text) {
  std::vector<fe::ParseError> errors;
  std::shared_ptr<fe::ProgramAST> p = fe::parse(text, &errors);
  diagnostics.clear();
  for (auto &e : errors) diagnostics += e.str() + "\n";
  if (!p) return 1;
  ast = p;
  program.clear();
  return 0;
}
int Program::loadBuiltin(const std::string &name) {
  if (std::find(kPrograms.begin(), kPrograms.end(), name) == kPrograms.end()) return 1;
  program = name;
  ast.reset();
  return 0;
}
std::vector<std::string> Program::getFunctionNames() const {
  if (ast) {
    std::vector<std::string> names;
    for (auto &f : ast->funcs) names.push_back(f->name);
    return names;
  }
  if (program.empty()) return {};
  if (program == "femTet") return {"initializeTet", "main"};
  return {"main"};
}
std::vector<std::string> Program::getExportedFunctionNames() const {
  std::vector<std::string> names;
  if (ast) {
    for (auto &f : ast->funcs)
      if (f->exported) names.push_back(f->name);
    return names;
  }
  return getFunctionNames();
}
std::string Program::getGeneratedSource(const std::string &function) const {
  if (!ast) throw SimitException("no parsed program loaded");
  return hipimpl::generatedSource(*ast, function);
}
Function Program::compile(const std::string &function) {
  if (ast) return Function(hipimpl::makeSimFunction(ast, function));
  if (program.empty()) throw SimitException("no program loaded");
  const std::vector<std::string> fs = getFunctionNames();
  if (std::find(fs.begin(), fs.end(), function) == fs.end() && function != "timestep" &&
      function != program)
    throw SimitException("no function " + function + " in program " + program);
  return Function(makeFunction(program, function));
}

}  // namespace simit
