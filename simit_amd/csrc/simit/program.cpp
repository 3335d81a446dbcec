// HIP backend functions for the hot-path Simit programs (see program.h).
#include "program.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "simit_hip.h"

namespace simit {

namespace {
std::string g_backend = "hip";

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

// device mirror of one host field (floats are widened to double)
struct DevField {
  FieldData *host = nullptr;
  void *dev = nullptr;
  size_t devBytes = 0;
  bool output = false;
};

struct FieldSpec {
  std::string name;
  FieldType type;  // Float64 means "a float field" (Float32 also accepted)
  int ncomp;
  bool output;
};
struct ExternSpec {
  std::string name;
  int cardinality;
  std::vector<FieldSpec> fields;
};

size_t devElemBytes(FieldType t) {
  return (t == FieldType::Float64 || t == FieldType::Float32) ? 8 : fieldTypeSize(t);
}

class HipFunction : public backend::Function {
 public:
  explicit HipFunction(std::vector<ExternSpec> spec) : spec(std::move(spec)) {}
  ~HipFunction() override { release(); }

  void bind(const std::string &name, Set *set) override {
    const ExternSpec *es = findExtern(name);
    if (!es) throw SimitException("no extern set named " + name);
    if (set->getCardinality() != es->cardinality)
      throw SimitException("set " + name + " has cardinality " +
                           std::to_string(set->getCardinality()) + ", expected " +
                           std::to_string(es->cardinality));
    for (const FieldSpec &f : es->fields) {
      const FieldData *fd = set->field(f.name);
      if (!fd) throw SimitException("set " + name + " has no field " + f.name);
      const bool isfloat = fd->type == FieldType::Float64 || fd->type == FieldType::Float32;
      const bool ok = (f.type == FieldType::Float64 ? isfloat : fd->type == f.type) &&
                      fd->ncomp == f.ncomp;
      if (!ok) throw SimitException("field " + name + "." + f.name + " has the wrong type");
    }
    sets[name] = set;
    initialized = false;
  }
  void bind(const std::string &name, void *data) override {
    scalars[name] = data;
    initialized = false;
  }
  bool isInitialized() const override {
    if (!initialized) return false;
    for (auto &kv : sets)
      if (kv.second->structureVersion() != versions.at(kv.first)) return false;
    return true;
  }
  std::vector<std::string> externs() const override {
    std::vector<std::string> n;
    for (auto &e : spec) n.push_back(e.name);
    return n;
  }
  void setSolverOptions(const SolverOptions &o) override { opts = o; }
  SolveStats lastSolve() const override { return stats; }

  std::function<void()> init() override {
    release();
    for (auto &e : spec)
      if (!sets.count(e.name)) throw SimitException("extern " + e.name + " is not bound");
    smh_ctx *ctx = device_ctx();
    for (auto &e : spec) {
      Set *s = sets[e.name];
      versions[e.name] = s->structureVersion();
      for (const FieldSpec &f : e.fields) {
        DevField d;
        d.host = s->field(f.name);
        d.output = f.output;
        d.devBytes = devElemBytes(d.host->type) * d.host->ncomp * (size_t)s->getSize();
        check(smh_malloc(ctx, d.devBytes, &d.dev), "device allocation");
        fields[e.name + "." + f.name] = d;
      }
      if (e.cardinality > 0) {
        void *ep = nullptr;
        const size_t bytes = sizeof(int32_t) * (size_t)s->getSize() * e.cardinality;
        check(smh_malloc(ctx, bytes, &ep), "device allocation");
        check(smh_memcpy_htod(ctx, ep, s->getEndpointsData(), bytes), "endpoint upload");
        endpoints[e.name] = (int32_t *)ep;
      }
    }
    unmapArgs(true);
    setup(ctx);
    initialized = true;
    return [this] { this->runDevice(device_ctx()); };
  }

  // host -> device (the reference re-uploads every pushed buffer)
  void unmapArgs(bool updated) override {
    if (!updated) return;
    smh_ctx *ctx = device_ctx();
    std::vector<unsigned char> tmp;
    for (auto &kv : fields) {
      DevField &d = kv.second;
      const size_t n = d.host->bytes.size() / fieldTypeSize(d.host->type);
      if (d.host->type == FieldType::Float32) {
        tmp.resize(n * 8);
        const float *src = (const float *)d.host->data();
        double *dst = (double *)tmp.data();
        for (size_t i = 0; i < n; ++i) dst[i] = src[i];
        check(smh_memcpy_htod(ctx, d.dev, tmp.data(), n * 8), "field upload");
      } else {
        check(smh_memcpy_htod(ctx, d.dev, d.host->data(), d.host->bytes.size()), "field upload");
      }
    }
  }
  // device -> host for the fields a program writes
  void mapArgs() override {
    smh_ctx *ctx = device_ctx();
    std::vector<unsigned char> tmp;
    for (auto &kv : fields) {
      DevField &d = kv.second;
      if (!d.output) continue;
      const size_t n = d.host->bytes.size() / fieldTypeSize(d.host->type);
      if (d.host->type == FieldType::Float32) {
        tmp.resize(n * 8);
        check(smh_memcpy_dtoh(ctx, tmp.data(), d.dev, n * 8), "field download");
        const double *src = (const double *)tmp.data();
        float *dst = (float *)d.host->data();
        for (size_t i = 0; i < n; ++i) dst[i] = (float)src[i];
      } else {
        check(smh_memcpy_dtoh(ctx, d.host->data(), d.dev, d.host->bytes.size()), "field download");
      }
    }
  }

 protected:
  virtual void setup(smh_ctx *ctx) = 0;
  virtual void runDevice(smh_ctx *ctx) = 0;

  void *dev(const std::string &key) { return fields.at(key).dev; }
  Set *set(const std::string &name) { return sets.at(name); }
  double scalar(const std::string &name, double dflt) {
    auto it = scalars.find(name);
    return it == scalars.end() ? dflt : *(const double *)it->second;
  }
  void *alloc(size_t bytes) {
    void *p = nullptr;
    check(smh_malloc(device_ctx(), bytes, &p), "device allocation");
    owned.push_back(p);
    return p;
  }
  void buildIndex(const std::string &vertices, const std::string &edges) {
    Set *v = set(vertices), *e = set(edges);
    const uint64_t ne = (uint64_t)e->getSize();
    const uint32_t card = (uint32_t)e->getCardinality();
    const int32_t *eps[1] = {endpoints.at(edges)};
    check(smh_index_build(device_ctx(), (uint32_t)v->getSize(), 1, &ne, &card, eps, 0, &index),
          "index build");
    check(smh_index_get(index, &view), "index view");
  }
  void solve(const smh_bsr &A, const double *b, double *x, smh_cg_opts o) {
    if (opts.maxiters >= 0) o.maxiters = opts.maxiters;
    if (opts.tol >= 0) o.tol = opts.tol;
    if (opts.precond >= 0) o.precond = opts.precond;
    if (opts.mixed >= 0) o.mixed = opts.mixed;
    smh_cg_stats st{};
    check(smh_cg_solve(device_ctx(), &A, b, x, &o, &st), "CG solve");
    stats.iters = st.iters;
    stats.reliable_updates = st.reliable_updates;
    stats.status = st.status;
    stats.true_rel = st.true_rel;
    stats.seconds = st.seconds;
    stats.precond_used = st.precond_used;
    stats.fallbacks = st.fallbacks;
  }
  void release() {
    smh_ctx *ctx = nullptr;
    if (!fields.empty() || !owned.empty() || index || !endpoints.empty()) ctx = device_ctx();
    if (!ctx) return;
    smh_ctx_synchronize(ctx);
    for (auto &kv : fields) smh_free(ctx, kv.second.dev);
    for (auto &kv : endpoints) smh_free(ctx, kv.second);
    for (void *p : owned) smh_free(ctx, p);
    if (index) smh_index_destroy(ctx, index);
    fields.clear();
    endpoints.clear();
    owned.clear();
    index = nullptr;
  }
  const ExternSpec *findExtern(const std::string &n) const {
    for (auto &e : spec)
      if (e.name == n) return &e;
    return nullptr;
  }

  std::vector<ExternSpec> spec;
  std::map<std::string, Set *> sets;
  std::map<std::string, uint64_t> versions;
  std::map<std::string, void *> scalars;
  std::map<std::string, DevField> fields;
  std::map<std::string, int32_t *> endpoints;
  std::vector<void *> owned;
  smh_index *index = nullptr;
  smh_index_view view{};
  bool initialized = false;
  SolverOptions opts;
  SolveStats stats;
};

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
                      {{"k", FieldType::Float64, 1, false}, {"l0", FieldType::Float64, 1, false}}}}),
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
                     {"springs", 2, {{"a", FieldType::Float64, 1, false}}}}) {}

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
                      {{"u", FieldType::Float64, 1, false}, {"l", FieldType::Float64, 1, false}}}}) {}

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
                     {"hexes", 8, {{"k", FieldType::Float64, 1, false}}}}) {}

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

const std::vector<std::string> kPrograms = {"isprings", "springs_implicit", "cg", "fem_static",
                                            "thermal_hex"};

std::shared_ptr<backend::Function> makeFunction(const std::string &program) {
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
  if (impl) impl->setSolverOptions(o);
}
SolveStats Function::lastSolve() const { return impl ? impl->lastSolve() : SolveStats{}; }

int Program::loadFile(const std::string &path) {
  std::string base = path.substr(path.find_last_of('/') == std::string::npos
                                     ? 0
                                     : path.find_last_of('/') + 1);
  if (base.size() > 4 && base.substr(base.size() - 4) == ".sim") base.resize(base.size() - 4);
  return loadBuiltin(base);
}
int Program::loadBuiltin(const std::string &name) {
  if (std::find(kPrograms.begin(), kPrograms.end(), name) == kPrograms.end()) return 1;
  program = name;
  return 0;
}
std::vector<std::string> Program::getFunctionNames() const {
  if (program.empty()) return {};
  return {"main"};
}
Function Program::compile(const std::string &function) {
  if (program.empty()) throw SimitException("no program loaded");
  if (function != "main" && function != "timestep" && function != program)
    throw SimitException("no function " + function + " in program " + program);
  return Function(makeFunction(program));
}

}  // namespace simit
