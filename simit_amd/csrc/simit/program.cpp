// HIP backend functions for the hot-path Simit programs (see program.h).
#include "program.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include <dlfcn.h>

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

// An extern sparse matrix tensor[rows, rows](tensor[b, b](T)) bound as CSR
// TensorData, and an extern dense vector tensor[rows](tensor[b](T)) bound by
// pointer.  Values are widened to double on the device and narrowed back by
// mapArgs (as float fields are).
struct MatrixSpec {
  std::string name, rows;
  int bdim;
  ComponentType type;
  bool output;
};
struct VectorSpec {
  std::string name, rows;
  int bdim;
  ComponentType type;
  bool output;
};
struct DevMatrix {
  TensorData *host = nullptr;  // the caller's (borrowed) CSR arrays
  uint32_t *rowptr = nullptr, *colidx = nullptr;
  double *vals = nullptr;
  uint32_t nrows = 0;
  uint64_t nnz = 0;
};

class HipFunction : public backend::Function, public hip::SolverControl {
 public:
  explicit HipFunction(std::vector<ExternSpec> spec, std::vector<MatrixSpec> mats = {},
                       std::vector<VectorSpec> vecs = {}, std::string pipeline = "")
      : spec(std::move(spec)), mspec(std::move(mats)), vspec(std::move(vecs)),
        pipeline(std::move(pipeline)) {
    for (auto &e : this->spec) bindables.push_back({e.name, "set"});
    for (auto &m : mspec) bindables.push_back({m.name, "matrix"});
    for (auto &v : vspec) bindables.push_back({v.name, "vector"});
  }
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
  // a dense extern vector (host pointer) or a scalar global
  void bind(const std::string &name, void *data) override {
    if (!data) throw SimitException("null data bound to " + name);
    if (findVector(name)) vectors[name] = data;
    else scalars[name] = data;
    initialized = false;
  }
  // an extern sparse matrix in CSR form (function.h:58, tensor_data.h)
  void bind(const std::string &name, TensorData &data) override {
    const MatrixSpec *ms = findMatrix(name);
    if (!ms) throw SimitException("no extern sparse matrix named " + name);
    if (data.getKind() != TensorData::Sparse || !data.getRowPtr() || data.getRowLen() < 1 ||
        data.getDataLen() < 0 || (data.getDataLen() > 0 && (!data.getColInd() || !data.getData())))
      throw SimitException("sparse tensor data bound to " + name + " is malformed");
    matrices[name].host = &data;
    initialized = false;
  }
  bool isInitialized() override {
    if (!initialized) return false;
    for (auto &kv : sets)
      if (kv.second->structureVersion() != versions.at(kv.first)) return false;
    return true;
  }
  void setSolverOptions(const SolverOptions &o) override { opts = o; }
  SolveStats lastSolve() const override { return stats; }

  void print(std::ostream &os) const override {
    os << "hip function {\n";
    for (auto &e : spec) {
      os << "  extern " << e.name << " : set";
      if (e.cardinality) os << " of " << e.cardinality << "-element tuples";
      os << " {";
      for (size_t i = 0; i < e.fields.size(); ++i)
        os << (i ? ", " : "") << e.fields[i].name << (e.fields[i].output ? " (out)" : "");
      os << "}";
      auto it = sets.find(e.name);
      if (it != sets.end()) os << "  bound: " << it->second->getSize() << " elements";
      os << "\n";
    }
    for (auto &m : mspec) {
      os << "  extern " << m.name << " : matrix[" << m.rows << "," << m.rows << "] of "
         << m.bdim << "x" << m.bdim << " blocks, CSR" << (m.output ? " (out)" : "");
      auto it = matrices.find(m.name);
      if (it != matrices.end() && it->second.host)
        os << "  bound: " << it->second.host->getRowLen() - 1 << " rows, "
           << it->second.host->getDataLen() << " blocks";
      os << "\n";
    }
    for (auto &v : vspec)
      os << "  extern " << v.name << " : vector[" << v.rows << "] of " << v.bdim
         << "-vectors" << (v.output ? " (out)" : "") << "\n";
    os << "  run: " << pipeline << "\n}\n";
  }
  void printMachine(std::ostream &os) const override {
    hipInfo(os);
    os << "  launches per run: " << pipeline << "\n";
  }

  std::function<void()> init() override {
    release();
    for (auto &e : spec)
      if (!sets.count(e.name)) throw SimitException("extern " + e.name + " is not bound");
    for (auto &m : mspec)
      if (!matrices.count(m.name) || !matrices[m.name].host)
        throw SimitException("extern " + m.name + " is not bound");
    for (auto &v : vspec)
      if (!vectors.count(v.name)) throw SimitException("extern " + v.name + " is not bound");
    smh_ctx *ctx = device_ctx();
    for (auto &m : mspec) uploadMatrixStructure(ctx, m);
    for (auto &v : vspec) {
      DevVector d;
      d.host = vectors[v.name];
      d.n = (size_t)rowsOf(v.rows) * v.bdim;
      check(smh_malloc(ctx, 8 * (d.n ? d.n : 1), &d.dev), "device allocation");
      devVectors[v.name] = d;
    }
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
    for (auto &m : mspec) {
      DevMatrix &d = matrices[m.name];
      widenUpload(ctx, d.vals, d.host->getData(), m.type, d.nnz * m.bdim * m.bdim);
    }
    for (auto &v : vspec) {
      DevVector &d = devVectors[v.name];
      widenUpload(ctx, (double *)d.dev, d.host, v.type, d.n);
    }
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
    for (auto &m : mspec)
      if (m.output) {
        DevMatrix &d = matrices[m.name];
        narrowDownload(ctx, d.host->getData(), d.vals, m.type, d.nnz * m.bdim * m.bdim);
      }
    for (auto &v : vspec)
      if (v.output) {
        DevVector &d = devVectors[v.name];
        narrowDownload(ctx, d.host, (const double *)d.dev, v.type, d.n);
      }
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

  struct DevVector {
    void *host = nullptr;
    void *dev = nullptr;
    size_t n = 0;
  };
  uint32_t rowsOf(const std::string &setName) {
    auto it = sets.find(setName);
    if (it == sets.end()) throw SimitException("set " + setName + " is not bound");
    return (uint32_t)it->second->getSize();
  }
  // checks the CSR structure against its row set and uploads it (u32)
  void uploadMatrixStructure(smh_ctx *ctx, const MatrixSpec &m) {
    DevMatrix &d = matrices[m.name];
    TensorData &t = *d.host;
    const uint32_t n = rowsOf(m.rows);
    if ((uint32_t)t.getRowLen() != n + 1)
      throw SimitException("matrix " + m.name + ": rowLen " + std::to_string(t.getRowLen()) +
                           " for " + std::to_string(n) + " rows of " + m.rows);
    const int *rp = t.getRowPtr(), *ci = t.getColInd();
    if (rp[0] != 0 || rp[n] != t.getDataLen())
      throw SimitException("matrix " + m.name + ": rowPtr must run from 0 to dataLen");
    for (uint32_t i = 0; i < n; ++i)
      if (rp[i + 1] < rp[i]) throw SimitException("matrix " + m.name + ": rowPtr decreases");
    for (int k = 0; k < t.getDataLen(); ++k)
      if (ci[k] < 0 || (uint32_t)ci[k] >= n)
        throw SimitException("matrix " + m.name + ": column index out of range");
    d.nrows = n;
    d.nnz = (uint64_t)t.getDataLen();
    const size_t bb = (size_t)m.bdim * m.bdim;
    void *p = nullptr;
    check(smh_malloc(ctx, 4 * (size_t)(n + 1), &p), "device allocation");
    d.rowptr = (uint32_t *)p;
    check(smh_memcpy_htod(ctx, p, rp, 4 * (size_t)(n + 1)), "rowPtr upload");
    check(smh_malloc(ctx, 4 * (d.nnz ? d.nnz : 1), &p), "device allocation");
    d.colidx = (uint32_t *)p;
    if (d.nnz) check(smh_memcpy_htod(ctx, p, ci, 4 * d.nnz), "colInd upload");
    check(smh_malloc(ctx, 8 * (d.nnz ? d.nnz * bb : 1), &p), "device allocation");
    d.vals = (double *)p;
  }
  static void widenUpload(smh_ctx *ctx, double *dev, const void *host, ComponentType t,
                          size_t n) {
    if (!n) return;
    if (t == ComponentType::Double) {
      check(smh_memcpy_htod(ctx, dev, host, 8 * n), "tensor upload");
      return;
    }
    std::vector<double> tmp(n);
    for (size_t i = 0; i < n; ++i)
      tmp[i] = t == ComponentType::Float ? (double)((const float *)host)[i]
               : t == ComponentType::Int ? (double)((const int *)host)[i]
                                          : (double)((const bool *)host)[i];
    check(smh_memcpy_htod(ctx, dev, tmp.data(), 8 * n), "tensor upload");
  }
  static void narrowDownload(smh_ctx *ctx, void *host, const double *dev, ComponentType t,
                             size_t n) {
    if (!n) return;
    if (t == ComponentType::Double) {
      check(smh_memcpy_dtoh(ctx, host, dev, 8 * n), "tensor download");
      return;
    }
    std::vector<double> tmp(n);
    check(smh_memcpy_dtoh(ctx, tmp.data(), dev, 8 * n), "tensor download");
    for (size_t i = 0; i < n; ++i) {
      if (t == ComponentType::Float) ((float *)host)[i] = (float)tmp[i];
      else if (t == ComponentType::Int) ((int *)host)[i] = (int)tmp[i];
      else ((bool *)host)[i] = tmp[i] != 0.0;
    }
  }
  // the bound matrix as an FP64 blocked-CSR operand of the ABI
  smh_bsr bsrOf(const std::string &name) {
    const MatrixSpec *ms = findMatrix(name);
    DevMatrix &d = matrices.at(name);
    return smh_bsr{d.nrows, (uint32_t)ms->bdim, d.nnz, d.rowptr, d.colidx, SMH_F64, d.vals,
                   nullptr};
  }
  double *devVector(const std::string &name) { return (double *)devVectors.at(name).dev; }
  const MatrixSpec *findMatrix(const std::string &n) const {
    for (auto &m : mspec)
      if (m.name == n) return &m;
    return nullptr;
  }
  const VectorSpec *findVector(const std::string &n) const {
    for (auto &v : vspec)
      if (v.name == n) return &v;
    return nullptr;
  }
  static void hipInfo(std::ostream &os);

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
    bool any = !fields.empty() || !owned.empty() || index || !endpoints.empty() ||
               !devVectors.empty();
    for (auto &kv : matrices) any |= kv.second.rowptr != nullptr;
    if (any) ctx = device_ctx();
    if (!ctx) return;
    smh_ctx_synchronize(ctx);
    for (auto &kv : fields) smh_free(ctx, kv.second.dev);
    for (auto &kv : endpoints) smh_free(ctx, kv.second);
    for (void *p : owned) smh_free(ctx, p);
    for (auto &kv : matrices) {
      DevMatrix &d = kv.second;
      if (d.rowptr) smh_bsr_forget(ctx, d.rowptr);
      smh_free(ctx, d.rowptr);
      smh_free(ctx, d.colidx);
      smh_free(ctx, d.vals);
      d.rowptr = d.colidx = nullptr;
      d.vals = nullptr;
    }
    for (auto &kv : devVectors) smh_free(ctx, kv.second.dev);
    if (index) smh_index_destroy(ctx, index);
    fields.clear();
    endpoints.clear();
    owned.clear();
    devVectors.clear();
    index = nullptr;
  }
  const ExternSpec *findExtern(const std::string &n) const {
    for (auto &e : spec)
      if (e.name == n) return &e;
    return nullptr;
  }

  std::vector<ExternSpec> spec;
  std::vector<MatrixSpec> mspec;
  std::vector<VectorSpec> vspec;
  std::string pipeline;  // the library calls of one run, for print()
  std::map<std::string, DevMatrix> matrices;
  std::map<std::string, void *> vectors;  // bound host pointers
  std::map<std::string, DevVector> devVectors;
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

void HipFunction::hipInfo(std::ostream &os) {
  char buf[512] = {0};
  if (smh_ctx_describe(device_ctx(), buf, sizeof(buf)) == SMH_OK) os << buf << "\n";
  Dl_info info{};
  if (dladdr((const void *)&smh_cg_solve, &info) && info.dli_fname)
    os << "code object: " << info.dli_fname << " (hand-written HIP kernels for gfx950)\n";
}

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
  if (program == "femTet") return {"initializeTet", "main"};
  return {"main"};
}
Function Program::compile(const std::string &function) {
  if (program.empty()) throw SimitException("no program loaded");
  const std::vector<std::string> fs = getFunctionNames();
  if (std::find(fs.begin(), fs.end(), function) == fs.end() && function != "timestep" &&
      function != program)
    throw SimitException("no function " + function + " in program " + program);
  return Function(makeFunction(program, function));
}

}  // namespace simit
