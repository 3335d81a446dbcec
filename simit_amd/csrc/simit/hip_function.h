// The HIP backend::Function base the programs of this backend derive from:
// binding (sets, dense externs, CSR TensorData), device mirrors of the bound
// fields, init / mapArgs / unmapArgs, and the solve and index helpers.
// Internal to the host library (program.cpp and the compiled-program runtime).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "program.h"
#include "simit_hip.h"

namespace simit {
namespace hipimpl {

// the process's context on device 0, created on first use
smh_ctx *device_ctx();
// throws SimitException(what + the context's last error) for rc != SMH_OK
void check(int rc, const char *what);

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

inline size_t devElemBytes(FieldType t) {
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


}  // namespace hipimpl
}  // namespace simit
