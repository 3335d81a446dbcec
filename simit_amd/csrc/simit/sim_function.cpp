// Runtime of a parsed Simit program's function on the HIP backend.
//
// The reference compiles a function to machine code at Program::compile and
// builds its indices at Function::init (function.cpp:22-139,
// llvm_function.cpp:248-465).  Here:
//   compile  checks the function and generates the HIP source of every map
//            it runs (codegen.cpp);
//   init     uploads the bound sets (HipFunction), compiles the source for
//            gfx950 with hiprtc (cached per process), builds the sparsity
//            pattern of each matrix type (the union of the edge sets that
//            assemble into it, every diagonal present) and, per map output,
//            the table of staged writes each entry sums;
//   run      interprets the function's statements on the host.  Every tensor
//            operation is a device kernel behind the C ABI: maps are the
//            generated kernels + smh_stage_reduce, A*x is smh_bsrmv in the
//            reference's row order, x'*y / dot / norm are smh_dot, and the
//            elementwise statements are smh_vec_op (a*x + b*y with one
//            rounding per operation, so `r - alpha*Ap` is one launch and the
//            same value as evaluating it step by step).  Scalars come back to
//            the host for control flow (while conditions).
#include "sim_function.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <iostream>
#include <mutex>
#include <set>
#include <sstream>

#include "codegen.h"
#include "hip_function.h"

namespace simit {
namespace hipimpl {
namespace {

using namespace fe;

struct DBuf {
  double *p = nullptr;
  size_t n = 0;
};

// sparsity of the [V,V] matrices of one vertex set
struct Pattern {
  std::string set;
  std::vector<std::string> edgeSets;
  smh_index *idx = nullptr;
  uint32_t *rowptr = nullptr, *colidx = nullptr;  // device
  uint32_t nrows = 0;
  uint64_t nnz = 0;
  std::vector<uint32_t> hrow, hcol;
  uint64_t loc(uint32_t r, uint32_t c) const {
    const uint32_t *b = hcol.data() + hrow[r], *e = hcol.data() + hrow[r + 1];
    const uint32_t *p = std::lower_bound(b, e, c);
    if (p == e || *p != c) throw SimitException("internal: block missing from the pattern");
    return (uint64_t)(p - hcol.data());
  }
};

struct Value {
  enum K { None, Scalar, Dense, Vec, Mat, Str, SetRef, Tuple } k = None;
  char t = 'f';  // Scalar / Dense: 'f', 'i', 'b'
  double s = 0.0;
  std::vector<double> d;  // Dense (row-major)
  std::vector<int> dims;
  bool row = false;
  std::string set;          // Vec / Mat rows, SetRef
  std::vector<int> block;   // Vec / Mat entry dims
  std::shared_ptr<DBuf> buf;
  double scale = 1.0;       // Vec / Mat: value = scale * buf (one pending product)
  bool scaled = false;
  bool transposed = false;  // Vec'
  const Pattern *pat = nullptr;
  bool stencil = false;  // Mat over a grid's points: 27 periodic slots per row
  uint32_t sd[3] = {1, 1, 1};
  std::vector<Value> tuple;
  std::string str;
  bool sys() const { return k == Vec || k == Mat; }
  int bs() const {
    int n = 1;
    for (int x : block) n *= x;
    return n;
  }
};

Value sval(double v, char t = 'f') {
  Value r;
  r.k = Value::Scalar;
  r.s = t == 'i' ? std::trunc(v) : t == 'b' ? (v != 0.0 ? 1.0 : 0.0) : v;
  r.t = t;
  return r;
}

[[noreturn]] void error(Loc l, const std::string &m) {
  throw SimitException("line " + std::to_string(l.line) + ": " + m);
}

// ---- static analysis of a function ---------------------------------------------------
void forEach(const Block &b, const std::function<void(const ExprP &)> &fe,
             const std::function<void(const StmtP &)> &fs) {
  std::function<void(const ExprP &)> ex = [&](const ExprP &e) {
    if (!e) return;
    fe(e);
    for (auto &a : e->args) ex(a);
  };
  for (auto &s : b) {
    fs(s);
    ex(s->init);
    ex(s->rhs);
    ex(s->lo);
    ex(s->hi);
    for (auto &x : s->lhs) ex(x);
    for (auto &x : s->conds) ex(x);
    for (auto &x : s->args) ex(x);
    for (auto &blk : s->blocks) forEach(blk, fe, fs);
    forEach(s->body, fe, fs);
  }
}

struct Analysis {
  const ProgramAST &P;
  std::vector<const Func *> hostFuncs;
  std::vector<MapRequest> reqs;
  std::map<const void *, size_t> mapIndex;  // Map expression / apply statement -> request
  std::set<std::pair<std::string, std::string>> hostFieldWrites;
  explicit Analysis(const ProgramAST &p) : P(p) {}
  void visit(const Func *f) {
    if (std::find(hostFuncs.begin(), hostFuncs.end(), f) != hostFuncs.end()) return;
    hostFuncs.push_back(f);
    std::vector<const Func *> calls;
    forEach(
        f->body,
        [&](const ExprP &e) {
          if (e->kind == Expr::Map) {
            mapIndex[e.get()] = reqs.size();
            reqs.push_back({e->name, e->target, e->through, e->reduce, (int)e->args.size()});
          }
          if (e->kind == Expr::Apply && e->args[0]->kind == Expr::Var) {
            const Func *g = P.func(e->args[0]->name);
            if (g) calls.push_back(g);
          }
        },
        [&](const StmtP &s) {
          if (s->kind == Stmt::Apply) {
            mapIndex[s.get()] = reqs.size();
            reqs.push_back({s->func, s->target, "", false, (int)s->args.size()});
          }
          if (s->kind == Stmt::Assign)
            for (auto &l : s->lhs)
              if (l->kind == Expr::Field && l->args[0]->kind == Expr::Var)
                hostFieldWrites.insert({l->args[0]->name, l->name});
        });
    for (const Func *g : calls) visit(g);
  }
};

FieldType fieldTypeOf(const Type &t) {
  if (t.comp == "int") return FieldType::Int32;
  if (t.comp == "bool") return FieldType::Bool;
  return FieldType::Float64;
}

// compiled code objects by generated source (one hiprtc run per program text)
std::mutex g_imageMutex;
std::map<std::string, std::vector<char>> g_images;

class SimFunction : public HipFunction {
 public:
  SimFunction(std::shared_ptr<const ProgramAST> prog, const Func *fn, std::vector<ExternSpec> spec,
              Analysis an, std::string source, std::vector<MapKernel> kernels,
              std::vector<std::string> denseExterns)
      : HipFunction(std::move(spec), {}, {}, "interpreted " + fn->name),
        prog(std::move(prog)),
        fn(fn),
        an(std::move(an)),
        source(std::move(source)),
        kernels(std::move(kernels)),
        denseExterns(std::move(denseExterns)) {
    for (auto &d : this->denseExterns) bindables.push_back({d, "scalar"});
  }
  ~SimFunction() override { freeMine(); }

  FuncType init() override {
    freeMine();
    return HipFunction::init();
  }

  void print(std::ostream &os) const override {
    HipFunction::print(os);
    os << "compiled " << fn->name << ": " << kernels.size() << " map kernel(s)\n";
    for (auto &k : kernels) {
      os << "  " << k.kernel << ": map " << k.req.func << " to " << k.req.target
         << (k.req.reduce ? " reduce +" : "") << " ->";
      for (auto &o : k.outputs)
        os << " " << o.name << (o.matrix ? "[" + o.rowSet + "," + o.rowSet + "]" : "[" + o.rowSet + "]")
           << " (" << o.ninst << " staged writes/element)";
      os << "\n";
    }
  }
  void printMachine(std::ostream &os) const override {
    hipInfo(os);
    os << "generated kernels (hiprtc, gfx950):";
    for (auto &k : kernels) os << " " << k.kernel;
    os << "\n";
  }

 protected:
  void setup(smh_ctx *ctx) override {
    // the element kernels
    std::vector<char> image;
    {
      std::lock_guard<std::mutex> lock(g_imageMutex);
      auto it = g_images.find(source);
      if (it == g_images.end()) {
        void *img = nullptr;
        size_t bytes = 0;
        std::vector<char> log(1 << 16);
        const int rc = smh_hiprtc_compile(source.c_str(), (fn->name + ".hip").c_str(), "gfx950",
                                          &img, &bytes, log.data(), log.size());
        if (rc != SMH_OK) throw SimitException("compiling the element kernels failed:\n" +
                                               std::string(log.data()));
        it = g_images.emplace(source, std::vector<char>((char *)img, (char *)img + bytes)).first;
        smh_image_free(img);
      }
      image = it->second;
    }
    check(smh_module_load(ctx, image.data(), image.size(), &module), "loading the element kernels");
    // the matrix patterns
    for (auto &k : kernels) {
      if (!k.gridDims) continue;
      Set *g = sets.at(k.req.through);
      if (g->getKind() != Set::Grid || (int)g->getDimensions().size() != k.gridDims)
        throw SimitException(k.req.through + " must be bound to a grid edge set of " +
                             std::to_string(k.gridDims) + " dimension(s)");
      size_t n = 1;
      for (int d : g->getDimensions()) n *= (size_t)d;
      if (n != (size_t)sets.at(k.req.target)->getSize() || g->getEndpointSet(0) != sets.at(k.req.target))
        throw SimitException(k.req.through + " is not the grid over the bound " + k.req.target);
    }
    for (auto &k : kernels)
      for (auto &o : k.outputs) {
        if (!o.matrix || o.stencil) continue;
        Pattern &p = patterns[o.rowSet];
        p.set = o.rowSet;
        if (k.card && std::find(p.edgeSets.begin(), p.edgeSets.end(), k.req.target) ==
                          p.edgeSets.end())
          p.edgeSets.push_back(k.req.target);
      }
    for (auto &kv : patterns) buildPattern(ctx, kv.second);
    // per map: staging buffers and the write tables of its outputs
    maps.resize(kernels.size());
    for (size_t m = 0; m < kernels.size(); ++m) buildMap(ctx, kernels[m], maps[m]);
  }

  void runDevice(smh_ctx *ctx) override {
    this->ctx = ctx;
    Frame fr;
    exec(fn->body, fr);
    check(smh_ctx_synchronize(ctx), "program run");
  }

 private:
  struct MapRt {
    uint64_t ne = 0;
    std::vector<uint64_t *> ptr;
    std::vector<uint32_t *> src;
    std::vector<double *> stage;
    std::vector<uint8_t *> taken;
    std::vector<uint64_t> nout;
  };
  struct Frame {
    std::map<std::string, Value> vars;
    std::map<std::string, std::string> generics;
  };

  std::shared_ptr<const ProgramAST> prog;
  const Func *fn;
  Analysis an;
  std::string source;
  std::vector<MapKernel> kernels;
  std::vector<std::string> denseExterns;
  smh_module *module = nullptr;
  std::map<std::string, Pattern> patterns;
  std::vector<MapRt> maps;
  std::vector<void *> mine;  // device allocations of this function
  std::multimap<size_t, double *> pool;
  std::map<std::string, Value> globals;
  smh_ctx *ctx = nullptr;

  void *dalloc(size_t bytes) {
    void *p = nullptr;
    check(smh_malloc(device_ctx(), bytes ? bytes : 8, &p), "device allocation");
    mine.push_back(p);
    return p;
  }
  void freeMine() {
    smh_ctx *c = nullptr;
    if (!mine.empty() || !pool.empty() || module || !patterns.empty()) c = device_ctx();
    if (!c) return;
    smh_ctx_synchronize(c);
    globals.clear();  // their buffers go back to the pool first
    for (auto &kv : patterns) {
      if (kv.second.idx) smh_index_destroy(c, kv.second.idx);
      else if (kv.second.rowptr) smh_bsr_forget(c, kv.second.rowptr);
    }
    patterns.clear();
    for (void *p : mine) smh_free(c, p);
    mine.clear();
    for (auto &kv : pool) smh_free(c, kv.second);
    pool.clear();
    if (module) smh_module_destroy(c, module);
    module = nullptr;
    maps.clear();
  }

  std::shared_ptr<DBuf> newBuf(size_t n) {
    double *p = nullptr;
    auto it = pool.find(n);
    if (it != pool.end()) {
      p = it->second;
      pool.erase(it);
    } else {
      void *q = nullptr;
      check(smh_malloc(device_ctx(), 8 * (n ? n : 1), &q), "device allocation");
      p = (double *)q;
    }
    DBuf *b = new DBuf{p, n};
    return std::shared_ptr<DBuf>(b, [this](DBuf *x) {
      pool.emplace(x->n, x->p);
      delete x;
    });
  }

  // ---- init helpers ----
  void buildPattern(smh_ctx *c, Pattern &p) {
    Set *v = sets.at(p.set);
    p.nrows = (uint32_t)v->getSize();
    if (p.edgeSets.size() > 4)
      throw SimitException("matrices over " + p.set + " assembled from more than 4 edge sets");
    if (!p.edgeSets.empty()) {
      std::vector<uint64_t> ne;
      std::vector<uint32_t> card;
      std::vector<const int32_t *> eps;
      for (auto &e : p.edgeSets) {
        Set *s = sets.at(e);
        ne.push_back((uint64_t)s->getSize());
        card.push_back((uint32_t)s->getCardinality());
        eps.push_back(endpoints.at(e));
      }
      check(smh_index_build(c, p.nrows, (uint32_t)ne.size(), ne.data(), card.data(), eps.data(),
                            SMH_INDEX_FORCE_DIAG, &p.idx),
            "index build");
      smh_index_view view{};
      check(smh_index_get(p.idx, &view), "index view");
      p.nnz = view.nnz;
      p.rowptr = const_cast<uint32_t *>(view.rowptr);
      p.colidx = const_cast<uint32_t *>(view.colidx);
      p.hrow.resize(p.nrows + 1);
      p.hcol.resize(p.nnz);
      check(smh_memcpy_dtoh(c, p.hrow.data(), p.rowptr, 4 * (size_t)(p.nrows + 1)), "index download");
      if (p.nnz) check(smh_memcpy_dtoh(c, p.hcol.data(), p.colidx, 4 * p.nnz), "index download");
    } else {  // only maps over the vertex set itself: the diagonal
      p.nnz = p.nrows;
      p.hrow.resize(p.nrows + 1);
      p.hcol.resize(p.nrows);
      for (uint32_t i = 0; i <= p.nrows; ++i) p.hrow[i] = i;
      for (uint32_t i = 0; i < p.nrows; ++i) p.hcol[i] = i;
      p.rowptr = (uint32_t *)dalloc(4 * (size_t)(p.nrows + 1));
      p.colidx = (uint32_t *)dalloc(4 * (size_t)p.nrows);
      check(smh_memcpy_htod(c, p.rowptr, p.hrow.data(), 4 * (size_t)(p.nrows + 1)), "upload");
      if (p.nrows) check(smh_memcpy_htod(c, p.colidx, p.hcol.data(), 4 * (size_t)p.nrows), "upload");
    }
  }

  // the grid dimensions of a map through a grid (padded with 1)
  std::vector<uint64_t> gridDims(const MapKernel &k) {
    std::vector<uint64_t> d(3, 1);
    if (k.gridDims) {
      const std::vector<int> &g = sets.at(k.req.through)->getDimensions();
      for (size_t i = 0; i < g.size() && i < 3; ++i) d[i] = (uint64_t)g[i];
    }
    return d;
  }
  static uint64_t gridNbr(uint64_t e, const int *o, const std::vector<uint64_t> &d) {
    int64_t c[3] = {(int64_t)(e % d[0]), (int64_t)((e / d[0]) % d[1]), (int64_t)(e / (d[0] * d[1]))};
    for (int k = 0; k < 3; ++k) c[k] = ((c[k] + o[k]) % (int64_t)d[k] + (int64_t)d[k]) % (int64_t)d[k];
    return (uint64_t)c[0] + d[0] * ((uint64_t)c[1] + d[1] * (uint64_t)c[2]);
  }

  void buildMap(smh_ctx *c, const MapKernel &k, MapRt &rt) {
    Set *tgt = sets.at(k.req.target);
    rt.ne = (uint64_t)tgt->getSize();
    const int32_t *ep = k.card ? tgt->getEndpointsData() : nullptr;
    const std::vector<uint64_t> gd = gridDims(k);
    for (size_t o = 0; o < k.outputs.size(); ++o) {
      const MapOutput &mo = k.outputs[o];
      const Pattern *pat = mo.matrix && !mo.stencil ? &patterns.at(mo.rowSet) : nullptr;
      const uint64_t nrow = (uint64_t)sets.at(mo.rowSet)->getSize();
      const uint64_t nout = mo.stencil ? 27 * nrow : mo.matrix ? pat->nnz : nrow;
      if (rt.ne * (uint64_t)mo.ninst >= 0xffffffffull)
        throw SimitException("map " + k.req.func + ": too many staged writes");
      auto target = [&](uint64_t e, const MapOutput::Write &w) -> uint64_t {
        if (mo.stencil) {  // slot-major [27][n]: slot of the column's offset, row e
          const int *d = w.off[1];
          const uint64_t slot = w.lj == -2 ? (uint64_t)((d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1))
                                           : 13;
          return slot * nrow + e;
        }
        const uint64_t r = w.li == -2   ? gridNbr(e, w.off[0], gd)
                           : w.li < 0 ? e
                                      : (uint64_t)ep[e * k.card + w.li];
        if (!mo.matrix) return r;
        const uint64_t col = w.lj == -2   ? gridNbr(e, w.off[1], gd)
                             : w.lj < 0 ? e
                                        : (uint64_t)ep[e * k.card + w.lj];
        return pat->loc((uint32_t)r, (uint32_t)col);
      };
      std::vector<uint64_t> ptr(nout + 1, 0);
      for (uint64_t e = 0; e < rt.ne; ++e)
        for (auto &w : mo.order) ++ptr[target(e, w) + 1];
      for (uint64_t i = 0; i < nout; ++i) ptr[i + 1] += ptr[i];
      std::vector<uint32_t> src(ptr[nout] ? ptr[nout] : 1);
      std::vector<uint64_t> fill(ptr.begin(), ptr.end() - 1);
      for (uint64_t e = 0; e < rt.ne; ++e)
        for (auto &w : mo.order) src[fill[target(e, w)]++] = (uint32_t)(e * mo.ninst + w.inst);
      uint64_t *dptr = (uint64_t *)dalloc(8 * (nout + 1));
      uint32_t *dsrc = (uint32_t *)dalloc(4 * src.size());
      check(smh_memcpy_htod(c, dptr, ptr.data(), 8 * (nout + 1)), "upload");
      check(smh_memcpy_htod(c, dsrc, src.data(), 4 * src.size()), "upload");
      rt.ptr.push_back(dptr);
      rt.src.push_back(dsrc);
      rt.nout.push_back(nout);
      const uint64_t slots = rt.ne * (uint64_t)mo.ninst;
      rt.stage.push_back((double *)dalloc(8 * slots * mo.bs));
      rt.taken.push_back(k.req.reduce ? nullptr : (uint8_t *)dalloc(slots));
    }
  }

  // ---- values ----
  Value materialize(Value v) {
    if (!v.sys() || !v.scaled) return v;
    auto b = newBuf(v.buf->n);
    check(smh_vec_op(ctx, SMH_VEC_LINCOMB, b->n, v.scale, v.buf->p, 0.0, nullptr, b->p),
          "vector statement");
    v.buf = b;
    v.scale = 1.0;
    v.scaled = false;
    return v;
  }
  Value sysLike(const Value &like) {
    Value r = like;
    r.buf = newBuf(like.buf->n);
    r.scale = 1.0;
    r.scaled = false;
    r.transposed = false;
    return r;
  }
  void sameSpace(const Value &a, const Value &b, Loc l) {
    if (a.k != b.k || a.set != b.set || a.bs() != b.bs() || a.pat != b.pat ||
        a.stencil != b.stencil || a.buf->n != b.buf->n)
      error(l, "operands are tensors of different types");
  }
  Value zeros(const Type &t, const Frame &fr, Loc l) {
    Value v;
    std::vector<int> dd;
    if (t.kind == Type::Scalar) return sval(0.0, t.comp == "int" ? 'i' : t.comp == "bool" ? 'b' : 'f');
    if (t.kind != Type::Tensor) error(l, "cannot make a value of type " + t.str());
    if (denseDims(t, &dd)) {
      v.k = Value::Dense;
      v.dims = dd;
      size_t n = 1;
      for (int x : dd) n *= (size_t)x;
      v.d.assign(n, 0.0);
      v.t = t.comp == "int" ? 'i' : t.comp == "bool" ? 'b' : 'f';
      v.row = t.transposed && dd.size() == 1;
      return v;
    }
    auto setOf = [&](const std::string &d) {
      auto g = fr.generics.find(d);
      const std::string s = g != fr.generics.end() ? g->second : d;
      if (!sets.count(s)) error(l, d + " is not a set");
      return s;
    };
    if (t.dims.size() > 2) error(l, "system tensors of order > 2 are not supported");
    v.set = setOf(t.dims[0]);
    if (t.block && !denseDims(*t.block, &v.block)) error(l, "blocks must be dense");
    if (t.dims.size() == 2) {
      if (setOf(t.dims[1]) != v.set) error(l, "matrices between different sets are not supported");
      auto it = patterns.find(v.set);
      if (it == patterns.end()) error(l, "no map assembles a matrix over " + v.set);
      v.k = Value::Mat;
      v.pat = &it->second;
      v.buf = newBuf(it->second.nnz * (size_t)v.bs());
    } else {
      v.k = Value::Vec;
      v.buf = newBuf(sets.at(v.set)->getSize() * (size_t)v.bs());
      v.transposed = t.transposed;
    }
    check(smh_vec_op(ctx, SMH_VEC_FILL, v.buf->n, 0.0, nullptr, 0.0, nullptr, v.buf->p), "fill");
    return v;
  }
  // a scalar (or a value of the same type) assigned to a declared type
  Value convert(const Value &v, const Type &t, const Frame &fr, Loc l) {
    if (t.kind == Type::Scalar) {
      if (v.k != Value::Scalar) error(l, "expected a scalar");
      return sval(v.s, t.comp == "int" ? 'i' : t.comp == "bool" ? 'b' : 'f');
    }
    Value z = zeros(t, fr, l);
    if (v.k == Value::Scalar) {
      if (z.k == Value::Dense) std::fill(z.d.begin(), z.d.end(), v.s);
      else check(smh_vec_op(ctx, SMH_VEC_FILL, z.buf->n, v.s, nullptr, 0.0, nullptr, z.buf->p), "fill");
      return z;
    }
    if (v.k != z.k) error(l, "value does not have the declared type " + t.str());
    return materialize(v);
  }

  // ---- fields ----
  DevField &fieldOf(const std::string &set, const std::string &f, Loc l) {
    auto it = fields.find(set + "." + f);
    if (it == fields.end()) error(l, set + " has no field " + f);
    return it->second;
  }
  Value fieldRead(const std::string &set, const std::string &f, Loc l) {
    DevField &d = fieldOf(set, f, l);
    if (d.host->type != FieldType::Float64 && d.host->type != FieldType::Float32)
      error(l, "only float fields are read as vectors");
    Value v;
    v.k = Value::Vec;
    v.set = set;
    const ElementDecl *ed = prog->element(prog->externType(set)->elem);
    for (auto &fd : ed->fields)
      if (fd.first == f) denseDims(*fd.second, &v.block);
    const size_t n = (size_t)sets.at(set)->getSize() * (size_t)d.host->ncomp;
    v.buf = newBuf(n);
    check(smh_vec_op(ctx, SMH_VEC_LINCOMB, n, 1.0, (const double *)d.dev, 0.0, nullptr, v.buf->p),
          "field read");
    return v;
  }
  void fieldWrite(const std::string &set, const std::string &f, Value v, Loc l) {
    DevField &d = fieldOf(set, f, l);
    if (d.host->type != FieldType::Float64 && d.host->type != FieldType::Float32)
      error(l, "only float fields are written");
    const size_t n = (size_t)sets.at(set)->getSize() * (size_t)d.host->ncomp;
    if (v.k == Value::Scalar) {
      check(smh_vec_op(ctx, SMH_VEC_FILL, n, v.s, nullptr, 0.0, nullptr, (double *)d.dev), "field write");
      return;
    }
    if (v.k != Value::Vec || v.set != set || v.buf->n != n)
      error(l, "value written to " + set + "." + f + " is not a vector over " + set +
                   " of the field's shape");
    check(smh_vec_op(ctx, SMH_VEC_LINCOMB, n, v.scaled ? v.scale : 1.0, v.buf->p, 0.0, nullptr,
                     (double *)d.dev),
          "field write");
  }

  // ---- maps ----
  Value runMap(size_t m, const std::vector<ExprP> &extra, Frame &fr, Loc l,
               const std::string *initSet = nullptr, const std::string *initField = nullptr) {
    const MapKernel &k = kernels[m];
    MapRt &rt = maps[m];
    std::vector<double> xs;
    for (auto &x : extra) {
      Value v = eval(x, fr);
      if (v.k != Value::Scalar) error(l, "map arguments must be scalars");
      xs.push_back(v.s);
    }
    for (size_t o = 0; o < k.outputs.size(); ++o) {
      const uint64_t slots = rt.ne * (uint64_t)k.outputs[o].ninst;
      if (!slots) continue;
      check(smh_vec_op(ctx, SMH_VEC_FILL, slots * k.outputs[o].bs, 0.0, nullptr, 0.0, nullptr,
                       rt.stage[o]),
            "stage clear");
      if (rt.taken[o]) check(smh_memset(ctx, rt.taken[o], 0, slots), "stage clear");
    }
    std::vector<uint64_t> args(k.args.size(), 0);
    for (size_t s = 0; s < k.args.size(); ++s) {
      const KernelArg &a = k.args[s];
      switch (a.kind) {
        case KernelArg::Count: args[s] = rt.ne; break;
        case KernelArg::Endpoints:
          args[s] = k.card ? (uint64_t)(uintptr_t)endpoints.at(k.req.target) : 0;
          break;
        case KernelArg::Stage: args[s] = (uint64_t)(uintptr_t)rt.stage[a.index]; break;
        case KernelArg::Taken: args[s] = (uint64_t)(uintptr_t)rt.taken[a.index]; break;
        case KernelArg::Field:
          args[s] = (uint64_t)(uintptr_t)fieldOf(a.set, a.field, l).dev;
          break;
        case KernelArg::Scalar: std::memcpy(&args[s], &xs.at(a.index), 8); break;
        case KernelArg::GridDim: args[s] = gridDims(k)[a.index]; break;
      }
    }
    check(smh_module_launch(ctx, module, k.kernel.c_str(), rt.ne, args.data(), 8 * args.size()),
          "map launch");
    std::vector<Value> outs;
    for (size_t o = 0; o < k.outputs.size(); ++o) {
      const MapOutput &mo = k.outputs[o];
      Value v;
      v.k = mo.matrix ? Value::Mat : Value::Vec;
      v.set = mo.rowSet;
      v.block = mo.blockDims;
      v.stencil = mo.stencil;
      v.pat = mo.matrix && !mo.stencil ? &patterns.at(mo.rowSet) : nullptr;
      if (mo.stencil) {
        const std::vector<uint64_t> gd = gridDims(k);
        for (int i = 0; i < 3; ++i) v.sd[i] = (uint32_t)gd[i];
      }
      v.buf = newBuf(rt.nout[o] * (size_t)mo.bs);
      if (!k.req.reduce) {  // plain writes: unwritten entries keep the target's values
        if (initSet && k.outputs.size() == 1 && !mo.matrix) {
          DevField &d = fieldOf(*initSet, *initField, l);
          check(smh_vec_op(ctx, SMH_VEC_LINCOMB, v.buf->n, 1.0, (const double *)d.dev, 0.0,
                           nullptr, v.buf->p),
                "map output");
        } else {
          check(smh_vec_op(ctx, SMH_VEC_FILL, v.buf->n, 0.0, nullptr, 0.0, nullptr, v.buf->p),
                "map output");
        }
      }
      check(smh_stage_reduce(ctx, rt.nout[o], (uint32_t)mo.bs, rt.ptr[o], rt.src[o], rt.stage[o],
                             rt.taken[o], v.buf->p),
            "map reduction");
      outs.push_back(v);
    }
    if (outs.size() == 1) return outs[0];
    Value t;
    t.k = Value::Tuple;
    t.tuple = outs;
    return t;
  }

  // ---- expressions ----
  Value lookupVar(const std::string &n, Frame &fr, Loc l) {
    auto it = fr.vars.find(n);
    if (it != fr.vars.end()) return it->second;
    auto g = globals.find(n);
    if (g != globals.end()) return g->second;
    for (auto &c : prog->consts)
      if (c->name == n) {
        Frame empty;
        Value v = c->type ? convert(eval(c->init, empty), *c->type, empty, c->loc)
                          : materialize(eval(c->init, empty));
        globals[n] = v;
        return v;
      }
    if (sets.count(n)) {
      Value v;
      v.k = Value::SetRef;
      v.set = n;
      return v;
    }
    if (TypeP t = prog->externType(n)) {
      auto sc = scalars.find(n);
      if (sc == scalars.end()) error(l, "extern " + n + " is not bound");
      const double *p = (const double *)sc->second;
      std::vector<int> dd;
      if (t->kind == Type::Scalar) return sval(p[0], t->comp == "int" ? 'i' : 'f');
      if (!denseDims(*t, &dd)) error(l, "extern " + n + " has an unsupported type");
      Value v;
      v.k = Value::Dense;
      v.dims = dd;
      size_t cnt = 1;
      for (int x : dd) cnt *= (size_t)x;
      v.d.assign(p, p + cnt);
      return v;
    }
    error(l, "unknown name " + n);
  }

  Value scalarOp(const std::string &op, const Value &a, const Value &b, Loc l) {
    const bool ints = a.t == 'i' && b.t == 'i';
    const double x = a.s, y = b.s;
    if (op == "+") return sval(x + y, ints ? 'i' : 'f');
    if (op == "-") return sval(x - y, ints ? 'i' : 'f');
    if (op == "*") return sval(x * y, ints ? 'i' : 'f');
    if (op == "/" || op == "./") {
      if (ints) {
        if (y == 0.0) error(l, "integer division by zero");
        return sval((double)((long long)x / (long long)y), 'i');
      }
      return sval(x / y);
    }
    if (op == ".*") return sval(x * y, ints ? 'i' : 'f');
    if (op == "^") return sval(std::pow(x, y));
    if (op == "==") return sval(x == y, 'b');
    if (op == "!=") return sval(x != y, 'b');
    if (op == "<") return sval(x < y, 'b');
    if (op == ">") return sval(x > y, 'b');
    if (op == "<=") return sval(x <= y, 'b');
    if (op == ">=") return sval(x >= y, 'b');
    if (op == "and") return sval(x != 0.0 && y != 0.0, 'b');
    if (op == "or") return sval(x != 0.0 || y != 0.0, 'b');
    if (op == "xor") return sval((x != 0.0) != (y != 0.0), 'b');
    error(l, "unsupported operator " + op);
  }

  Value denseOp(const std::string &op, const Value &a, const Value &b, Loc l) {
    Value r = a.k == Value::Dense ? a : b;
    const bool as = a.k == Value::Scalar, bs = b.k == Value::Scalar;
    if (!as && !bs && a.d.size() != b.d.size()) error(l, "operands have different dimensions");
    if (op != "+" && op != "-" && op != "*" && op != "/" && op != ".*" && op != "./")
      error(l, "operator " + op + " on dense tensors");
    if (op == "*" && !as && !bs) error(l, "dense tensor products at the host level are not supported");
    for (size_t i = 0; i < r.d.size(); ++i) {
      const double x = as ? a.s : a.d[i], y = bs ? b.s : b.d[i];
      r.d[i] = op == "+" ? x + y : op == "-" ? x - y : (op == "*" || op == ".*") ? x * y : x / y;
    }
    r.t = 'f';
    return r;
  }

  Value sysOp(const std::string &op, Value a, Value b, Loc l) {
    const bool as = a.k == Value::Scalar, bs = b.k == Value::Scalar;
    if (op == "*") {
      if (as || bs) {
        Value s = as ? a : b, v = as ? b : a;
        if (!v.sys()) error(l, "product of a scalar and a non-tensor");
        if (v.scaled) v = materialize(v);
        v.scale = s.s;
        v.scaled = true;
        return v;
      }
      if (a.k == Value::Mat && b.k == Value::Vec) {
        if (b.transposed) error(l, "matrix times a row vector");
        a = materialize(a);
        b = materialize(b);
        if (a.set != b.set || a.bs() != b.bs() * b.bs() || (b.bs() != 1 && b.bs() != 3))
          error(l, "matrix and vector do not match (blocks of 1 or 3 components)");
        Value y = sysLike(b);
        if (a.stencil) {
          if (b.bs() != 1) error(l, "a grid matrix multiplies a vector of scalars");
          smh_stencil27 S{a.sd[0], a.sd[1], a.sd[2], SMH_F64, a.buf->p, nullptr,
                          SMH_STENCIL_PERIODIC};
          check(smh_stencil27_mv(ctx, &S, b.buf->p, y.buf->p), "A*x");
          return y;
        }
        smh_bsr A{a.pat->nrows, (uint32_t)(b.bs() == 1 ? 1 : 3), a.pat->nnz, a.pat->rowptr,
                  a.pat->colidx, SMH_F64, a.buf->p, nullptr};
        check(smh_bsrmv(ctx, &A, b.buf->p, y.buf->p, 1), "A*x");
        return y;
      }
      if (a.k == Value::Vec && b.k == Value::Vec && a.transposed && !b.transposed) {
        sameSpace(a, b, l);
        a = materialize(a);
        b = materialize(b);
        double r = 0.0;
        check(smh_dot(ctx, SMH_F64, a.buf->n, a.buf->p, b.buf->p, &r), "x'*y");
        return sval(r);
      }
      error(l, "unsupported product of system tensors");
    }
    if (op == "+" || op == "-") {
      if (as || bs) {
        Value v = materialize(as ? b : a), s = as ? a : b;
        if (v.k != Value::Vec) error(l, "scalar added to a matrix");
        Value r = sysLike(v);
        if (as && op == "-") {  // s - v = (-1 * v) + s
          check(smh_vec_op(ctx, SMH_VEC_LINCOMB, v.buf->n, -1.0, v.buf->p, 0.0, nullptr, r.buf->p), "vector statement");
          check(smh_vec_op(ctx, SMH_VEC_ADDS, v.buf->n, s.s, r.buf->p, 0.0, nullptr, r.buf->p), "vector statement");
        } else {
          check(smh_vec_op(ctx, SMH_VEC_ADDS, v.buf->n, op == "-" ? -s.s : s.s, v.buf->p, 0.0,
                           nullptr, r.buf->p),
                "vector statement");
        }
        return r;
      }
      sameSpace(a, b, l);
      if (a.k == Value::Vec && a.transposed != b.transposed) error(l, "row and column vectors added");
      Value r = sysLike(a);
      r.transposed = a.transposed;
      const double sa = a.scaled ? a.scale : 1.0, sb0 = b.scaled ? b.scale : 1.0;
      const double sb = op == "-" ? -sb0 : sb0;  // x - s*y = x + (-s)*y exactly
      check(smh_vec_op(ctx, SMH_VEC_LINCOMB, a.buf->n, sa, a.buf->p, sb, b.buf->p, r.buf->p),
            "vector statement");
      return r;
    }
    if (op == "/" && bs) {
      a = materialize(a);
      Value r = sysLike(a);
      r.transposed = a.transposed;
      check(smh_vec_op(ctx, SMH_VEC_DIVS, a.buf->n, b.s, a.buf->p, 0.0, nullptr, r.buf->p),
            "vector statement");
      return r;
    }
    if ((op == ".*" || op == "./") && !as && !bs) {
      sameSpace(a, b, l);
      a = materialize(a);
      b = materialize(b);
      Value r = sysLike(a);
      check(smh_vec_op(ctx, op == ".*" ? SMH_VEC_EMUL : SMH_VEC_EDIV, a.buf->n, 0.0, a.buf->p, 0.0,
                       b.buf->p, r.buf->p),
            "vector statement");
      return r;
    }
    error(l, "unsupported operation " + op + " on system tensors");
  }

  Value eval(const ExprP &e, Frame &fr) {
    switch (e->kind) {
      case Expr::Num: return sval(e->num, e->isInt ? 'i' : 'f');
      case Expr::Bool: return sval(e->num, 'b');
      case Expr::Str: {
        Value v;
        v.k = Value::Str;
        v.str = e->name;
        return v;
      }
      case Expr::Var: return lookupVar(e->name, fr, e->loc);
      case Expr::Literal: {
        Value v;
        v.k = Value::Dense;
        v.d = e->lit;
        v.dims = e->litDims;
        v.t = e->litInt ? 'i' : 'f';
        v.row = e->litDims.size() == 1;
        return v;
      }
      case Expr::Unary: {
        Value a = eval(e->args[0], fr);
        if (e->op == "+") return a;
        if (e->op == "not") {
          if (a.k != Value::Scalar) error(e->loc, "not of a tensor");
          return sval(a.s == 0.0, 'b');
        }
        if (a.k == Value::Scalar) return sval(-a.s, a.t);
        if (a.k == Value::Dense) {
          for (auto &x : a.d) x = -x;
          return a;
        }
        if (a.sys()) {  // -(s*x) = (-s)*x exactly
          a.scale = a.scaled ? -a.scale : -1.0;
          a.scaled = true;
          return a;
        }
        error(e->loc, "negation of a non-numeric value");
      }
      case Expr::Transpose: {
        Value a = eval(e->args[0], fr);
        if (a.k == Value::Scalar) return a;
        if (a.k == Value::Vec) {
          a.transposed = !a.transposed;
          return a;
        }
        if (a.k == Value::Dense && a.dims.size() == 1) {
          a.row = !a.row;
          return a;
        }
        if (a.k == Value::Dense && a.dims.size() == 2) {
          Value r = a;
          std::swap(r.dims[0], r.dims[1]);
          for (int i = 0; i < a.dims[0]; ++i)
            for (int j = 0; j < a.dims[1]; ++j) r.d[(size_t)j * a.dims[0] + i] = a.d[(size_t)i * a.dims[1] + j];
          return r;
        }
        error(e->loc, "transpose of a system matrix is not supported");
      }
      case Expr::Binary: {
        Value a = eval(e->args[0], fr), b = eval(e->args[1], fr);
        if (a.sys() || b.sys()) return sysOp(e->op, a, b, e->loc);
        if (a.k == Value::Dense || b.k == Value::Dense) return denseOp(e->op, a, b, e->loc);
        if (a.k != Value::Scalar || b.k != Value::Scalar) error(e->loc, "operands are not numbers");
        return scalarOp(e->op, a, b, e->loc);
      }
      case Expr::Field: {
        const ExprP &b = e->args[0];
        if (b->kind == Expr::Var) {
          Value base = lookupVar(b->name, fr, b->loc);
          if (base.k == Value::SetRef) return fieldRead(base.set, e->name, e->loc);
        }
        error(e->loc, "fields are read from sets");
      }
      case Expr::Map: {
        auto it = an.mapIndex.find(e.get());
        return runMap(it->second, e->args, fr, e->loc);
      }
      case Expr::Apply: return apply(e, fr);
    }
    error(e->loc, "unsupported expression");
  }

  Value apply(const ExprP &e, Frame &fr) {
    const ExprP &base = e->args[0];
    std::vector<ExprP> args(e->args.begin() + 1, e->args.end());
    if (base->kind == Expr::Var && !fr.vars.count(base->name)) {
      const std::string &n = base->name;
      if (const Func *f = prog->func(n)) {
        std::vector<Value> rs = call(*f, args, fr, e->loc);
        if (rs.size() != 1) error(e->loc, n + " does not return one value");
        return rs[0];
      }
      std::vector<Value> a;
      for (auto &x : args) {
        if (!x) error(e->loc, "slices are not arguments");
        a.push_back(eval(x, fr));
      }
      if (n == "dot" || n == "norm") {
        if (a.size() != (n == "dot" ? 2u : 1u)) error(e->loc, n + ": wrong number of arguments");
        const Value &x = a[0], &y = a.back();
        if (x.k == Value::Vec && y.k == Value::Vec) {
          sameSpace(x, y, e->loc);
          const Value xm = materialize(x), ym = materialize(y);
          double r = 0.0;
          check(smh_dot(ctx, SMH_F64, xm.buf->n, xm.buf->p, ym.buf->p, &r), n.c_str());
          return sval(n == "norm" ? std::sqrt(r) : r);
        }
        if (x.k == Value::Dense && y.k == Value::Dense && x.d.size() == y.d.size()) {
          double r = 0.0;
          for (size_t i = 0; i < x.d.size(); ++i) r = r + x.d[i] * y.d[i];
          return sval(n == "norm" ? std::sqrt(r) : r);
        }
        error(e->loc, n + " of unsupported operands");
      }
      static const std::map<std::string, double (*)(double)> unary = {
          {"sqrt", std::sqrt}, {"log", std::log}, {"exp", std::exp}, {"sin", std::sin},
          {"cos", std::cos},   {"tan", std::tan}, {"asin", std::asin}, {"acos", std::acos},
          {"atan", std::atan}, {"abs", std::fabs}};
      auto u = unary.find(n);
      if (u != unary.end()) {
        if (a.size() != 1 || a[0].k != Value::Scalar) error(e->loc, n + " takes a scalar");
        return sval(u->second(a[0].s));
      }
      if ((n == "pow" || n == "atan2") && a.size() == 2 && a[0].k == Value::Scalar &&
          a[1].k == Value::Scalar)
        return sval(n == "pow" ? std::pow(a[0].s, a[1].s) : std::atan2(a[0].s, a[1].s));
      error(e->loc, "unknown function " + n);
    }
    // a read of a dense tensor
    Value t = eval(base, fr);
    if (t.k != Value::Dense) error(e->loc, "elements of system tensors are not read on the host");
    if (args.size() != t.dims.size()) error(e->loc, "wrong number of indices");
    size_t off = 0;
    for (size_t k = 0; k < args.size(); ++k) {
      if (!args[k]) error(e->loc, "slices of host tensors are not supported");
      Value i = eval(args[k], fr);
      if (i.k != Value::Scalar || i.s < 0 || i.s >= t.dims[k]) error(e->loc, "index out of range");
      off = off * (size_t)t.dims[k] + (size_t)i.s;
    }
    return sval(t.d[off], t.t);
  }

  std::vector<Value> call(const Func &f, const std::vector<ExprP> &args, Frame &caller, Loc l) {
    if (args.size() != f.params.size())
      error(l, f.name + " takes " + std::to_string(f.params.size()) + " arguments");
    Frame fr;
    std::vector<Value> vals;
    for (size_t k = 0; k < args.size(); ++k) {
      if (!args[k]) error(l, "slices are not arguments");
      Value v = materialize(eval(args[k], caller));
      const Type &t = *f.params[k].type;
      if (t.kind == Type::Tensor && v.sys()) {  // bind generic set parameters
        for (size_t d = 0; d < t.dims.size(); ++d)
          if (std::find(f.generics.begin(), f.generics.end(), t.dims[d]) != f.generics.end()) {
            auto g = fr.generics.find(t.dims[d]);
            if (g != fr.generics.end() && g->second != v.set)
              error(l, "generic " + t.dims[d] + " bound to two sets");
            fr.generics[t.dims[d]] = v.set;
          }
      }
      fr.vars[f.params[k].name] = v;
      vals.push_back(v);
    }
    exec(f.body, fr);
    for (size_t k = 0; k < args.size(); ++k)
      if (f.params[k].inout) assignTo(args[k], fr.vars[f.params[k].name], caller, l);
    std::vector<Value> rs;
    for (auto &r : f.results) {
      auto it = fr.vars.find(r.name);
      if (it == fr.vars.end()) rs.push_back(zeros(*r.type, fr, l));
      else rs.push_back(it->second);
    }
    return rs;
  }

  // ---- statements ----
  void assignTo(const ExprP &lhs, const Value &v, Frame &fr, Loc l) {
    if (lhs->kind == Expr::Var) {
      if (sets.count(lhs->name)) error(l, "a set cannot be assigned");
      fr.vars[lhs->name] = materialize(v);
      return;
    }
    if (lhs->kind == Expr::Field && lhs->args[0]->kind == Expr::Var) {
      Value base = lookupVar(lhs->args[0]->name, fr, l);
      if (base.k != Value::SetRef) error(l, "fields of sets are assigned");
      fieldWrite(base.set, lhs->name, v, l);
      return;
    }
    if (lhs->kind == Expr::Apply && lhs->args[0]->kind == Expr::Var) {
      auto it = fr.vars.find(lhs->args[0]->name);
      if (it == fr.vars.end() || it->second.k != Value::Dense)
        error(l, "only elements of host tensors can be assigned");
      Value &t = it->second;
      if (v.k != Value::Scalar) error(l, "assigning a tensor to one element");
      size_t off = 0;
      if (lhs->args.size() - 1 != t.dims.size()) error(l, "wrong number of indices");
      for (size_t k = 1; k < lhs->args.size(); ++k) {
        if (!lhs->args[k]) error(l, "slices are not assigned on the host");
        Value i = eval(lhs->args[k], fr);
        if (i.s < 0 || i.s >= t.dims[k - 1]) error(l, "index out of range");
        off = off * (size_t)t.dims[k - 1] + (size_t)i.s;
      }
      t.d[off] = v.s;
      return;
    }
    error(l, "unsupported assignment target");
  }

  bool truth(const Value &v, Loc l) {
    if (v.k != Value::Scalar) error(l, "condition is not a scalar");
    return v.s != 0.0;
  }

  void exec(const Block &b, Frame &fr) {
    for (auto &s : b) execStmt(s, fr);
  }

  void execStmt(const StmtP &s, Frame &fr) {
    switch (s->kind) {
      case Stmt::VarDecl:
      case Stmt::ConstDecl: {
        Value v;
        if (s->init) v = eval(s->init, fr);
        if (s->type) v = s->init ? convert(v, *s->type, fr, s->loc) : zeros(*s->type, fr, s->loc);
        fr.vars[s->name] = materialize(v);
        return;
      }
      case Stmt::Assign: {
        if (s->lhs.size() == 1) {
          const ExprP &lhs = s->lhs[0];
          // a map without a reduction written straight into a field keeps
          // the field's values where no element writes
          if (s->rhs->kind == Expr::Map && !s->rhs->reduce && lhs->kind == Expr::Field &&
              lhs->args[0]->kind == Expr::Var && sets.count(lhs->args[0]->name)) {
            const std::string set = lhs->args[0]->name;
            Value v = runMap(an.mapIndex.at(s->rhs.get()), s->rhs->args, fr, s->rhs->loc, &set,
                             &lhs->name);
            return assignTo(lhs, v, fr, s->loc);
          }
          return assignTo(lhs, eval(s->rhs, fr), fr, s->loc);
        }
        Value v = eval(s->rhs, fr);
        if (v.k != Value::Tuple || v.tuple.size() != s->lhs.size())
          error(s->loc, "the right-hand side does not have " + std::to_string(s->lhs.size()) +
                            " results");
        for (size_t k = 0; k < s->lhs.size(); ++k) assignTo(s->lhs[k], v.tuple[k], fr, s->loc);
        return;
      }
      case Stmt::ExprStmt: {
        if (s->rhs->kind == Expr::Apply && s->rhs->args[0]->kind == Expr::Var) {
          if (const Func *f = prog->func(s->rhs->args[0]->name)) {
            std::vector<ExprP> args(s->rhs->args.begin() + 1, s->rhs->args.end());
            call(*f, args, fr, s->loc);
            return;
          }
        }
        eval(s->rhs, fr);
        return;
      }
      case Stmt::If: {
        for (size_t k = 0; k < s->conds.size(); ++k)
          if (truth(eval(s->conds[k], fr), s->conds[k]->loc)) return exec(s->blocks[k], fr);
        if (s->blocks.size() > s->conds.size()) exec(s->blocks.back(), fr);
        return;
      }
      case Stmt::While:
        while (truth(eval(s->rhs, fr), s->rhs->loc)) exec(s->body, fr);
        return;
      case Stmt::For: {
        if (!s->overSet.empty()) error(s->loc, "loops over sets are not supported");
        const Value lo = eval(s->lo, fr), hi = eval(s->hi, fr);
        if (lo.k != Value::Scalar || hi.k != Value::Scalar) error(s->loc, "loop bounds are scalars");
        for (long long i = (long long)lo.s; i < (long long)hi.s; ++i) {
          fr.vars[s->name] = sval((double)i, 'i');
          exec(s->body, fr);
        }
        return;
      }
      case Stmt::Print: {
        for (auto &a : s->args) {
          Value v = materialize(eval(a, fr));
          if (v.k == Value::Str) std::cout << v.str;
          else if (v.k == Value::Scalar) std::cout << v.s;
          else if (v.k == Value::Dense) {
            for (size_t i = 0; i < v.d.size(); ++i) std::cout << (i ? " " : "") << v.d[i];
          } else if (v.k == Value::Vec) {
            std::vector<double> h(v.buf->n);
            check(smh_memcpy_dtoh(ctx, h.data(), v.buf->p, 8 * h.size()), "print");
            for (size_t i = 0; i < h.size(); ++i) std::cout << (i ? " " : "") << h[i];
          } else {
            error(s->loc, "cannot print this value");
          }
        }
        if (s->newline) std::cout << "\n";
        std::cout.flush();
        return;
      }
      case Stmt::Apply: {
        runMap(an.mapIndex.at(s.get()), s->args, fr, s->loc);
        return;
      }
    }
  }
};

}  // namespace

std::shared_ptr<backend::Function> makeSimFunction(std::shared_ptr<const ProgramAST> prog,
                                                   const std::string &function) {
  const Func *f = prog->func(function);
  if (!f) throw SimitException("Attempting to compile an unknown function (" + function + ")");
  if (!f->params.empty() || !f->results.empty())
    throw SimitException("function " + function +
                         " takes arguments or returns results: bind externs and compile a "
                         "function without them");
  Analysis an(*prog);
  an.visit(f);
  std::vector<MapKernel> kernels;
  std::string source;
  try {
    source = generateModule(*prog, an.reqs, &kernels);
  } catch (const CodegenError &e) {
    throw SimitException("line " + std::to_string(e.loc.line) + ": " + e.what());
  }
  std::set<std::pair<std::string, std::string>> written = an.hostFieldWrites;
  for (auto &k : kernels)
    for (auto &w : k.fieldsWritten) written.insert(w);
  std::vector<ExternSpec> spec;
  std::vector<std::string> dense;
  for (auto &ex : prog->externs) {
    const Type &t = *ex.second;
    if (t.kind != Type::Set) {
      if (t.kind == Type::Scalar || t.kind == Type::Tensor) {
        std::vector<int> dd;
        if (!denseDims(t, &dd))
          throw SimitException("extern " + ex.first + ": sparse extern tensors are not supported "
                               "by compiled programs");
        dense.push_back(ex.first);
        continue;
      }
      throw SimitException("extern " + ex.first + " has an unsupported type");
    }
    const ElementDecl *ed = prog->element(t.elem);
    if (!ed) throw SimitException("unknown element type " + t.elem);
    ExternSpec es;
    es.name = ex.first;
    es.cardinality = (int)t.endpoints.size();
    for (auto &fd : ed->fields) {
      std::vector<int> dd;
      if (!denseDims(*fd.second, &dd))
        throw SimitException("field " + t.elem + "." + fd.first + " is not a dense tensor");
      int n = 1;
      for (int x : dd) n *= x;
      es.fields.push_back({fd.first, fieldTypeOf(*fd.second), n,
                           written.count({ex.first, fd.first}) > 0});
    }
    spec.push_back(es);
  }
  return std::make_shared<SimFunction>(prog, f, spec, std::move(an), source, kernels, dense);
}

std::string generatedSource(const ProgramAST &prog, const std::string &function) {
  const Func *f = prog.func(function);
  if (!f) throw SimitException("unknown function " + function);
  Analysis an(prog);
  an.visit(f);
  std::vector<MapKernel> kernels;
  try {
    return generateModule(prog, an.reqs, &kernels);
  } catch (const CodegenError &e) {
    throw SimitException("line " + std::to_string(e.loc.line) + ": " + e.what());
  }
}

}  // namespace hipimpl
}  // namespace simit
