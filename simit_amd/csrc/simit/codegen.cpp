// HIP C++ generation for the element functions of a Simit program's maps
// (see codegen.h).  Expressions over small dense tensors become explicit
// loops in row-major order with one rounding per operation (the module is
// compiled with -ffp-contract=off), products sum from 0 in ascending index
// order; det/inv/log are the library's own (cofactor det/inv as in
// misc/gpu/linalg.c, log by an atanh series) so that generated and
// hand-written element code agree bit for bit.
#include "codegen.h"

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <set>
#include <sstream>

namespace simit {
namespace fe {

namespace {

struct Shape {
  char t = 'f';  // 'f' double, 'i' int, 'b' bool
  std::vector<int> d;
  bool row = false;  // a row vector (1-D only)
  bool scalar() const { return d.empty(); }
  int size() const {
    int n = 1;
    for (int x : d) n *= x;
    return n;
  }
};

using Idx = std::vector<std::string>;

struct CVal {
  Shape sh;
  std::string s;                               // scalar value (C expression)
  std::function<std::string(const Idx &)> at;  // array element
  std::string arr;  // a row-major C array holding the value, if there is one
  // the element / an endpoint or grid neighbour / a grid link (s = its index)
  enum Ref { None, Elem, Endpoint, Link } ref = None;
};

std::string ctype(char t) { return t == 'i' ? "int" : t == 'b' ? "bool" : "double"; }

std::string num(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  std::string s = buf;
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

std::string flat(const Idx &idx, const std::vector<int> &dims) {
  if (idx.empty()) return "0";
  std::string e = "(" + idx[0] + ")";
  for (size_t k = 1; k < idx.size(); ++k)
    e = "(" + e + "*" + std::to_string(dims[k]) + "+(" + idx[k] + "))";
  return e;
}

struct Sym {
  enum K { Local, Param, ArrParam, Result, ElemParam, TupleParam, GridParam, Output, Loop } k = Local;
  std::string c;
  Shape sh;
  int out = -1;
  bool inout = false;
};

struct LoopInfo {
  std::string name, c;
  bool constant = false;
  long lo = 0, trip = 0;
};

struct Site {
  int out = 0, base = 0;
  std::vector<LoopInfo> loops;
  std::vector<ExprP> idx;
};

class Gen {
 public:
  explicit Gen(const ProgramAST &p) : P(p) {}

  std::string module(const std::vector<MapRequest> &reqs, std::vector<MapKernel> *ks) {
    std::ostringstream kern;
    ks->clear();
    for (size_t r = 0; r < reqs.size(); ++r) {
      MapKernel mk;
      mk.req = reqs[r];
      mk.kernel = "simit_map_" + std::to_string(r) + "_" + reqs[r].func;
      kern << mapKernel(mk);
      ks->push_back(mk);
    }
    return prelude() + dev.str() + kern.str();
  }

 private:
  const ProgramAST &P;
  std::ostringstream dev;
  std::set<std::string> devDone;
  int tmp = 0;

  struct Fn {
    std::map<std::string, Sym> syms;
    std::vector<std::pair<std::string, Shape>> decls;
    std::ostringstream body;
    int indent = 1;
    MapKernel *mk = nullptr;
    std::string elemSet, epSet, elemType, epType;
    std::string gridSet, gridType;  // a map through a grid: its edge set
    std::string dims[3];            // ... and its dimension arguments
    std::vector<LoopInfo> loops;
    int whileDepth = 0;
    std::map<const Stmt *, Site> sites;
    std::vector<int> nextInst;
    std::map<std::string, int> fieldSlot;  // "set.field" -> argument slot
  };
  Fn *F = nullptr;

  [[noreturn]] void fail(Loc l, const std::string &m) { throw CodegenError(l, m); }

  static std::string prelude() {
    return R"(// generated by the simit_amd front-end: element kernels of one program
#define SIMIT_DEV __device__ __forceinline__
// an index the generator could not bound at compile time, kept in range
SIMIT_DEV int simit_clamp(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
// the point at offset (o0, o1, o2) from point e of an n0 x n1 x n2 periodic grid
// (first coordinate fastest)
SIMIT_DEV unsigned long long simit_nbr(unsigned long long e, long long o0, long long o1,
                                       long long o2, unsigned long long n0,
                                       unsigned long long n1, unsigned long long n2) {
  const long long m0 = (long long)n0, m1 = (long long)n1, m2 = (long long)n2;
  const long long c0 = (((long long)(e % n0) + o0) % m0 + m0) % m0;
  const long long c1 = (((long long)((e / n0) % n1) + o1) % m1 + m1) % m1;
  const long long c2 = (((long long)(e / (n0 * n1)) + o2) % m2 + m2) % m2;
  return (unsigned long long)c0 + n0 * ((unsigned long long)c1 + n1 * (unsigned long long)c2);
}
SIMIT_DEV double simit_det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
SIMIT_DEV void simit_inv3(const double *a, double *r) {
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = -a[3] * a[8] + a[5] * a[6];
  const double c02 = a[3] * a[7] - a[4] * a[6], c10 = -a[1] * a[8] + a[2] * a[7];
  const double c11 = a[0] * a[8] - a[2] * a[6], c12 = -a[0] * a[7] + a[1] * a[6];
  const double c20 = a[1] * a[5] - a[2] * a[4], c21 = -a[0] * a[5] + a[2] * a[3];
  const double c22 = a[0] * a[4] - a[1] * a[3];
  double d = a[0] * c00 + a[1] * c01 + a[2] * c02;
  d = 1.0 / d;
  r[0] = c00 * d; r[1] = c10 * d; r[2] = c20 * d;
  r[3] = c01 * d; r[4] = c11 * d; r[5] = c21 * d;
  r[6] = c02 * d; r[7] = c12 * d; r[8] = c22 * d;
}
// ln x = k ln2 + 2 atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2))
SIMIT_DEV double simit_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  int k = 0;
  double m = frexp(x, &k);
  if (m < 0.70710678118654752440) { m = m * 2.0; k = k - 1; }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double t = 1.0 / 25.0;
  for (int n = 11; n >= 0; --n) t = t * s2 + 1.0 / (double)(2 * n + 1);
  return (double)k * 0.69314718055994530942 + 2.0 * s * t;
}
)";
  }

  // ---- emission helpers ----
  void line(const std::string &s) { F->body << std::string(2 * F->indent, ' ') << s << "\n"; }
  std::string fresh(const char *p = "t") { return std::string(p) + std::to_string(tmp++); }
  std::string declTemp(const Shape &sh) {
    const std::string n = fresh();
    if (sh.scalar()) line(ctype(sh.t) + " " + n + ";");
    else line(ctype(sh.t) + " " + n + "[" + std::to_string(sh.size()) + "];");
    return n;
  }
  // nested loops over dims; body(idx) returns the innermost statement
  void loopNest(const std::vector<int> &dims, const std::function<std::string(const Idx &)> &body) {
    Idx idx;
    const int base = F->indent;
    for (int d : dims) {
      const std::string v = fresh("i");
      line("for (int " + v + " = 0; " + v + " < " + std::to_string(d) + "; ++" + v + ") {");
      ++F->indent;
      idx.push_back(v);
    }
    line(body(idx));
    for (size_t k = 0; k < dims.size(); ++k) {
      --F->indent;
      line("}");
    }
    F->indent = base;
  }
  CVal arrayVal(const std::string &name, const Shape &sh) {
    CVal v;
    v.sh = sh;
    v.arr = name;
    const std::vector<int> d = sh.d;
    v.at = [name, d](const Idx &i) { return name + "[" + flat(i, d) + "]"; };
    return v;
  }
  CVal scalarVal(const std::string &s, char t) {
    CVal v;
    v.sh.t = t;
    v.s = s;
    return v;
  }
  // a named C array holding v (copies a view into a temporary)
  std::string materialize(const CVal &v) {
    if (!v.arr.empty()) return v.arr;
    const std::string n = declTemp(v.sh);
    loopNest(v.sh.d, [&](const Idx &i) { return n + "[" + flat(i, v.sh.d) + "] = " + v.at(i) + ";"; });
    return n;
  }
  std::string elem(const CVal &v, const Idx &i) { return v.sh.scalar() ? v.s : v.at(i); }

  // ---- types ----
  Shape shapeOf(const Type &t, Loc l) {
    Shape s;
    s.t = t.comp == "int" ? 'i' : t.comp == "bool" ? 'b' : 'f';
    if (t.kind == Type::Scalar) return s;
    std::vector<int> dims;
    if (t.kind != Type::Tensor || !denseDims(t, &dims))
      fail(l, "expected a dense tensor type, found " + t.str());
    s.d = dims;
    s.row = t.transposed && dims.size() == 1;
    return s;
  }
  const ElementDecl &elementDecl(const std::string &n, Loc l) {
    const ElementDecl *e = P.element(n);
    if (!e) fail(l, "unknown element type " + n);
    return *e;
  }
  // set name -> its element type and endpoint set
  void setInfo(const std::string &set, Loc l, std::string *elemType, std::string *epSet,
               int *card) {
    TypeP t = P.externType(set);
    if (!t || t->kind != Type::Set) fail(l, set + " is not an extern set");
    *elemType = t->elem;
    *card = (int)t->endpoints.size();
    epSet->clear();
    for (auto &ep : t->endpoints) {
      if (!epSet->empty() && ep != *epSet)
        fail(l, "edge sets over several vertex sets are not supported (" + set + ")");
      *epSet = ep;
    }
  }

  // ---- symbols ----
  Sym *lookup(const std::string &n) {
    auto it = F->syms.find(n);
    return it == F->syms.end() ? nullptr : &it->second;
  }
  Sym &declareLocal(const std::string &n, const Shape &sh, Loc l) {
    if (Sym *s = lookup(n)) {
      if (s->k != Sym::Local && s->k != Sym::Result) fail(l, n + " cannot be assigned");
      if (s->sh.size() != sh.size() || s->sh.d.size() != sh.d.size())
        fail(l, n + " changes its shape");
      return *s;
    }
    Sym s;
    s.k = Sym::Local;
    s.c = "s_" + n;
    s.sh = sh;
    F->decls.push_back({s.c, sh});
    return F->syms[n] = s;
  }
  CVal symVal(const Sym &s) {
    switch (s.k) {
      case Sym::ElemParam: {
        CVal v;
        v.ref = CVal::Elem;
        v.s = "e";
        return v;
      }
      case Sym::TupleParam: fail({}, "a tuple of endpoints is read by index");
      case Sym::GridParam: fail({}, "a grid's links are read as l[offset; offset]");
      case Sym::Output: fail({}, "results of a map function are written, not read");
      default: break;
    }
    if (s.sh.scalar()) return scalarVal(s.c, s.sh.t);
    return arrayVal(s.c, s.sh);
  }

  // ---- fields ----
  std::string fieldSlot(const std::string &set, const std::string &field, const Shape &sh) {
    const std::string key = set + "." + field;
    auto it = F->fieldSlot.find(key);
    if (it != F->fieldSlot.end()) return "a.a" + std::to_string(it->second);
    KernelArg a;
    a.kind = KernelArg::Field;
    a.set = set;
    a.field = field;
    a.index = sh.t == 'i' ? 1 : sh.t == 'b' ? 2 : 0;
    F->mk->args.push_back(a);
    const int slot = (int)F->mk->args.size() - 1;
    F->fieldSlot[key] = slot;
    return "a.a" + std::to_string(slot);
  }
  CVal fieldRead(const CVal &base, const std::string &field, Loc l) {
    if (!F->mk) fail(l, "fields are read in map functions only");
    std::string set, type;
    if (base.ref == CVal::Elem) {
      set = F->elemSet;
      type = F->elemType;
    } else if (base.ref == CVal::Endpoint) {
      set = F->epSet;
      type = F->epType;
    } else if (base.ref == CVal::Link) {
      set = F->gridSet;
      type = F->gridType;
    } else {
      fail(l, "only the element, its endpoints and grid links have fields");
    }
    const ElementDecl &ed = elementDecl(type, l);
    TypeP ft;
    for (auto &f : ed.fields)
      if (f.first == field) ft = f.second;
    if (!ft) fail(l, type + " has no field " + field);
    const Shape sh = shapeOf(*ft, l);
    const std::string p = fieldSlot(set, field, sh);
    const std::string n = std::to_string(sh.size());
    const std::string row = "(unsigned long long)(" + base.s + ")";
    auto load = [sh](const std::string &x) { return sh.t == 'b' ? "(" + x + " != 0)" : x; };
    if (sh.scalar()) return scalarVal(load(p + "[" + row + "]"), sh.t);
    CVal v;
    v.sh = sh;
    const std::vector<int> d = sh.d;
    v.at = [p, row, n, d, load](const Idx &i) {
      return load(p + "[" + row + "*" + n + "ull+" + flat(i, d) + "]");
    };
    return v;
  }

  // ---- expressions ----
  CVal eval(const ExprP &e) {
    switch (e->kind) {
      case Expr::Num:
        return e->isInt ? scalarVal(std::to_string((long long)e->num), 'i')
                        : scalarVal(num(e->num), 'f');
      case Expr::Bool: return scalarVal(e->num != 0 ? "true" : "false", 'b');
      case Expr::Str: fail(e->loc, "strings are not values in element functions");
      case Expr::Var: {
        if (Sym *s = lookup(e->name)) {
          if (s->k == Sym::TupleParam) fail(e->loc, "the endpoint tuple is read by index");
          if (s->k == Sym::Output) fail(e->loc, e->name + " is a result of the map function");
          return symVal(*s);
        }
        for (auto &c : P.consts)
          if (c->name == e->name) return eval(c->init);
        fail(e->loc, "unknown name " + e->name);
      }
      case Expr::Literal: {
        Shape sh;
        sh.t = e->litInt ? 'i' : 'f';
        sh.d = e->litDims;
        sh.row = sh.d.size() == 1;  // [a, b, c] is a row vector
        const std::string n = fresh();
        std::string init;
        for (size_t k = 0; k < e->lit.size(); ++k)
          init += (k ? ", " : "") +
                  (e->litInt ? std::to_string((long long)e->lit[k]) : num(e->lit[k]));
        line("const " + ctype(sh.t) + " " + n + "[" + std::to_string(sh.size()) + "] = {" + init +
             "};");
        return arrayVal(n, sh);
      }
      case Expr::Unary: {
        CVal a = eval(e->args[0]);
        if (e->op == "+") return a;
        if (e->op == "not") {
          if (!a.sh.scalar()) fail(e->loc, "not of a tensor");
          return scalarVal("(!" + a.s + ")", 'b');
        }
        if (a.sh.scalar()) return scalarVal("(-" + a.s + ")", a.sh.t);
        const std::string n = declTemp(a.sh);
        loopNest(a.sh.d,
                 [&](const Idx &i) { return n + "[" + flat(i, a.sh.d) + "] = -" + a.at(i) + ";"; });
        CVal r = arrayVal(n, a.sh);
        r.sh.row = a.sh.row;
        return r;
      }
      case Expr::Transpose: {
        CVal a = eval(e->args[0]);
        if (a.sh.scalar()) return a;
        if (a.sh.d.size() == 1) {
          a.sh.row = !a.sh.row;
          return a;
        }
        if (a.sh.d.size() != 2) fail(e->loc, "transpose of a tensor of order > 2");
        CVal r;
        r.sh = a.sh;
        std::swap(r.sh.d[0], r.sh.d[1]);
        auto at = a.at;
        r.at = [at](const Idx &i) { return at({i[1], i[0]}); };
        return r;
      }
      case Expr::Binary: return binary(e);
      case Expr::Field: return fieldRead(eval(e->args[0]), e->name, e->loc);
      case Expr::Apply: return apply(e);
      case Expr::Map: fail(e->loc, "a map inside an element function");
      case Expr::SetRead: return setRead(e);
    }
    fail(e->loc, "unsupported expression");
  }

  CVal elementwise(const std::string &op, const CVal &a, const CVal &b, Loc l) {
    const CVal &arr = a.sh.scalar() ? b : a;
    if (!a.sh.scalar() && !b.sh.scalar()) {
      if (a.sh.d != b.sh.d) fail(l, "operands of " + op + " have different dimensions");
      if (a.sh.d.size() == 1 && a.sh.row != b.sh.row)
        fail(l, "operands of " + op + " are a row and a column vector");
    }
    Shape sh = arr.sh;
    sh.t = (a.sh.t == 'f' || b.sh.t == 'f') ? 'f' : a.sh.t;
    const std::string cop = op == ".*" ? "*" : op == "./" ? "/" : op;
    const std::string n = declTemp(sh);
    loopNest(sh.d, [&](const Idx &i) {
      return n + "[" + flat(i, sh.d) + "] = " + elem(a, i) + " " + cop + " " + elem(b, i) + ";";
    });
    CVal r = arrayVal(n, sh);
    r.sh.row = arr.sh.row;
    return r;
  }

  // s = sum_k a(k) b(k) in ascending k, from 0
  std::string sumProducts(int K, const std::function<std::string(const std::string &)> &ta,
                          const std::function<std::string(const std::string &)> &tb) {
    const std::string s = fresh("s"), k = fresh("k");
    line("double " + s + " = 0.0;");
    line("for (int " + k + " = 0; " + k + " < " + std::to_string(K) + "; ++" + k + ") " + s +
         " = " + s + " + " + ta(k) + " * " + tb(k) + ";");
    return s;
  }

  CVal product(const CVal &a, const CVal &b, Loc l) {
    if (a.sh.scalar() || b.sh.scalar()) {
      if (a.sh.scalar() && b.sh.scalar()) {
        const char t = (a.sh.t == 'f' || b.sh.t == 'f') ? 'f' : 'i';
        return scalarVal("(" + a.s + " * " + b.s + ")", t);
      }
      return elementwise("*", a, b, l);
    }
    const auto &da = a.sh.d, &db = b.sh.d;
    if (da.size() == 1 && db.size() == 1) {
      if (da[0] != db[0]) fail(l, "vector lengths differ");
      if (a.sh.row && !b.sh.row)  // inner product
        return scalarVal(sumProducts(da[0], [&](const std::string &k) { return a.at({k}); },
                                     [&](const std::string &k) { return b.at({k}); }),
                         'f');
      if (!a.sh.row && b.sh.row) {  // outer product
        Shape sh;
        sh.d = {da[0], db[0]};
        const std::string n = declTemp(sh);
        loopNest(sh.d, [&](const Idx &i) {
          return n + "[" + flat(i, sh.d) + "] = " + a.at({i[0]}) + " * " + b.at({i[1]}) + ";";
        });
        return arrayVal(n, sh);
      }
      fail(l, "product of two column (or two row) vectors; use .* or a transpose");
    }
    if (da.size() == 2 && db.size() == 1) {  // matrix * column vector
      if (b.sh.row || da[1] != db[0]) fail(l, "matrix-vector dimensions do not match");
      Shape sh;
      sh.d = {da[0]};
      const std::string n = declTemp(sh), i = fresh("i");
      line("for (int " + i + " = 0; " + i + " < " + std::to_string(da[0]) + "; ++" + i + ") {");
      ++F->indent;
      const std::string s = sumProducts(da[1], [&](const std::string &k) { return a.at({i, k}); },
                                        [&](const std::string &k) { return b.at({k}); });
      line(n + "[" + i + "] = " + s + ";");
      --F->indent;
      line("}");
      return arrayVal(n, sh);
    }
    if (da.size() == 1 && db.size() == 2) {  // row vector * matrix
      if (!a.sh.row || da[0] != db[0]) fail(l, "vector-matrix dimensions do not match");
      Shape sh;
      sh.d = {db[1]};
      sh.row = true;
      const std::string n = declTemp(sh), j = fresh("j");
      line("for (int " + j + " = 0; " + j + " < " + std::to_string(db[1]) + "; ++" + j + ") {");
      ++F->indent;
      const std::string s = sumProducts(da[0], [&](const std::string &k) { return a.at({k}); },
                                        [&](const std::string &k) { return b.at({k, j}); });
      line(n + "[" + j + "] = " + s + ";");
      --F->indent;
      line("}");
      CVal r = arrayVal(n, sh);
      r.sh.row = true;
      return r;
    }
    if (da.size() == 2 && db.size() == 2) {
      if (da[1] != db[0]) fail(l, "matrix dimensions do not match");
      Shape sh;
      sh.d = {da[0], db[1]};
      const std::string n = declTemp(sh), i = fresh("i"), j = fresh("j");
      line("for (int " + i + " = 0; " + i + " < " + std::to_string(da[0]) + "; ++" + i + ")");
      line("for (int " + j + " = 0; " + j + " < " + std::to_string(db[1]) + "; ++" + j + ") {");
      ++F->indent;
      const std::string s = sumProducts(da[1], [&](const std::string &k) { return a.at({i, k}); },
                                        [&](const std::string &k) { return b.at({k, j}); });
      line(n + "[" + i + "*" + std::to_string(db[1]) + "+" + j + "] = " + s + ";");
      --F->indent;
      line("}");
      return arrayVal(n, sh);
    }
    fail(l, "products of tensors of order > 2 are not supported");
  }

  CVal binary(const ExprP &e) {
    CVal a = eval(e->args[0]), b = eval(e->args[1]);
    const std::string &op = e->op;
    if (a.ref != CVal::None || b.ref != CVal::None) fail(e->loc, "arithmetic on an element");
    if (op == "*") return product(a, b, e->loc);
    if (op == "+" || op == "-" || op == ".*" || op == "./" || op == "/") {
      if (a.sh.scalar() && b.sh.scalar()) {
        const char t = (a.sh.t == 'f' || b.sh.t == 'f') ? 'f' : 'i';
        const std::string c = op == ".*" ? "*" : op == "./" ? "/" : op;
        return scalarVal("(" + a.s + " " + c + " " + b.s + ")", t);
      }
      if (op == "/" && !b.sh.scalar()) fail(e->loc, "division by a tensor; use ./");
      return elementwise(op, a, b, e->loc);
    }
    if (!a.sh.scalar() || !b.sh.scalar()) fail(e->loc, op + " of tensors");
    if (op == "^") return scalarVal("pow((double)" + a.s + ", (double)" + b.s + ")", 'f');
    if (op == "and") return scalarVal("(" + a.s + " && " + b.s + ")", 'b');
    if (op == "or") return scalarVal("(" + a.s + " || " + b.s + ")", 'b');
    if (op == "xor") return scalarVal("((bool)" + a.s + " != (bool)" + b.s + ")", 'b');
    return scalarVal("(" + a.s + " " + op + " " + b.s + ")", 'b');
  }

  static bool isIntrinsic(const std::string &n) {
    static const std::set<std::string> k = {"norm", "dot",  "det",  "inv",  "log",   "sqrt",
                                            "exp",  "sin",  "cos",  "tan",  "asin",  "acos",
                                            "atan", "atan2", "pow", "abs",  "cross", "transpose"};
    return k.count(n) > 0;
  }

  CVal intrinsic(const std::string &n, const std::vector<ExprP> &args, Loc l) {
    std::vector<CVal> a;
    for (auto &x : args) {
      if (!x) fail(l, "slices are not arguments");
      a.push_back(eval(x));
    }
    auto want = [&](size_t k) {
      if (a.size() != k) fail(l, n + " takes " + std::to_string(k) + " argument(s)");
    };
    if (n == "norm") {
      want(1);
      if (a[0].sh.d.size() != 1) fail(l, "norm of a non-vector");
      const std::string s = sumProducts(a[0].sh.d[0], [&](const std::string &k) { return a[0].at({k}); },
                                        [&](const std::string &k) { return a[0].at({k}); });
      return scalarVal("sqrt(" + s + ")", 'f');
    }
    if (n == "dot") {
      want(2);
      if (a[0].sh.d.size() != 1 || a[0].sh.d != a[1].sh.d) fail(l, "dot of unequal vectors");
      return scalarVal(sumProducts(a[0].sh.d[0], [&](const std::string &k) { return a[0].at({k}); },
                                   [&](const std::string &k) { return a[1].at({k}); }),
                       'f');
    }
    if (n == "det" || n == "inv") {
      want(1);
      if (a[0].sh.d != std::vector<int>{3, 3}) fail(l, n + " of a matrix other than 3x3");
      const std::string m = materialize(a[0]);
      if (n == "det") return scalarVal("simit_det3(" + m + ")", 'f');
      Shape sh;
      sh.d = {3, 3};
      const std::string r = declTemp(sh);
      line("simit_inv3(" + m + ", " + r + ");");
      return arrayVal(r, sh);
    }
    if (n == "transpose") {
      want(1);
      auto e = std::make_shared<Expr>();
      e->kind = Expr::Transpose;
      e->args = {args[0]};
      return eval(e);
    }
    if (n == "cross") {
      want(2);
      if (a[0].sh.d != std::vector<int>{3} || a[1].sh.d != std::vector<int>{3})
        fail(l, "cross of vectors other than 3-vectors");
      Shape sh;
      sh.d = {3};
      const std::string r = declTemp(sh);
      auto x = [&](int i, int k) { return a[i].at({std::to_string(k)}); };
      line(r + "[0] = " + x(0, 1) + " * " + x(1, 2) + " - " + x(0, 2) + " * " + x(1, 1) + ";");
      line(r + "[1] = " + x(0, 2) + " * " + x(1, 0) + " - " + x(0, 0) + " * " + x(1, 2) + ";");
      line(r + "[2] = " + x(0, 0) + " * " + x(1, 1) + " - " + x(0, 1) + " * " + x(1, 0) + ";");
      return arrayVal(r, sh);
    }
    for (auto &v : a)
      if (!v.sh.scalar()) fail(l, n + " of a tensor");
    if (n == "atan2" || n == "pow") {
      want(2);
      return scalarVal(n + "((double)" + a[0].s + ", (double)" + a[1].s + ")", 'f');
    }
    want(1);
    if (n == "log") return scalarVal("simit_log((double)" + a[0].s + ")", 'f');
    if (n == "abs") return scalarVal("fabs((double)" + a[0].s + ")", 'f');
    return scalarVal(n + "((double)" + a[0].s + ")", 'f');
  }

  // grid offsets of a set read: constants, one per grid dimension
  std::vector<long> offsets(const ExprP &e, size_t from, size_t count) {
    std::vector<long> o;
    for (size_t k = 0; k < count; ++k) o.push_back(evalConstInt(e->args[from + k], {}));
    return o;
  }
  std::string nbr(const std::vector<long> &o) {
    std::string c = "simit_nbr(e";
    for (int k = 0; k < 3; ++k) c += ", " + std::to_string(k < (int)o.size() ? o[k] : 0) + "ll";
    return c + ", " + F->dims[0] + ", " + F->dims[1] + ", " + F->dims[2] + ")";
  }
  // points[o] (a neighbour) or l[a;b] (the link between the points at a and b)
  CVal setRead(const ExprP &e) {
    if (!F->mk || !F->mk->gridDims) fail(e->loc, "grid reads need a map through a grid");
    const int nd = F->mk->gridDims;
    const ExprP &base = e->args[0];
    const size_t n = e->args.size() - 1;
    Sym *s = lookup(base->name);
    CVal v;
    if (s && s->k == Sym::GridParam) {
      if (e->split != nd || (int)n != 2 * nd)
        fail(e->loc, "a link is read as l[a;b] with two offsets of " + std::to_string(nd) +
                         " coordinates");
      const std::vector<long> a = offsets(e, 1, nd), b = offsets(e, 1 + nd, nd);
      int axis = -1;
      long dir = 0;
      for (int k = 0; k < nd; ++k)
        if (b[k] != a[k]) {
          if (axis >= 0 || (b[k] - a[k] != 1 && b[k] - a[k] != -1))
            fail(e->loc, "l[a;b] must name two neighbouring grid points");
          axis = k;
          dir = b[k] - a[k];
        }
      if (axis < 0) fail(e->loc, "l[a;b] must name two neighbouring grid points");
      // the link leaves the lower of the two points along its axis
      v.ref = CVal::Link;
      v.s = "(" + nbr(dir > 0 ? a : b) + "*" + std::to_string(nd) + "ull+" + std::to_string(axis) +
            "ull)";
      return v;
    }
    if (!s && base->name == F->elemSet) {
      if (e->split != (int)n || (int)n != nd)
        fail(e->loc, base->name + "[...] takes one offset of " + std::to_string(nd) +
                         " coordinates");
      v.ref = CVal::Endpoint;
      v.s = nbr(offsets(e, 1, nd));
      return v;
    }
    fail(e->loc, "only the grid argument and the grid's point set are read with [...]");
  }

  CVal apply(const ExprP &e) {
    const ExprP &base = e->args[0];
    std::vector<ExprP> args(e->args.begin() + 1, e->args.end());
    if (base->kind == Expr::Var) {
      Sym *s = lookup(base->name);
      if (s && s->k == Sym::TupleParam) {
        if (args.size() != 1 || !args[0]) fail(e->loc, "an endpoint tuple takes one index");
        CVal k = eval(args[0]);
        if (!k.sh.scalar() || k.sh.t != 'i') fail(e->loc, "endpoint index must be an int");
        CVal v;
        v.ref = CVal::Endpoint;
        v.s = "a.a1[e*" + std::to_string(F->mk->card) + "ull+(unsigned long long)(" +
              index(args[0], k.s, F->mk->card) + ")]";
        return v;
      }
      if (s && s->k == Sym::Output)
        fail(e->loc, base->name + " is a result of the map function (written, not read)");
      if (!s && isIntrinsic(base->name)) return intrinsic(base->name, args, e->loc);
      if (!s && P.func(base->name)) {
        std::vector<CVal> rs = callUser(*P.func(base->name), args, e->loc);
        if (rs.size() != 1) fail(e->loc, base->name + " does not return one value");
        return rs[0];
      }
    }
    CVal t = eval(base);
    if (t.sh.scalar() || t.ref != CVal::None) fail(e->loc, "indexing a non-tensor");
    if (args.size() != t.sh.d.size()) fail(e->loc, "wrong number of indices");
    Idx fixed(args.size());
    std::vector<int> freeDims, freePos;
    for (size_t k = 0; k < args.size(); ++k) {
      if (!args[k]) {
        freeDims.push_back(t.sh.d[k]);
        freePos.push_back((int)k);
        continue;
      }
      CVal i = eval(args[k]);
      if (!i.sh.scalar() || i.sh.t != 'i') fail(e->loc, "tensor indices must be ints");
      fixed[k] = index(args[k], i.s, t.sh.d[k]);
    }
    if (freeDims.empty()) return scalarVal(t.at(fixed), t.sh.t);
    CVal v;
    v.sh.t = t.sh.t;
    v.sh.d = freeDims;
    v.sh.row = freeDims.size() == 1 && freePos[0] == 1 && args.size() == 2;
    auto at = t.at;
    v.at = [at, fixed, freePos](const Idx &i) {
      Idx full = fixed;
      for (size_t k = 0; k < freePos.size(); ++k) full[freePos[k]] = i[k];
      return at(full);
    };
    return v;
  }

  // ---- user functions called from element code: __device__ functions ----
  std::vector<CVal> callUser(const Func &f, const std::vector<ExprP> &args, Loc l) {
    if (!f.generics.empty()) fail(l, "generic functions in element code are not supported");
    if (args.size() != f.params.size())
      fail(l, f.name + " takes " + std::to_string(f.params.size()) + " arguments");
    deviceFunction(f);
    std::vector<std::string> cargs;
    for (size_t k = 0; k < args.size(); ++k) {
      if (!args[k]) fail(l, "slices are not arguments");
      if (f.params[k].inout) fail(l, "inout arguments in element code are not supported");
      CVal v = eval(args[k]);
      const Shape ps = shapeOf(*f.params[k].type, l);
      if (ps.size() != v.sh.size() || ps.d.size() != v.sh.d.size())
        fail(args[k]->loc, "argument " + f.params[k].name + " of " + f.name + " has the wrong shape");
      cargs.push_back(v.sh.scalar() ? v.s : materialize(v));
    }
    std::vector<CVal> rs;
    for (auto &r : f.results) {
      const Shape rsh = shapeOf(*r.type, l);
      const std::string n = declTemp(rsh);
      cargs.push_back(rsh.scalar() ? "&" + n : n);
      rs.push_back(rsh.scalar() ? scalarVal(n, rsh.t) : arrayVal(n, rsh));
    }
    std::string call = "u_" + f.name + "(";
    for (size_t k = 0; k < cargs.size(); ++k) call += (k ? ", " : "") + cargs[k];
    line(call + ");");
    return rs;
  }

  void deviceFunction(const Func &f) {
    if (devDone.count(f.name)) return;
    devDone.insert(f.name);
    Fn fn;
    Fn *saved = F;
    F = &fn;
    std::string sig = "__device__ void u_" + f.name + "(";
    bool first = true;
    auto add = [&](const std::string &s) {
      sig += (first ? "" : ", ") + s;
      first = false;
    };
    for (auto &p : f.params) {
      Sym s;
      s.sh = shapeOf(*p.type, f.loc);
      s.c = "p_" + p.name;
      s.k = s.sh.scalar() ? Sym::Param : Sym::ArrParam;
      add((s.sh.scalar() ? ctype(s.sh.t) + " " : "const " + ctype(s.sh.t) + " *") + s.c);
      fn.syms[p.name] = s;
    }
    for (auto &r : f.results) {
      Sym s;
      s.sh = shapeOf(*r.type, f.loc);
      s.k = Sym::Result;
      add(ctype(s.sh.t) + " *r_" + r.name);
      s.c = s.sh.scalar() ? "(*r_" + r.name + ")" : "r_" + r.name;
      fn.syms[r.name] = s;
      if (s.sh.scalar()) line(s.c + " = 0;");
      else line("for (int z = 0; z < " + std::to_string(s.sh.size()) + "; ++z) " + s.c + "[z] = 0;");
    }
    block(f.body);
    std::ostringstream out;
    out << sig << ") {\n" << decls() << fn.body.str() << "}\n";
    F = saved;
    dev << out.str();
  }

  std::string decls() {
    std::string s;
    for (auto &d : F->decls) {
      if (d.second.scalar()) s += "  " + ctype(d.second.t) + " " + d.first + " = 0;\n";
      else
        s += "  " + ctype(d.second.t) + " " + d.first + "[" + std::to_string(d.second.size()) +
             "] = {};\n";
    }
    return s;
  }

  // ---- statements ----
  void block(const Block &b) {
    for (auto &s : b) stmt(s);
  }

  void assignTo(Sym &dst, const CVal &v, Loc l) {
    if (dst.sh.scalar()) {
      if (!v.sh.scalar()) fail(l, "assigning a tensor to a scalar");
      line(dst.c + " = " + v.s + ";");
      return;
    }
    if (!v.sh.scalar() && v.sh.size() != dst.sh.size()) fail(l, "assignment changes a shape");
    const std::vector<int> d = dst.sh.d;
    const std::string c = dst.c;
    if (v.sh.scalar()) {
      line("for (int z = 0; z < " + std::to_string(dst.sh.size()) + "; ++z) " + c + "[z] = " +
           v.s + ";");
      return;
    }
    const std::vector<int> vd = v.sh.d;
    loopNest(d, [&](const Idx &i) {
      // same component count; index v by its own dims when they differ
      if (vd == d) return c + "[" + flat(i, d) + "] = " + v.at(i) + ";";
      fail(l, "assignment between tensors of different dimensions");
    });
  }

  long evalConstInt(const ExprP &e, const std::map<std::string, long> &vals) {
    switch (e->kind) {
      case Expr::Num:
        if (e->isInt) return (long)e->num;
        break;
      case Expr::Var: {
        auto it = vals.find(e->name);
        if (it != vals.end()) return it->second;
        break;
      }
      case Expr::Unary:
        if (e->op == "-") return -evalConstInt(e->args[0], vals);
        if (e->op == "+") return evalConstInt(e->args[0], vals);
        break;
      case Expr::Binary: {
        const long a = evalConstInt(e->args[0], vals), b = evalConstInt(e->args[1], vals);
        if (e->op == "+") return a + b;
        if (e->op == "-") return a - b;
        if (e->op == "*") return a * b;
        break;
      }
      default: break;
    }
    throw CodegenError(e->loc, "not a constant integer");
  }

  // [lo, hi] of an integer index expression over constants and the
  // variables of loops with constant bounds; false if it cannot be bounded
  bool indexRange(const ExprP &e, long *lo, long *hi) {
    switch (e->kind) {
      case Expr::Num:
        if (!e->isInt) return false;
        *lo = *hi = (long)e->num;
        return true;
      case Expr::Var: {
        Sym *s = lookup(e->name);
        if (!s || s->k != Sym::Loop) return false;
        for (auto &L : F->loops)
          if (L.c == s->c && L.constant) {
            *lo = L.lo;
            *hi = L.lo + (L.trip > 0 ? L.trip - 1 : 0);
            return true;
          }
        return false;
      }
      case Expr::Unary: {
        long a, b;
        if (e->op != "-" || !indexRange(e->args[0], &a, &b)) return false;
        *lo = -b;
        *hi = -a;
        return true;
      }
      case Expr::Binary: {
        long a0, a1, b0, b1;
        if (!indexRange(e->args[0], &a0, &a1) || !indexRange(e->args[1], &b0, &b1)) return false;
        if (e->op == "+") {
          *lo = a0 + b0;
          *hi = a1 + b1;
        } else if (e->op == "-") {
          *lo = a0 - b1;
          *hi = a1 - b0;
        } else if (e->op == "*") {
          const long c[4] = {a0 * b0, a0 * b1, a1 * b0, a1 * b1};
          *lo = *std::min_element(c, c + 4);
          *hi = *std::max_element(c, c + 4);
        } else {
          return false;
        }
        return true;
      }
      default: return false;
    }
  }
  // an index into a dimension of n: checked here when it can be bounded,
  // clamped in the kernel otherwise (the kernel never leaves its arrays)
  std::string index(const ExprP &e, const std::string &c, int n) {
    long lo, hi;
    if (indexRange(e, &lo, &hi)) {
      if (lo < 0 || hi >= n)
        fail(e->loc, "index out of range [0, " + std::to_string(n) + ")");
      return c;
    }
    return "simit_clamp(" + c + ", " + std::to_string(n) + ")";
  }

  void outputWrite(const StmtP &s, const ExprP &lhs, Sym &out) {
    MapOutput &mo = F->mk->outputs[out.out];
    std::vector<ExprP> idx(lhs->args.begin() + 1, lhs->args.end());
    if (idx.size() != (mo.matrix ? 2u : 1u))
      fail(lhs->loc, mo.name + " takes " + std::string(mo.matrix ? "two" : "one") + " index(es)");
    for (auto &i : idx) {
      if (!i) fail(lhs->loc, "a result is written at an element or an endpoint");
      CVal v = eval(i);
      if (v.ref == CVal::None || v.ref == CVal::Link)
        fail(i->loc, "a result is written at an element or an endpoint");
    }
    if (F->whileDepth) fail(s->loc, "results written inside a while loop");
    for (auto &L : F->loops)
      if (!L.constant) fail(s->loc, "results written inside a loop with non-constant bounds");
    if (F->sites.count(s.get())) fail(s->loc, "one statement writes one result");
    Site site;
    site.out = out.out;
    site.base = F->nextInst[out.out];
    site.loops = F->loops;
    site.idx = idx;
    long count = 1;
    for (auto &L : F->loops) count *= L.trip;
    F->nextInst[out.out] += (int)count;
    // slot = base + sum (var - lo) * stride
    std::string inst = std::to_string(site.base);
    long stride = 1;
    for (size_t k = F->loops.size(); k-- > 0;) {
      const LoopInfo &L = F->loops[k];
      inst += "+(" + L.c + "-(" + std::to_string(L.lo) + "))*" + std::to_string(stride);
      stride *= L.trip;
    }
    F->sites[s.get()] = site;
    const CVal v = eval(s->rhs);
    if (!v.sh.scalar() && v.sh.size() != mo.bs)
      fail(s->loc, "value written to " + mo.name + " has the wrong shape");
    const std::string ninst = "@NINST" + std::to_string(out.out) + "@";
    const std::string st = "a.a" + std::to_string(stageSlot(out.out));
    const std::string o = fresh("o");
    line("{");
    ++F->indent;
    line("const unsigned long long " + o + " = (e*" + ninst + "ull+(" + inst + "))*" +
         std::to_string(mo.bs) + "ull;");
    if (v.sh.scalar()) {
      line("for (int z = 0; z < " + std::to_string(mo.bs) + "; ++z) " + st + "[" + o +
           "+z] = " + v.s + ";");
    } else {
      loopNest(v.sh.d, [&](const Idx &i) {
        return st + "[" + o + "+" + flat(i, v.sh.d) + "] = " + v.at(i) + ";";
      });
    }
    if (!F->mk->req.reduce)
      line("a.a" + std::to_string(stageSlot(out.out) + 1) + "[e*" + ninst + "ull+(" + inst +
           ")] = 1;");
    --F->indent;
    line("}");
  }
  int stageSlot(int out) {
    for (size_t k = 0; k < F->mk->args.size(); ++k)
      if (F->mk->args[k].kind == KernelArg::Stage && F->mk->args[k].index == out) return (int)k;
    return -1;
  }

  void stmt(const StmtP &s) {
    switch (s->kind) {
      case Stmt::VarDecl:
      case Stmt::ConstDecl: {
        Shape sh;
        CVal init;
        if (s->init) init = eval(s->init);
        if (s->type) sh = shapeOf(*s->type, s->loc);
        else sh = init.sh;
        if (init.ref != CVal::None) fail(s->loc, "an element is not a value");
        Sym &d = declareLocal(s->name, sh, s->loc);
        if (s->init) {
          assignTo(d, init, s->loc);
        } else if (sh.scalar()) {
          line(d.c + " = 0;");
        } else {
          line("for (int z = 0; z < " + std::to_string(sh.size()) + "; ++z) " + d.c + "[z] = 0;");
        }
        return;
      }
      case Stmt::Assign: return assign(s);
      case Stmt::ExprStmt: {
        if (s->rhs->kind != Expr::Apply) fail(s->loc, "expression statement has no effect");
        eval(s->rhs);
        return;
      }
      case Stmt::If: {
        const int depth = F->indent;
        for (size_t k = 0; k < s->conds.size(); ++k) {
          CVal c = eval(s->conds[k]);
          if (!c.sh.scalar()) fail(s->conds[k]->loc, "condition is not a scalar");
          line("if (" + c.s + ") {");
          ++F->indent;
          block(s->blocks[k]);
          --F->indent;
          line("} else {");
          ++F->indent;
        }
        if (s->blocks.size() > s->conds.size()) block(s->blocks.back());
        while (F->indent > depth) {
          --F->indent;
          line("}");
        }
        return;
      }
      case Stmt::While: {
        line("while (true) {");
        ++F->indent;
        ++F->whileDepth;
        CVal c = eval(s->rhs);
        if (!c.sh.scalar()) fail(s->rhs->loc, "condition is not a scalar");
        line("if (!(" + c.s + ")) break;");
        block(s->body);
        --F->whileDepth;
        --F->indent;
        line("}");
        return;
      }
      case Stmt::For: {
        if (!s->overSet.empty()) fail(s->loc, "loops over sets in element code");
        CVal lo = eval(s->lo), hi = eval(s->hi);
        if (!lo.sh.scalar() || !hi.sh.scalar() || lo.sh.t != 'i' || hi.sh.t != 'i')
          fail(s->loc, "loop bounds must be ints");
        LoopInfo L;
        L.name = s->name;
        L.c = "s_" + s->name + "_" + std::to_string(tmp++);
        try {
          L.lo = evalConstInt(s->lo, {});
          L.trip = evalConstInt(s->hi, {}) - L.lo;
          if (L.trip < 0) L.trip = 0;
          L.constant = true;
        } catch (const CodegenError &) {
          L.constant = false;
        }
        line("for (int " + L.c + " = " + lo.s + "; " + L.c + " < " + hi.s + "; ++" + L.c + ") {");
        ++F->indent;
        Sym prev;
        const bool had = F->syms.count(s->name) > 0;
        if (had) prev = F->syms[s->name];
        Sym v;
        v.k = Sym::Loop;
        v.c = L.c;
        v.sh.t = 'i';
        F->syms[s->name] = v;
        F->loops.push_back(L);
        block(s->body);
        F->loops.pop_back();
        if (had) F->syms[s->name] = prev;
        else F->syms.erase(s->name);
        --F->indent;
        line("}");
        return;
      }
      case Stmt::Print: line("// print (not executed on the device)"); return;
      case Stmt::Apply: fail(s->loc, "apply inside an element function");
    }
  }

  void assign(const StmtP &s) {
    if (s->lhs.size() > 1) {
      if (s->rhs->kind != Expr::Apply || s->rhs->args[0]->kind != Expr::Var ||
          !P.func(s->rhs->args[0]->name))
        fail(s->loc, "several targets need a function call with several results");
      std::vector<ExprP> args(s->rhs->args.begin() + 1, s->rhs->args.end());
      std::vector<CVal> rs = callUser(*P.func(s->rhs->args[0]->name), args, s->loc);
      if (rs.size() != s->lhs.size()) fail(s->loc, "result count does not match");
      for (size_t k = 0; k < rs.size(); ++k) {
        if (s->lhs[k]->kind != Expr::Var) fail(s->loc, "targets must be names");
        assignTo(declareLocal(s->lhs[k]->name, rs[k].sh, s->loc), rs[k], s->loc);
      }
      return;
    }
    const ExprP &lhs = s->lhs[0];
    if (lhs->kind == Expr::Var) {
      Sym *d = lookup(lhs->name);
      if (d && d->k == Sym::Output) fail(s->loc, lhs->name + " is written at an element index");
      if (d && d->k != Sym::Local && d->k != Sym::Result)
        fail(s->loc, lhs->name + " cannot be assigned");
      const CVal v = eval(s->rhs);
      if (v.ref != CVal::None) fail(s->loc, "an element is not a value");
      assignTo(d ? *d : declareLocal(lhs->name, v.sh, s->loc), v, s->loc);
      return;
    }
    if (lhs->kind == Expr::Field) {  // inout element field
      const ExprP &b = lhs->args[0];
      Sym *es = b->kind == Expr::Var ? lookup(b->name) : nullptr;
      if (!es || es->k != Sym::ElemParam || !es->inout)
        fail(s->loc, "only fields of an inout element can be written");
      const ElementDecl &ed = elementDecl(F->elemType, s->loc);
      TypeP ft;
      for (auto &f : ed.fields)
        if (f.first == lhs->name) ft = f.second;
      if (!ft) fail(s->loc, F->elemType + " has no field " + lhs->name);
      const Shape sh = shapeOf(*ft, s->loc);
      if (sh.t != 'f') fail(s->loc, "only float fields can be written");
      const std::string p = fieldSlot(F->elemSet, lhs->name, sh);
      F->mk->fieldsWritten.push_back({F->elemSet, lhs->name});
      const CVal v = eval(s->rhs);
      if (!v.sh.scalar() && v.sh.size() != sh.size()) fail(s->loc, "field has another shape");
      const std::string n = std::to_string(sh.size());
      if (sh.scalar()) {
        if (!v.sh.scalar()) fail(s->loc, "field has another shape");
        line(p + "[e] = " + v.s + ";");
      } else {
        loopNest(sh.d, [&](const Idx &i) {
          return p + "[e*" + n + "ull+" + flat(i, sh.d) + "] = " + elem(v, i) + ";";
        });
      }
      return;
    }
    if (lhs->kind == Expr::Apply && lhs->args[0]->kind == Expr::Var) {
      Sym *d = lookup(lhs->args[0]->name);
      if (d && d->k == Sym::Output) return outputWrite(s, lhs, *d);
      if (!d || (d->k != Sym::Local && d->k != Sym::Result) || d->sh.scalar())
        fail(s->loc, "cannot write an element of " + lhs->args[0]->name);
      std::vector<ExprP> idx(lhs->args.begin() + 1, lhs->args.end());
      if (idx.size() != d->sh.d.size()) fail(s->loc, "wrong number of indices");
      Idx fixed(idx.size());
      std::vector<int> freePos, freeDims;
      for (size_t k = 0; k < idx.size(); ++k) {
        if (!idx[k]) {
          freePos.push_back((int)k);
          freeDims.push_back(d->sh.d[k]);
          continue;
        }
        CVal i = eval(idx[k]);
        if (!i.sh.scalar() || i.sh.t != 'i') fail(s->loc, "tensor indices must be ints");
        fixed[k] = index(idx[k], i.s, d->sh.d[k]);
      }
      const CVal v = eval(s->rhs);
      const std::string c = d->c;
      const std::vector<int> dd = d->sh.d;
      if (freePos.empty()) {
        if (!v.sh.scalar()) fail(s->loc, "assigning a tensor to one component");
        line(c + "[" + flat(fixed, dd) + "] = " + v.s + ";");
        return;
      }
      if (!v.sh.scalar() && v.sh.d != freeDims) fail(s->loc, "slice has another shape");
      loopNest(freeDims, [&](const Idx &i) {
        Idx full = fixed;
        for (size_t k = 0; k < freePos.size(); ++k) full[freePos[k]] = i[k];
        return c + "[" + flat(full, dd) + "] = " + elem(v, i) + ";";
      });
      return;
    }
    fail(s->loc, "unsupported assignment target");
  }

  // ---- one map kernel ----
  std::string mapKernel(MapKernel &mk) {
    Fn fn;
    F = &fn;
    fn.mk = &mk;
    const Func *f = P.func(mk.req.func);
    const Loc l = f ? f->loc : Loc{};
    if (!f) fail(l, "map of unknown function " + mk.req.func);
    setInfo(mk.req.target, l, &fn.elemType, &fn.epSet, &mk.card);
    fn.elemSet = mk.req.target;
    mk.endpointSet = fn.epSet;
    if (!fn.epSet.empty()) {
      TypeP et = P.externType(fn.epSet);
      if (!et || et->kind != Type::Set) fail(l, fn.epSet + " is not a set");
      fn.epType = et->elem;
    }
    // slots: 0 count, 1 endpoints, then stage (+ taken) per output
    mk.args.push_back({KernelArg::Count, "", "", 0});
    mk.args.push_back({KernelArg::Endpoints, "", "", 0});
    const bool grid = !mk.req.through.empty();
    if (grid) {
      TypeP gt = P.externType(mk.req.through);
      if (mk.card) fail(l, "a map through a grid runs over the grid's point set");
      if (!gt || gt->kind != Type::Set || !gt->gridDims)
        fail(l, mk.req.through + " is not a grid edge set");
      if (gt->endpoints[0] != mk.req.target)
        fail(l, "grid " + mk.req.through + " is over " + gt->endpoints[0] + ", not " +
                    mk.req.target);
      mk.gridDims = gt->gridDims;
      fn.gridSet = mk.req.through;
      fn.gridType = gt->elem;
      fn.epSet = mk.req.target;  // neighbours are points of the same set
      fn.epType = fn.elemType;
      for (int k = 0; k < 3; ++k) {
        mk.args.push_back({KernelArg::GridDim, "", "", k});
        fn.dims[k] = "a.a" + std::to_string(mk.args.size() - 1);
      }
    }
    const size_t np = f->params.size();
    const size_t expect = (size_t)mk.req.nextra + (mk.card ? 2 : 1) + (grid ? 1 : 0);
    if (np != expect)
      fail(f->loc, "map " + f->name + " to " + mk.req.target + ": the function takes " +
                       std::to_string(np) + " arguments, the map supplies " +
                       std::to_string(expect));
    for (int k = 0; k < mk.req.nextra; ++k) {
      const Param &p = f->params[k];
      Shape sh = shapeOf(*p.type, f->loc);
      if (!sh.scalar()) fail(f->loc, "map arguments must be scalars");
      mk.args.push_back({KernelArg::Scalar, "", "", k});
      Sym s;
      s.k = Sym::Param;
      s.sh = sh;
      s.c = (sh.t == 'f' ? "a.x" : "(" + ctype(sh.t) + ")a.x") + std::to_string(mk.args.size() - 1);
      fn.syms[p.name] = s;
    }
    {
      const Param &p = f->params[mk.req.nextra];
      if (p.type->kind != Type::Element || p.type->elem != fn.elemType)
        fail(f->loc, "argument " + p.name + " must be a " + fn.elemType);
      Sym s;
      s.k = Sym::ElemParam;
      s.inout = p.inout;
      fn.syms[p.name] = s;
    }
    if (mk.card) {
      const Param &p = f->params[mk.req.nextra + 1];
      if (p.type->kind != Type::Tuple || p.type->elem != fn.epType || p.type->tupleLen != mk.card)
        fail(f->loc, "argument " + p.name + " must be (" + fn.epType + "*" +
                         std::to_string(mk.card) + ")");
      Sym s;
      s.k = Sym::TupleParam;
      fn.syms[p.name] = s;
    }
    if (grid) {
      const Param &p = f->params[mk.req.nextra + 1];
      if (p.type->kind != Type::Set || p.type->gridDims != mk.gridDims ||
          p.type->elem != fn.gridType)
        fail(f->loc, "argument " + p.name + " must be grid[" + std::to_string(mk.gridDims) + "]{" +
                         fn.gridType + "}(" + mk.req.target + ")");
      Sym s;
      s.k = Sym::GridParam;
      fn.syms[p.name] = s;
    }
    for (auto &r : f->results) {
      MapOutput mo;
      mo.name = r.name;
      const Type &t = *r.type;
      if (t.kind != Type::Tensor || t.dims.empty() || t.dims.size() > 2)
        fail(f->loc, "result " + r.name + " of a map function must be a vector or matrix over a set");
      mo.matrix = t.dims.size() == 2;
      mo.rowSet = t.dims[0];
      if (mo.matrix && t.dims[1] != t.dims[0])
        fail(f->loc, "matrices between different sets are not supported");
      const std::string want = mk.card ? fn.epSet : fn.elemSet;
      if (mo.rowSet != want)
        fail(f->loc, "result " + r.name + " is indexed by " + mo.rowSet + ", not " + want);
      if (t.block) {
        std::vector<int> bd;
        if (!denseDims(*t.block, &bd)) fail(f->loc, "blocks must have integer dimensions");
        mo.blockDims = bd;
      }
      if (t.comp != "float") fail(f->loc, "results must be float tensors");
      mo.bs = 1;
      for (int d : mo.blockDims) mo.bs *= d;
      if (grid && mo.matrix) {
        if (mo.bs != 1) fail(f->loc, "matrices over a grid hold scalar entries");
        mo.stencil = true;
      }
      const int oi = (int)mk.outputs.size();
      mk.outputs.push_back(mo);
      mk.args.push_back({KernelArg::Stage, "", "", oi});
      if (!mk.req.reduce) mk.args.push_back({KernelArg::Taken, "", "", oi});
      Sym s;
      s.k = Sym::Output;
      s.out = oi;
      fn.syms[r.name] = s;
    }
    fn.nextInst.assign(mk.outputs.size(), 0);
    block(f->body);
    for (size_t o = 0; o < mk.outputs.size(); ++o) mk.outputs[o].ninst = fn.nextInst[o];
    // the writes in execution order
    std::map<std::string, long> vals;
    walk(f->body, vals, mk, fn);
    // the argument struct
    std::ostringstream k;
    const std::string A = "A_" + mk.kernel;
    k << "struct " << A << " {\n";
    for (size_t s = 0; s < mk.args.size(); ++s) {
      const KernelArg &a = mk.args[s];
      std::string t;
      switch (a.kind) {
        case KernelArg::Count: t = "unsigned long long"; break;
        case KernelArg::Endpoints: t = "const int *"; break;
        case KernelArg::Stage: t = "double *"; break;
        case KernelArg::Taken: t = "unsigned char *"; break;
        case KernelArg::Field: t = a.index == 1 ? "int *" : a.index == 2 ? "unsigned char *" : "double *"; break;
        case KernelArg::Scalar: t = "double"; break;
        case KernelArg::GridDim: t = "unsigned long long"; break;
      }
      k << "  " << t << (a.kind == KernelArg::Scalar ? " x" : " a") << s << ";\n";
    }
    k << "};\n";
    std::string body = decls() + fn.body.str();
    for (size_t o = 0; o < mk.outputs.size(); ++o) {
      const std::string tok = "@NINST" + std::to_string(o) + "@";
      const std::string val = std::to_string(mk.outputs[o].ninst);
      for (size_t p = body.find(tok); p != std::string::npos; p = body.find(tok, p))
        body.replace(p, tok.size(), val);
    }
    k << "extern \"C\" __global__ void __launch_bounds__(256) " << mk.kernel << "(" << A
      << " a) {\n"
      << "  const unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + "
         "threadIdx.x;\n"
      << "  if (e >= a.a0) return;\n"
      << body << "}\n";
    F = nullptr;
    return k.str();
  }

  // execution-order walk of the write sites: every branch, constant loops
  void walk(const Block &b, std::map<std::string, long> &vals, MapKernel &mk, Fn &fn) {
    for (auto &s : b) {
      switch (s->kind) {
        case Stmt::Assign: {
          auto it = fn.sites.find(s.get());
          if (it == fn.sites.end()) break;
          const Site &site = it->second;
          long inst = site.base, stride = 1;
          for (size_t k = site.loops.size(); k-- > 0;) {
            inst += (vals.at(site.loops[k].name) - site.loops[k].lo) * stride;
            stride *= site.loops[k].trip;
          }
          MapOutput &mo = mk.outputs[site.out];
          MapOutput::Write w{};
          w.inst = (int)inst;
          int li[2] = {-1, -1};
          for (size_t k = 0; k < site.idx.size(); ++k) {
            const ExprP &x = site.idx[k];
            if (x->kind == Expr::Apply) {
              li[k] = (int)evalConstInt(x->args[1], vals);
            } else if (x->kind == Expr::SetRead) {  // a grid neighbour
              li[k] = -2;
              for (int c = 0; c + 1 < (int)x->args.size() && c < 3; ++c)
                w.off[k][c] = (int)evalConstInt(x->args[1 + c], vals);
              if (mo.stencil) {
                if (k == 0) throw CodegenError(x->loc, "grid matrices are written in their own row");
                for (int c = 0; c < 3; ++c)
                  if (w.off[k][c] < -1 || w.off[k][c] > 1)
                    throw CodegenError(x->loc, "grid matrix entries lie within one step of the "
                                               "row's point (27-point stencil)");
              }
            } else {
              li[k] = -1;  // the element itself
            }
            if (li[k] >= mk.card && mk.card)
              throw CodegenError(x->loc, "endpoint index out of range");
          }
          w.li = li[0];
          w.lj = li[1];
          mo.order.push_back(w);
          break;
        }
        case Stmt::If:
          for (auto &blk : s->blocks) walk(blk, vals, mk, fn);
          break;
        case Stmt::For: {
          if (!s->overSet.empty()) break;
          long lo, hi;
          try {
            lo = evalConstInt(s->lo, {});
            hi = evalConstInt(s->hi, {});
          } catch (const CodegenError &) {
            break;  // no write sites inside (checked when generating)
          }
          const bool had = vals.count(s->name) > 0;
          const long prev = had ? vals[s->name] : 0;
          for (long v = lo; v < hi; ++v) {
            vals[s->name] = v;
            walk(s->body, vals, mk, fn);
          }
          if (had) vals[s->name] = prev;
          else vals.erase(s->name);
          break;
        }
        default: break;
      }
    }
  }
};

}  // namespace

std::string generateModule(const ProgramAST &prog, const std::vector<MapRequest> &reqs,
                           std::vector<MapKernel> *kernels) {
  Gen g(prog);
  return g.module(reqs, kernels);
}

}  // namespace fe
}  // namespace simit
