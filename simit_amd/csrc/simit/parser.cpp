// Lexer and recursive-descent parser for the Simit language subset (see
// frontend.h).  The grammar is the language's (src/frontend/parser.cpp:13-2074
// documents it in EBNF): precedence from loosest to tightest is map, or, and,
// xor, comparisons, not, + -, * / .* ./, unary + -, ^, transpose ', then
// calls / reads / field accesses.
#include <cctype>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

#include "frontend.h"

namespace simit {
namespace fe {

std::string Type::str() const {
  switch (kind) {
    case Scalar: return comp;
    case String: return "string";
    case Element: return elem;
    case Tuple: return "(" + elem + "*" + std::to_string(tupleLen) + ")";
    case Set: {
      if (gridDims)
        return "grid[" + std::to_string(gridDims) + "]{" + elem + "}(" + endpoints[0] + ")";
      std::string s = "set{" + elem + "}";
      if (!endpoints.empty()) {
        s += "(";
        for (size_t i = 0; i < endpoints.size(); ++i) s += (i ? "," : "") + endpoints[i];
        s += ")";
      }
      return s;
    }
    case Tensor: {
      std::string s = "tensor[";
      for (size_t i = 0; i < dims.size(); ++i) s += (i ? "," : "") + dims[i];
      s += "](" + (block ? block->str() : comp) + ")";
      if (transposed) s += "'";
      return s;
    }
  }
  return "?";
}

std::string ParseError::str() const {
  return std::to_string(loc.line) + ":" + std::to_string(loc.col) + ": " + msg;
}

const ElementDecl *ProgramAST::element(const std::string &n) const {
  for (auto &e : elements)
    if (e.name == n) return &e;
  return nullptr;
}
const Func *ProgramAST::func(const std::string &n) const {
  for (auto &f : funcs)
    if (f->name == n) return f.get();
  return nullptr;
}
TypeP ProgramAST::externType(const std::string &n) const {
  for (auto &e : externs)
    if (e.first == n) return e.second;
  return nullptr;
}

bool denseDims(const Type &t, std::vector<int> *dims) {
  dims->clear();
  if (t.kind == Type::Scalar) return true;
  if (t.kind != Type::Tensor) return false;
  for (auto &d : t.dims) {
    if (d.empty() || !std::isdigit((unsigned char)d[0])) return false;
    dims->push_back(std::atoi(d.c_str()));
  }
  if (t.block) {
    std::vector<int> b;
    if (!denseDims(*t.block, &b)) return false;
    dims->insert(dims->end(), b.begin(), b.end());
  }
  return true;
}

namespace {

struct Tok {
  enum K { Ident, Int, Float, Str, Op, End } k = End;
  std::string s;
  double v = 0.0;
  Loc loc;
};

struct SyntaxError : std::runtime_error {
  Loc loc;
  SyntaxError(Loc l, const std::string &m) : std::runtime_error(m), loc(l) {}
};

std::vector<Tok> lex(const std::string &src) {
  std::vector<Tok> out;
  int line = 1, col = 1;
  size_t i = 0;
  auto adv = [&](size_t n = 1) {
    for (size_t k = 0; k < n && i < src.size(); ++k, ++i) {
      if (src[i] == '\n') {
        ++line;
        col = 1;
      } else {
        ++col;
      }
    }
  };
  while (i < src.size()) {
    const char c = src[i];
    if (std::isspace((unsigned char)c)) {
      adv();
      continue;
    }
    if (c == '%') {
      if (i + 1 < src.size() && src[i + 1] == '{') {  // %{ block comment %}
        const size_t e = src.find("%}", i + 2);
        if (e == std::string::npos) throw SyntaxError({line, col}, "unterminated %{ comment");
        adv(e + 2 - i);
      } else {
        while (i < src.size() && src[i] != '\n') adv();
      }
      continue;
    }
    Tok t;
    t.loc = {line, col};
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i;
      while (j < src.size() && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
      t.k = Tok::Ident;
      t.s = src.substr(i, j - i);
      adv(j - i);
      out.push_back(t);
      continue;
    }
    if (std::isdigit((unsigned char)c) ||
        (c == '.' && i + 1 < src.size() && std::isdigit((unsigned char)src[i + 1]))) {
      size_t j = i;
      bool isFloat = false;
      while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
      if (j < src.size() && src[j] == '.' &&
          !(j + 1 < src.size() && (src[j + 1] == '*' || src[j + 1] == '/'))) {
        isFloat = true;
        ++j;
        while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
      }
      if (j < src.size() && (src[j] == 'e' || src[j] == 'E')) {
        size_t k = j + 1;
        if (k < src.size() && (src[k] == '+' || src[k] == '-')) ++k;
        if (k < src.size() && std::isdigit((unsigned char)src[k])) {
          isFloat = true;
          j = k;
          while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
        }
      }
      t.s = src.substr(i, j - i);
      t.k = isFloat ? Tok::Float : Tok::Int;
      t.v = std::strtod(t.s.c_str(), nullptr);
      adv(j - i);
      out.push_back(t);
      continue;
    }
    if (c == '"') {
      std::string s;
      size_t j = i + 1;
      while (j < src.size() && src[j] != '"') {
        if (src[j] == '\\' && j + 1 < src.size()) {
          const char e = src[j + 1];
          s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
          j += 2;
        } else {
          s += src[j++];
        }
      }
      if (j >= src.size()) throw SyntaxError(t.loc, "unterminated string");
      t.k = Tok::Str;
      t.s = s;
      adv(j + 1 - i);
      out.push_back(t);
      continue;
    }
    static const char *two[] = {"->", "==", "!=", "<=", ">=", ".*", "./"};
    bool matched = false;
    for (const char *op : two)
      if (src.compare(i, 2, op) == 0) {
        t.k = Tok::Op;
        t.s = op;
        adv(2);
        matched = true;
        break;
      }
    if (!matched) {
      if (std::string("+-*/\\^'()[]{},;:=<>.").find(c) == std::string::npos)
        throw SyntaxError(t.loc, std::string("unexpected character '") + c + "'");
      t.k = Tok::Op;
      t.s = std::string(1, c);
      adv();
    }
    out.push_back(t);
  }
  Tok e;
  e.k = Tok::End;
  e.loc = {line, col};
  out.push_back(e);
  return out;
}

class Parser {
 public:
  explicit Parser(std::vector<Tok> toks) : t(std::move(toks)) {}

  std::shared_ptr<ProgramAST> program() {
    auto p = std::make_shared<ProgramAST>();
    while (!atEnd()) {
      if (isKw("element")) {
        p->elements.push_back(elementDecl());
      } else if (isKw("extern")) {
        next();
        if (isKw("func")) throw err("extern functions are not supported");
        const std::string n = ident();
        expect(":");
        TypeP ty = type();
        expect(";");
        p->externs.push_back({n, ty});
      } else if (isKw("const")) {
        p->consts.push_back(constDecl());
      } else if (isKw("export") || isKw("func")) {
        p->funcs.push_back(funcDecl());
      } else {
        throw err("expected element, extern, const or func, found '" + cur().s + "'");
      }
    }
    return p;
  }

 private:
  std::vector<Tok> t;
  size_t i = 0;

  const Tok &cur() const { return t[i]; }
  const Tok &peek(size_t k = 1) const { return t[std::min(i + k, t.size() - 1)]; }
  bool atEnd() const { return cur().k == Tok::End; }
  Tok next() { return t[i < t.size() - 1 ? i++ : i]; }
  bool isOp(const char *s) const { return cur().k == Tok::Op && cur().s == s; }
  bool isKw(const char *s) const { return cur().k == Tok::Ident && cur().s == s; }
  SyntaxError err(const std::string &m) const { return SyntaxError(cur().loc, m); }
  void expect(const char *s) {
    if (!isOp(s)) throw err(std::string("expected '") + s + "', found '" + cur().s + "'");
    next();
  }
  void expectKw(const char *s) {
    if (!isKw(s)) throw err(std::string("expected '") + s + "', found '" + cur().s + "'");
    next();
  }
  bool accept(const char *s) {
    if (isOp(s)) {
      next();
      return true;
    }
    return false;
  }
  static bool reserved(const std::string &s) {
    static const char *kw[] = {"element", "extern", "const", "export", "func", "end", "var",
                               "if", "elif", "else", "while", "for", "in", "map", "to",
                               "reduce", "through", "apply", "print", "println", "and", "or",
                               "not", "xor", "true", "false", "inout", "do", "return"};
    for (const char *k : kw)
      if (s == k) return true;
    return false;
  }
  std::string ident() {
    if (cur().k != Tok::Ident || reserved(cur().s))
      throw err("expected an identifier, found '" + cur().s + "'");
    return next().s;
  }

  ElementDecl elementDecl() {
    expectKw("element");
    ElementDecl e;
    e.name = ident();
    while (!isKw("end")) {
      const std::string f = ident();
      expect(":");
      TypeP ty = type();
      expect(";");
      e.fields.push_back({f, ty});
    }
    expectKw("end");
    return e;
  }

  StmtP constDecl() {
    auto s = std::make_shared<Stmt>();
    s->loc = cur().loc;
    expectKw("const");
    s->kind = Stmt::ConstDecl;
    s->name = ident();
    if (accept(":")) s->type = type();
    expect("=");
    s->init = expr();
    expect(";");
    return s;
  }

  std::shared_ptr<Func> funcDecl() {
    auto f = std::make_shared<Func>();
    f->loc = cur().loc;
    if (isKw("export")) {
      next();
      f->exported = true;
    }
    expectKw("func");
    f->name = ident();
    if (accept("<")) {
      do {
        if (cur().k == Tok::Int) {  // [0:] ident
          next();
          expect(":");
        }
        f->generics.push_back(ident());
      } while (accept(","));
      expect(">");
    }
    expect("(");
    if (!isOp(")")) {
      do {
        Param p;
        if (isKw("inout")) {
          next();
          p.inout = true;
        }
        p.name = ident();
        expect(":");
        p.type = type();
        f->params.push_back(p);
      } while (accept(","));
    }
    expect(")");
    if (accept("->")) {
      if (accept("(")) {
        do {
          Param p;
          p.name = ident();
          expect(":");
          p.type = type();
          f->results.push_back(p);
        } while (accept(","));
        expect(")");
      } else {
        Param p;
        p.name = ident();
        expect(":");
        p.type = type();
        f->results.push_back(p);
      }
    }
    f->body = block({"end"});
    expectKw("end");
    return f;
  }

  // statements up to (not including) one of the terminating keywords
  Block block(std::initializer_list<const char *> ends) {
    Block b;
    for (;;) {
      if (atEnd()) throw err("unexpected end of input (missing 'end'?)");
      bool stop = false;
      for (const char *e : ends) stop |= isKw(e);
      if (stop) return b;
      b.push_back(stmt());
    }
  }

  StmtP stmt() {
    auto s = std::make_shared<Stmt>();
    s->loc = cur().loc;
    if (isKw("var")) {
      next();
      s->kind = Stmt::VarDecl;
      s->name = ident();
      if (accept(":")) s->type = type();
      if (accept("=")) s->init = expr();
      if (!s->type && !s->init) throw err("a var needs a type or an initial value");
      expect(";");
      return s;
    }
    if (isKw("const")) return constDecl();
    if (isKw("if")) {
      next();
      s->kind = Stmt::If;
      s->conds.push_back(expr());
      s->blocks.push_back(block({"elif", "else", "end"}));
      for (;;) {
        if (isKw("elif")) {
          next();
          s->conds.push_back(expr());
          s->blocks.push_back(block({"elif", "else", "end"}));
          continue;
        }
        if (isKw("else")) {
          next();
          s->blocks.push_back(block({"end"}));
        }
        break;
      }
      expectKw("end");
      return s;
    }
    if (isKw("while")) {
      next();
      s->kind = Stmt::While;
      s->rhs = expr();
      s->body = block({"end"});
      expectKw("end");
      return s;
    }
    if (isKw("for")) {
      next();
      s->kind = Stmt::For;
      s->name = ident();
      expectKw("in");
      ExprP a = expr();
      if (accept(":")) {
        s->lo = a;
        s->hi = expr();
      } else if (a->kind == Expr::Var) {
        s->overSet = a->name;
      } else {
        throw err("expected a set or a range lo:hi after 'in'");
      }
      s->body = block({"end"});
      expectKw("end");
      return s;
    }
    if (isKw("print") || isKw("println")) {
      s->kind = Stmt::Print;
      s->newline = next().s == "println";
      do s->args.push_back(expr());
      while (accept(","));
      expect(";");
      return s;
    }
    if (isKw("apply")) {
      next();
      s->kind = Stmt::Apply;
      s->func = ident();
      if (accept("(")) {
        if (!isOp(")")) do s->args.push_back(expr());
          while (accept(","));
        expect(")");
      }
      expectKw("to");
      s->target = ident();
      expect(";");
      return s;
    }
    std::vector<ExprP> es;
    es.push_back(expr());
    while (accept(",")) es.push_back(expr());
    if (accept("=")) {
      s->kind = Stmt::Assign;
      s->lhs = es;
      s->rhs = expr();
    } else {
      if (es.size() != 1) throw err("expected '='");
      s->kind = Stmt::ExprStmt;
      s->rhs = es[0];
    }
    expect(";");
    return s;
  }

  // ---- types ----
  TypeP type() {
    auto ty = std::make_shared<Type>();
    if (isKw("float") || isKw("int") || isKw("bool")) {
      ty->kind = Type::Scalar;
      ty->comp = next().s;
      return ty;
    }
    if (isKw("string")) {
      next();
      ty->kind = Type::String;
      return ty;
    }
    if (isKw("complex") || isKw("opaque")) throw err("type '" + cur().s + "' is not supported");
    if (isKw("grid")) {  // grid[n]{E}(points)
      next();
      ty->kind = Type::Set;
      expect("[");
      if (cur().k != Tok::Int) throw err("expected the grid's dimension count");
      ty->gridDims = (int)next().v;
      if (ty->gridDims < 1 || ty->gridDims > 3) throw err("grids have 1 to 3 dimensions");
      expect("]");
      expect("{");
      ty->elem = ident();
      expect("}");
      expect("(");
      const std::string pts = ident();
      expect(")");
      ty->endpoints = {pts, pts};
      return ty;
    }
    if (isKw("tensor") || isKw("vector") || isKw("matrix")) {
      const std::string k = next().s;
      ty->kind = Type::Tensor;
      if (accept("[")) {
        do {
          if (cur().k == Tok::Int) ty->dims.push_back(next().s);
          else ty->dims.push_back(ident());
        } while (accept(","));
        expect("]");
      }
      if ((k == "vector" && ty->dims.size() > 1) || (k == "matrix" && ty->dims.size() != 2 &&
                                                     !ty->dims.empty()))
        throw err(k + " has the wrong number of dimensions");
      expect("(");
      if (isKw("float") || isKw("int") || isKw("bool")) {
        ty->comp = next().s;
      } else {
        ty->block = type();
        if (ty->block->kind != Type::Tensor) throw err("tensor blocks must be tensors");
        ty->comp = ty->block->comp;
      }
      expect(")");
      if (accept("'")) ty->transposed = true;
      return ty;
    }
    if (isKw("set")) {
      next();
      ty->kind = Type::Set;
      expect("{");
      ty->elem = ident();
      expect("}");
      if (accept("(")) {
        const std::string first = ident();
        if (accept("*")) {
          if (cur().k != Tok::Int) throw err("expected an endpoint count");
          const int n = (int)next().v;
          for (int k = 0; k < n; ++k) ty->endpoints.push_back(first);
        } else {
          ty->endpoints.push_back(first);
          while (accept(",")) ty->endpoints.push_back(ident());
        }
        expect(")");
      }
      return ty;
    }
    if (accept("(")) {
      ty->kind = Type::Tuple;
      ty->elem = ident();
      expect("*");
      if (cur().k != Tok::Int) throw err("expected a tuple length");
      ty->tupleLen = (int)next().v;
      expect(")");
      return ty;
    }
    ty->kind = Type::Element;
    ty->elem = ident();
    return ty;
  }

  // ---- expressions ----
  ExprP mk(Expr::Kind k, Loc l) {
    auto e = std::make_shared<Expr>();
    e->kind = k;
    e->loc = l;
    return e;
  }
  ExprP bin(const std::string &op, ExprP a, ExprP b, Loc l) {
    ExprP e = mk(Expr::Binary, l);
    e->op = op;
    e->args = {a, b};
    return e;
  }

  ExprP expr() {
    if (isKw("map")) return mapExpr();
    return orExpr();
  }
  ExprP mapExpr() {
    ExprP e = mk(Expr::Map, cur().loc);
    expectKw("map");
    e->name = ident();
    if (isOp("<")) throw err("map endpoints <...> are not supported");
    if (accept("(")) {
      if (!isOp(")")) do e->args.push_back(expr());
        while (accept(","));
      expect(")");
    }
    expectKw("to");
    e->target = ident();
    if (isKw("through")) {
      next();
      e->through = ident();
    }
    if (isKw("reduce")) {
      next();
      expect("+");
      e->reduce = true;
    }
    return e;
  }
  ExprP orExpr() {
    ExprP a = andExpr();
    while (isKw("or")) {
      const Loc l = next().loc;
      a = bin("or", a, andExpr(), l);
    }
    return a;
  }
  ExprP andExpr() {
    ExprP a = xorExpr();
    while (isKw("and")) {
      const Loc l = next().loc;
      a = bin("and", a, xorExpr(), l);
    }
    return a;
  }
  ExprP xorExpr() {
    ExprP a = eqExpr();
    while (isKw("xor")) {
      const Loc l = next().loc;
      a = bin("xor", a, eqExpr(), l);
    }
    return a;
  }
  ExprP eqExpr() {
    ExprP a = term();
    while (isOp("==") || isOp("!=") || isOp("<") || isOp(">") || isOp("<=") || isOp(">=")) {
      const Tok op = next();
      a = bin(op.s, a, term(), op.loc);
    }
    return a;
  }
  ExprP term() {
    if (isKw("not")) {
      ExprP e = mk(Expr::Unary, next().loc);
      e->op = "not";
      e->args = {term()};
      return e;
    }
    return addExpr();
  }
  ExprP addExpr() {
    ExprP a = mulExpr();
    while (isOp("+") || isOp("-")) {
      const Tok op = next();
      a = bin(op.s, a, mulExpr(), op.loc);
    }
    return a;
  }
  ExprP mulExpr() {
    ExprP a = negExpr();
    while (isOp("*") || isOp("/") || isOp(".*") || isOp("./") || isOp("\\")) {
      const Tok op = next();
      if (op.s == "\\") throw SyntaxError(op.loc, "the \\ (solve) operator is not supported");
      a = bin(op.s, a, negExpr(), op.loc);
    }
    return a;
  }
  ExprP negExpr() {
    if (isOp("-") || isOp("+")) {
      const Tok op = next();
      ExprP e = mk(Expr::Unary, op.loc);
      e->op = op.s;
      e->args = {negExpr()};
      return e;
    }
    return expExpr();
  }
  ExprP expExpr() {
    ExprP a = postfix();
    if (isOp("^")) {
      const Tok op = next();
      return bin("^", a, expExpr(), op.loc);
    }
    return a;
  }
  ExprP postfix() {
    ExprP a = factor();
    for (;;) {
      if (isOp("(")) {
        if (a->kind != Expr::Var && a->kind != Expr::Field && a->kind != Expr::Apply)
          throw err("only names, fields and reads can be indexed");
        ExprP e = mk(Expr::Apply, cur().loc);
        next();
        e->args.push_back(a);  // args[0] is the callee / tensor
        if (!isOp(")")) {
          do {
            if (isOp(":") && (peek().k == Tok::Op && (peek().s == "," || peek().s == ")"))) {
              next();
              e->args.push_back(nullptr);
            } else {
              e->args.push_back(expr());
            }
          } while (accept(","));
        }
        expect(")");
        a = e;
      } else if (isOp("[") && a->kind == Expr::Var) {  // grid neighbour / link
        ExprP e = mk(Expr::SetRead, cur().loc);
        next();
        e->args.push_back(a);
        if (!isOp("]") && !isOp(";")) do e->args.push_back(expr());
          while (accept(","));
        e->split = (int)e->args.size() - 1;
        if (accept(";")) {
          do e->args.push_back(expr());
          while (accept(","));
        }
        expect("]");
        a = e;
      } else if (isOp(".") && peek().k == Tok::Ident) {
        const Loc l = next().loc;
        ExprP e = mk(Expr::Field, l);
        e->name = ident();
        e->args = {a};
        a = e;
      } else if (isOp("'")) {
        ExprP e = mk(Expr::Transpose, next().loc);
        e->args = {a};
        a = e;
      } else {
        return a;
      }
    }
  }
  ExprP factor() {
    const Tok &c = cur();
    if (isOp("(")) {
      next();
      ExprP e = expr();
      expect(")");
      return e;
    }
    if (c.k == Tok::Int || c.k == Tok::Float) {
      ExprP e = mk(Expr::Num, c.loc);
      e->num = c.v;
      e->isInt = c.k == Tok::Int;
      next();
      return e;
    }
    if (c.k == Tok::Str) {
      ExprP e = mk(Expr::Str, c.loc);
      e->name = c.s;
      next();
      return e;
    }
    if (isKw("true") || isKw("false")) {
      ExprP e = mk(Expr::Bool, c.loc);
      e->num = next().s == "true" ? 1.0 : 0.0;
      return e;
    }
    if (isOp("[")) return literal();
    if (c.k == Tok::Ident && !reserved(c.s)) {
      ExprP e = mk(Expr::Var, c.loc);
      e->name = next().s;
      return e;
    }
    throw err("expected an expression, found '" + c.s + "'");
  }

  // dense tensor literal: rows of signed numbers separated by ';', commas
  // optional; nested brackets add a dimension
  ExprP literal() {
    ExprP e = mk(Expr::Literal, cur().loc);
    expect("[");
    bool allInt = true;
    if (isOp("[")) {  // nested: [[...], [...]]
      std::vector<ExprP> subs;
      do {
        subs.push_back(literal());
        accept(",");
      } while (isOp("["));
      expect("]");
      for (auto &s : subs) {
        if (s->litDims != subs[0]->litDims) throw err("ragged tensor literal");
        e->lit.insert(e->lit.end(), s->lit.begin(), s->lit.end());
        allInt &= s->litInt;
      }
      e->litDims.push_back((int)subs.size());
      e->litDims.insert(e->litDims.end(), subs[0]->litDims.begin(), subs[0]->litDims.end());
      e->litInt = allInt;
      return e;
    }
    std::vector<std::vector<double>> rows(1);
    for (;;) {
      if (isOp("]")) break;
      if (accept(";")) {
        rows.emplace_back();
        continue;
      }
      if (accept(",")) continue;
      double sign = 1.0;
      if (isOp("-") || isOp("+")) sign = next().s == "-" ? -1.0 : 1.0;
      if (cur().k != Tok::Int && cur().k != Tok::Float)
        throw err("tensor literals hold numbers only");
      allInt &= cur().k == Tok::Int;
      rows.back().push_back(sign * next().v);
    }
    expect("]");
    for (auto &r : rows)
      if (r.size() != rows[0].size() || r.empty()) throw err("ragged tensor literal");
    for (auto &r : rows) e->lit.insert(e->lit.end(), r.begin(), r.end());
    if (rows.size() == 1) e->litDims = {(int)rows[0].size()};
    else e->litDims = {(int)rows.size(), (int)rows[0].size()};
    e->litInt = allInt;
    return e;
  }
};

}  // namespace

std::shared_ptr<ProgramAST> parse(const std::string &text, std::vector<ParseError> *errors) {
  try {
    Parser p(lex(text));
    return p.program();
  } catch (const SyntaxError &e) {
    if (errors) errors->push_back({e.loc, e.what()});
    return nullptr;
  }
}

}  // namespace fe
}  // namespace simit
