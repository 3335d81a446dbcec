// Simit-source front-end: syntax tree, parser, element-kernel generator and
// the runtime that executes a program's exported functions on the HIP
// backend.
//
// Reference: Frontend::parseFile / parseString (src/frontend/frontend.h:19-33)
// over the language grammar documented in src/frontend/parser.cpp:13-2074,
// and the lowering of maps (src/lower/lower.cpp:95-198,
// src/lower/lower_maps.cpp:35-120).  The grammar below is the language's;
// the implementation is this backend's own: a recursive-descent parser into
// a small tree, HIP C++ generated for every map's element function (compiled
// at run time for gfx950), and an interpreter for the exported function's
// statements whose tensor operations run as device kernels of the C ABI.
//
// Supported subset: element and extern declarations (sets, edge sets, dense
// extern tensors), constants, functions (generic set parameters, inout
// arguments, several results), statements var / const / assignment (also to
// fields, tensor elements and several targets) / if-elif-else / while /
// for over integer ranges / print / apply, and expressions over scalars,
// dense tensors (literals, slices, transposes, products), system vectors
// and matrices (map ... to ... [reduce +], A*x, x'*y, dot, norm, elementwise
// arithmetic).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace simit {
namespace fe {

struct Loc {
  int line = 0, col = 0;
};

// tensor[d0, d1](block) / vector / matrix / float / int / bool / string,
// an element type, set{E}(endpoints), a tuple (E*n)
struct Type {
  enum Kind { Scalar, Tensor, Element, Set, Tuple, String } kind = Scalar;
  std::string comp = "float";         // component type of Scalar / Tensor
  std::vector<std::string> dims;      // Tensor: index sets (integer text or a set name)
  std::shared_ptr<Type> block;        // Tensor with tensor-typed components
  bool transposed = false;            // a row vector
  std::string elem;                   // Element / Set / Tuple element type
  std::vector<std::string> endpoints;  // Set
  int gridDims = 0;                   // Set: a grid[n]{E}(points) edge set (n)
  int tupleLen = 0;                   // Tuple
  std::string str() const;
};
using TypeP = std::shared_ptr<Type>;

struct Expr;
using ExprP = std::shared_ptr<Expr>;
struct Expr {
  enum Kind {
    Num,        // num (isInt)
    Bool,       // num != 0
    Str,        // name holds the text
    Var,        // name
    Binary,     // op in {+ - * / .* ./ == != < > <= >= and or xor ^}, args[0], args[1]
    Unary,      // op in {- + not}, args[0]
    Transpose,  // args[0]'
    Apply,      // name(args): a call, a tensor read or a tuple read (null arg = ':')
    Field,      // args[0].name
    Literal,    // dense tensor literal: lit (row-major), litDims
    Map,        // map name(args) to target [through set] [reduce +]
    SetRead     // args[0][args[1..split]; args[split+1..]]: a grid neighbour
                // (one offset list) or link (two), offsets from the element
  } kind = Num;
  Loc loc;
  double num = 0.0;
  bool isInt = false;
  std::string op, name;
  std::vector<ExprP> args;
  std::vector<double> lit;
  std::vector<int> litDims;  // [n] (a row vector) or [rows, cols] or deeper
  bool litInt = false;
  std::string target;   // Map
  std::string through;  // Map over a grid: the grid edge set
  bool reduce = false;  // Map
  int split = 0;        // SetRead: offsets in the first list
};

struct Stmt;
using StmtP = std::shared_ptr<Stmt>;
using Block = std::vector<StmtP>;
struct Stmt {
  enum Kind { VarDecl, ConstDecl, Assign, ExprStmt, If, While, For, Print, Apply } kind = Assign;
  Loc loc;
  std::string name;          // VarDecl / ConstDecl / For variable
  TypeP type;                // VarDecl / ConstDecl (optional)
  ExprP init;                // VarDecl / ConstDecl initial value (optional)
  std::vector<ExprP> lhs;    // Assign targets
  ExprP rhs;                 // Assign value / ExprStmt / While condition
  std::vector<ExprP> conds;  // If: condition of each branch
  std::vector<Block> blocks;  // If: block of each branch, then the else block (if any)
  Block body;                // While / For
  ExprP lo, hi;              // For over lo:hi
  std::string overSet;       // For over a set (unsupported at run time)
  std::vector<ExprP> args;   // Print / Apply arguments
  bool newline = false;      // println
  std::string func, target;  // Apply
};

struct Param {
  std::string name;
  TypeP type;
  bool inout = false;
};

struct Func {
  std::string name;
  bool exported = false;
  std::vector<std::string> generics;
  std::vector<Param> params;
  std::vector<Param> results;
  Block body;
  Loc loc;
};

struct ElementDecl {
  std::string name;
  std::vector<std::pair<std::string, TypeP>> fields;
};

struct ProgramAST {
  std::vector<ElementDecl> elements;
  std::vector<std::pair<std::string, TypeP>> externs;  // in declaration order
  std::vector<StmtP> consts;                            // global const declarations
  std::vector<std::shared_ptr<Func>> funcs;
  const ElementDecl *element(const std::string &n) const;
  const Func *func(const std::string &n) const;
  TypeP externType(const std::string &n) const;
};

struct ParseError {
  Loc loc;
  std::string msg;
  std::string str() const;
};

// Parses Simit source; returns the program, or null with the errors.
std::shared_ptr<ProgramAST> parse(const std::string &text, std::vector<ParseError> *errors);

// Number of components of a dense tensor type with integer dimensions (and
// its integer dimensions); false if a dimension is a set.
bool denseDims(const Type &t, std::vector<int> *dims);

}  // namespace fe
}  // namespace simit
