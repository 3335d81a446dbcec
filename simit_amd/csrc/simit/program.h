// simit::init / Program / Function work-alike for the HIP backend.
//
// Reference interfaces: simit::init(Settings) (src/init.h:18-46), Program
// (src/program.h, program.cpp:30-107), Function (src/function.h:30-94,
// function.cpp:22-139) and the backend plugin contract backend::Function
// (src/backend/backend_function.h:27-88).  The reference compiles Simit source
// to LLVM.  Program::loadFile parses Simit source (frontend.h): each map's
// element function becomes a generated HIP kernel compiled for gfx950 at
// init, and the exported function's statements run over the C ABI.  The
// hot-path programs are also available as fixed pipelines of hand-written
// kernels (Program::loadBuiltin), by name:
//
//   "isprings"         test/input/program/isprings.sim main (implicit springs
//                      step: assembly of MD + K, inline CG, x += h v)
//   "springs_implicit" the same step with A = M - dt^2 K, gravity -9.8 z,
//                      relative 1e-8 CG (the 20^3 lattice workload)
//   "cg"               test/input/program/cg.sim main
//   "fem_static"       linear-elastic tets (fem_linear.sim element), clamped
//                      rows, body load, mixed-precision block-Jacobi PCG
//   "thermal_hex"      scalar trilinear hex Poisson, Dirichlet rows, PCG
//   "femTet"           test/input/program/femTet.sim: initializeTet (rest
//                      shape B, volume W, vertex masses) and main (neo-Hookean
//                      stiffness and force, inline CG, position update)
//   "extern_neg"       A = -A on an extern int matrix bound as CSR TensorData
//                      (test/function-tests.cpp:96-139)
//   "extern_spmv", "extern_spmv_3x3"
//                      y = A x, A an extern CSR matrix of 1x1 / 3x3 double
//                      blocks over the extern set V, x and y extern vectors
//   "extern_cg", "extern_cg_3x3"
//                      x = cg(A, b) (relative 1e-8, (block) Jacobi, FP64)
//
// The bind / init / run / runSafe / mapArgs / unmapArgs contract is the
// reference's: bind borrows the host sets, init uploads them and builds the
// sparsity index on the device, unmapArgs pushes host data to the device,
// mapArgs pulls results back (the reference's naming, springs.cpp:107-109).
#pragma once

#include <functional>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "graph.h"

namespace simit {

// Settings (init.h:18-22).  backend "hip" (or "gpu", routed here).
struct Settings {
  std::string backend = "hip";
  int floatSize = 8;
  bool indexlessStencils = false;
};
void init(const Settings &settings);
void init(const std::string &backend = "hip", int floatSize = 8);
const std::string &currentBackend();

// Solve options of the CG that a program runs (the values the .sim sources
// hard-code are the defaults of each program).
struct SolverOptions {
  int maxiters = -1;     // -1: the program's own value
  double tol = -1.0;     // -1: the program's own value
  int precond = -1;      // SMH_PRECOND_*; -1: the program's own choice
  int mixed = -1;        // 0/1; -1: the program's own choice
};

struct SolveStats {
  int iters = 0;
  int reliable_updates = 0;
  int status = 0;
  double true_rel = 0.0;
  double seconds = 0.0;
  int precond_used = 0;  // the rung of the fallback chain that produced x
  int fallbacks = 0;
};

// tensor_type.h: the component types of bound tensors, and a tensor's type
// (dimensions of its components, or of blocks of a block type)
enum class ComponentType { Float, Double, Int, Boolean };
inline size_t componentSize(ComponentType t) {
  switch (t) {
    case ComponentType::Float: return sizeof(float);
    case ComponentType::Double: return sizeof(double);
    case ComponentType::Int: return sizeof(int);
    case ComponentType::Boolean: return sizeof(bool);
  }
  return 0;
}
class TensorType {
 public:
  TensorType(ComponentType componentType, std::vector<int> dimensions = {})
      : componentType(componentType), dimensions(std::move(dimensions)) {}
  TensorType(const TensorType &blockType, std::vector<int> dimensions)
      : blocked(true), componentType(blockType.componentType), dimensions(std::move(dimensions)),
        blockDimensions(blockType.dimensions) {}
  bool isBlocked() const { return blocked; }
  ComponentType getComponentType() const { return componentType; }
  const std::vector<int> &getDimensions() const { return dimensions; }
  const std::vector<int> &getBlockDimensions() const { return blockDimensions; }
  size_t getSize() const;  // components

 private:
  bool blocked = false;
  ComponentType componentType;
  std::vector<int> dimensions, blockDimensions;
};

// tensor_data.h:7-50: sparse tensor data in CSR form bound to an extern
// matrix.  rowLen ints may be read from rowPtr, dataLen from colInd and
// dataLen blocks (each the extern's block of components) from data.
class TensorData {
 public:
  enum Kind { Sparse };
  TensorData(const int *rowPtr, const int *colInd, void *data, int rowLen, int dataLen)
      : kind(Sparse), rowPtr(rowPtr), colInd(colInd), data(data), rowLen(rowLen),
        dataLen(dataLen) {}
  TensorData(TensorData &td) = default;
  Kind getKind() { return kind; }
  void *getData() { return data; }
  const int *getRowPtr() { return rowPtr; }
  const int *getColInd() { return colInd; }
  int getRowLen() { return rowLen; }
  int getDataLen() { return dataLen; }

 private:
  Kind kind;
  const int *rowPtr;
  const int *colInd;
  void *data;
  int rowLen;
  int dataLen;
};

namespace backend {
// The backend plugin contract, backend_function.h:27-88: the virtuals are
// the reference's; the bindable metadata comes from the base.
class Function {
 public:
  typedef std::function<void()> FuncType;
  virtual ~Function() = default;
  Function(const Function &) = delete;
  Function &operator=(const Function &) = delete;

  /// Bind the given set to the set with the given name.
  virtual void bind(const std::string &name, Set *set) = 0;
  /// Bind the given data to the tensor with the given name.
  virtual void bind(const std::string &name, void *data) = 0;
  /// Bind the given data and indices to the sparse tensor with the given name.
  virtual void bind(const std::string &name, TensorData &data) = 0;
  /// Initialize the function.
  virtual FuncType init() = 0;
  /// Query whether the function requires initialization.
  virtual bool isInitialized() = 0;
  virtual void mapArgs() {}
  virtual void unmapArgs(bool updated = true) { (void)updated; }
  /// The function's pipeline: its bindables and the library calls of a run.
  virtual void print(std::ostream &os) const = 0;
  /// The machine it runs on and the gfx950 kernels a run launches.
  virtual void printMachine(std::ostream &os) const = 0;

  struct Bindable {
    std::string name;
    std::string kind;  // "set", "matrix", "vector" or "scalar"
  };
  bool hasBindable(const std::string &name) const;
  const std::vector<Bindable> &getBindables() const { return bindables; }

 protected:
  Function() = default;
  std::vector<Bindable> bindables;
};
}  // namespace backend

// HIP-specific controls of the functions this backend compiles (not part of
// the plugin contract): the CG options of a program's solve and its stats.
namespace hip {
class SolverControl {
 public:
  virtual ~SolverControl() = default;
  virtual void setSolverOptions(const SolverOptions &o) = 0;
  virtual SolveStats lastSolve() const = 0;
};
}  // namespace hip

// function.h:30-94
class Function {
 public:
  Function() = default;
  explicit Function(std::shared_ptr<backend::Function> impl) : impl(std::move(impl)) {}
  void clear() {
    impl.reset();
    funcPtr = nullptr;
  }
  void bind(const std::string &name, Set *set);
  /// Bind tensor data to the bindable with type checks.
  void bind(const std::string &name, const TensorType &ttype, void *data);
  void bind(const std::string &name, void *data);
  /// Bind sparse tensor data in the CSR format to the bindable.
  void bind(const std::string &name, TensorData &data);
  void init();
  inline void run() { funcPtr(); }
  void runSafe();
  void mapArgs();
  void unmapArgs(bool updated = true);
  bool defined() const { return impl != nullptr; }
  bool isInitialized() const { return impl && impl->isInitialized(); }
  void print(std::ostream &os) const;
  void printMachine(std::ostream &os) const;
  // HIP backend extensions (hip::SolverControl)
  void setSolverOptions(const SolverOptions &o);
  SolveStats lastSolve() const;

 private:
  std::shared_ptr<backend::Function> impl;
  std::function<void()> funcPtr;
};
std::ostream &operator<<(std::ostream &os, const Function &f);

namespace fe {
struct ProgramAST;
}

// Program (program.h:18-60, program.cpp:68-107): loadFile / loadString parse
// Simit source with the front-end (frontend.h); compile(f) returns f as a
// function of this backend: its maps as generated gfx950 kernels, its
// statements run over the C ABI.  Status: 0 ok, 1 parse errors (see
// getDiagnostics), 2 unreadable file.  loadBuiltin selects one of the fixed
// pipelines listed above instead (hand-written kernels end to end).
class Program {
 public:
  int loadFile(const std::string &path);
  int loadString(const std::string &text);
  int loadBuiltin(const std::string &name);
  std::vector<std::string> getFunctionNames() const;
  Function compile(const std::string &function);
  const std::string &getDiagnostics() const { return diagnostics; }
  // HIP backend extensions: the exported functions of a parsed program, and
  // the generated HIP source of a function's maps
  std::vector<std::string> getExportedFunctionNames() const;
  std::string getGeneratedSource(const std::string &function) const;

 private:
  std::string program;
  std::shared_ptr<const fe::ProgramAST> ast;
  std::string diagnostics;
};

}  // namespace simit
