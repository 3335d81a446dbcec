// simit::init / Program / Function work-alike for the HIP backend.
//
// Reference interfaces: simit::init(Settings) (src/init.h:18-46), Program
// (src/program.h, program.cpp:30-107), Function (src/function.h:30-94,
// function.cpp:22-139) and the backend plugin contract backend::Function
// (src/backend/backend_function.h:27-88).  The reference compiles Simit source
// to LLVM; this backend has no front-end (out of scope) and instead runs the
// hot-path programs as fixed HIP pipelines, looked up by name:
//
//   "isprings"         test/input/program/isprings.sim main (implicit springs
//                      step: assembly of MD + K, inline CG, x += h v)
//   "springs_implicit" the same step with A = M - dt^2 K, gravity -9.8 z,
//                      relative 1e-8 CG (the 20^3 lattice workload)
//   "cg"               test/input/program/cg.sim main
//   "fem_static"       linear-elastic tets (fem_linear.sim element), clamped
//                      rows, body load, mixed-precision block-Jacobi PCG
//   "thermal_hex"      scalar trilinear hex Poisson, Dirichlet rows, PCG
//
// The bind / init / run / runSafe / mapArgs / unmapArgs contract is the
// reference's: bind borrows the host sets, init uploads them and builds the
// sparsity index on the device, unmapArgs pushes host data to the device,
// mapArgs pulls results back (the reference's naming, springs.cpp:107-109).
#pragma once

#include <functional>
#include <memory>
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

namespace backend {
// backend_function.h:27-88
class Function {
 public:
  virtual ~Function() = default;
  virtual void bind(const std::string &name, Set *set) = 0;
  virtual void bind(const std::string &name, void *data) = 0;
  virtual std::function<void()> init() = 0;
  virtual bool isInitialized() const = 0;
  virtual void mapArgs() = 0;
  virtual void unmapArgs(bool updated = true) = 0;
  virtual void setSolverOptions(const SolverOptions &o) = 0;
  virtual SolveStats lastSolve() const = 0;
  virtual std::vector<std::string> externs() const = 0;
};
}  // namespace backend

class Function {
 public:
  Function() = default;
  explicit Function(std::shared_ptr<backend::Function> impl) : impl(std::move(impl)) {}
  void bind(const std::string &name, Set *set);
  void bind(const std::string &name, void *data);
  void init();
  inline void run() { funcPtr(); }
  void runSafe();
  void mapArgs();
  void unmapArgs(bool updated = true);
  bool defined() const { return impl != nullptr; }
  bool isInitialized() const { return impl && impl->isInitialized(); }
  void setSolverOptions(const SolverOptions &o);
  SolveStats lastSolve() const;

 private:
  std::shared_ptr<backend::Function> impl;
  std::function<void()> funcPtr;
};

class Program {
 public:
  // 0 ok, 1 errors, 2 unreadable (program.h:31-41).  A file is accepted when
  // its name is one of the programs above (isprings.sim, cg.sim, ...).
  int loadFile(const std::string &path);
  int loadBuiltin(const std::string &name);
  std::vector<std::string> getFunctionNames() const;
  Function compile(const std::string &function);

 private:
  std::string program;
};

}  // namespace simit
