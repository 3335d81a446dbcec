// Element-kernel generation for the maps of a parsed Simit program.
//
// Each `map f(args) to E [reduce +]` becomes one HIP kernel with a thread per
// element of E that runs f's body.  Writes into f's results (global vectors
// and matrices indexed by the element or its endpoints) are not scattered
// into the global tensor by the kernel: every write site has its own slot
// per element in a staging buffer, and the fixed kernel smh_stage_reduce
// combines each entry's writes in the reference's order (elements ascending,
// then the function's execution order), which the generator derives by
// walking f's loops and branches (lower_maps.cpp:37-49: += under reduce +,
// plain writes without it).  The result is deterministic and independent of
// the launch shape.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "frontend.h"

namespace simit {
namespace fe {

struct CodegenError : std::runtime_error {
  Loc loc;
  CodegenError(Loc l, const std::string &m) : std::runtime_error(m), loc(l) {}
};

// A map as the exported function uses it.
struct MapRequest {
  std::string func, target;
  std::string through;  // a grid edge set over target (map ... through)
  bool reduce = false;
  int nextra = 0;  // leading scalar arguments bound at the map
};

// One slot of a kernel's argument struct (every slot is 8 bytes).
struct KernelArg {
  enum Kind { Count, Endpoints, Stage, Taken, Field, Scalar, GridDim } kind = Count;
  std::string set, field;  // Field
  int index = 0;           // Stage / Taken: output; Scalar: extra argument; GridDim: axis
};

struct MapOutput {
  std::string name;
  bool matrix = false;        // tensor[V,V](block) (else tensor[V](block))
  bool stencil = false;       // a matrix over a grid's points: 27-point periodic stencil
  std::string rowSet;         // V
  std::vector<int> blockDims;  // dims of one entry ([] for float entries)
  int bs = 1;                 // doubles per entry
  int ninst = 0;              // write slots per element
  // the writes in execution order: slot, local row and column endpoint
  // (-1: the element itself, -2: the grid point at off[0] / off[1] from it;
  // col unused for vectors)
  struct Write {
    int inst, li, lj;
    int off[2][3];
  };
  std::vector<Write> order;
};

struct MapKernel {
  std::string kernel;       // extern "C" name in the generated module
  MapRequest req;
  int card = 0;             // endpoints per element (0: a map over a vertex set)
  std::string endpointSet;  // the set the endpoints index
  int gridDims = 0;         // a map through a grid of this many dimensions
  std::vector<MapOutput> outputs;
  std::vector<KernelArg> args;
  std::vector<std::pair<std::string, std::string>> fieldsWritten;  // (set, field) of inout writes
};

// Generates one module holding every requested map kernel (and the device
// functions they call).  Throws CodegenError for constructs outside the
// supported subset.
std::string generateModule(const ProgramAST &prog, const std::vector<MapRequest> &reqs,
                           std::vector<MapKernel> *kernels);

}  // namespace fe
}  // namespace simit
