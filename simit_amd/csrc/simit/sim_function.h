// The backend::Function of a parsed Simit program's function (sim_function.cpp).
#pragma once

#include <memory>
#include <string>

#include "frontend.h"
#include "program.h"

namespace simit {
namespace hipimpl {

// Throws SimitException if the function does not exist or uses constructs
// outside the supported subset (checked here, before any device work).
std::shared_ptr<backend::Function> makeSimFunction(std::shared_ptr<const fe::ProgramAST> prog,
                                                   const std::string &function);

// The HIP source generated for the maps of a function (for inspection and for
// compile checks that need no device).
std::string generatedSource(const fe::ProgramAST &prog, const std::string &function);

}  // namespace hipimpl
}  // namespace simit
