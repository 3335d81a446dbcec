// Tetgen volume-mesh loader of the C++ host API: the interface of the
// reference's simit::MeshVol (src/mesh.h:42-95) for the parts the hot path
// feeds on -- vertex positions, elements and edges read from .node / .ele /
// .edge files (loadTet, loadTetEdge; src/mesh.cpp:252-360 semantics: the
// first integer is the row count, '#' lines and lines shorter than three
// characters are skipped, each data line starts with its index, extra
// columns are ignored, vertex ids are kept as written).
#pragma once

#include <array>
#include <istream>
#include <string>
#include <vector>

namespace simit {

struct MeshVol {
  std::vector<std::array<double, 3>> v;   // vertex positions
  std::vector<std::vector<int>> e;        // elements (vertex ids)
  std::vector<std::array<int, 2>> edges;  // edges (vertex ids)

  // 0 on success, -1 if a file cannot be opened or is malformed
  int loadTet(const std::string &nodeFile, const std::string &eleFile);
  int loadTet(std::istream &nodeIn, std::istream &eleIn);
  int loadTetEdge(const std::string &edgeFile);
  int loadTetEdge(std::istream &edgeIn);
};

}  // namespace simit
