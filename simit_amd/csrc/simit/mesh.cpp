#include "simit/mesh.h"

#include <fstream>
#include <sstream>

namespace simit {
namespace {

// Reads the header count (returned through *count, the rest of the header
// line through *head) and then up to count data lines, each handed to
// row() as a stream positioned after the line's leading index.
template <class F>
int read_rows(std::istream &in, long *count, std::istringstream *head, F row) {
  std::string line;
  if (!std::getline(in, line)) return -1;
  std::istringstream h(line);
  if (!(h >> *count) || *count < 0) return -1;
  head->str(line);
  head->clear();
  long dummy;
  *head >> dummy;
  long got = 0;
  while (got < *count && std::getline(in, line)) {
    if (line.size() < 3 || line[0] == '#') continue;
    std::istringstream ss(line);
    long index;
    if (!(ss >> index)) return -1;
    if (!row(got, ss)) return -1;
    ++got;
  }
  return got == *count ? 0 : -1;
}

}  // namespace

int MeshVol::loadTet(std::istream &nodeIn, std::istream &eleIn) {
  long n = 0;
  std::istringstream head;
  int rc = read_rows(nodeIn, &n, &head, [&](long i, std::istringstream &ss) {
    if ((long)v.size() <= i) v.resize(i + 1);
    return bool(ss >> v[i][0] >> v[i][1] >> v[i][2]);
  });
  if (rc) return rc;
  v.resize(n);
  long ne = 0;
  int per = 4;
  std::istringstream eh;
  rc = read_rows(eleIn, &ne, &eh, [&](long i, std::istringstream &ss) {
    if (i == 0) {
      int p = 0;
      if (eh >> p && p > 0) per = p;
    }
    if ((long)e.size() <= i) e.resize(i + 1);
    e[i].resize(per);
    for (int a = 0; a < per; ++a)
      if (!(ss >> e[i][a])) return false;
    return true;
  });
  if (rc) return rc;
  e.resize(ne);
  return 0;
}

int MeshVol::loadTet(const std::string &nodeFile, const std::string &eleFile) {
  std::ifstream nodeIn(nodeFile), eleIn(eleFile);
  if (!nodeIn || !eleIn) return -1;
  return loadTet(nodeIn, eleIn);
}

int MeshVol::loadTetEdge(std::istream &edgeIn) {
  long n = 0;
  std::istringstream head;
  int rc = read_rows(edgeIn, &n, &head, [&](long i, std::istringstream &ss) {
    if ((long)edges.size() <= i) edges.resize(i + 1);
    return bool(ss >> edges[i][0] >> edges[i][1]);
  });
  if (rc) return rc;
  edges.resize(n);
  return 0;
}

int MeshVol::loadTetEdge(const std::string &edgeFile) {
  std::ifstream in(edgeFile);
  if (!in) return -1;
  return loadTetEdge(in);
}

}  // namespace simit
