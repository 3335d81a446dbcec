// Drives the REFERENCE's own PathIndexBuilder (src/path_indices.cpp:148-530)
// on meshes given in a small binary file, so that the oracle's and the HIP
// builder's vev indices can be compared with the reference's, bit for bit.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/ref/Makefile from the reference
// sources into oracle/_ref/ (never committed); only tests/ run it.
//
// Input (little endian):  int32 nv, int32 nsets,
//                         per set: int64 ne, int32 card, int32 ep[ne*card]
// Output:                 int32 nrows, int64 nnz, uint32 rowptr[nrows+1],
//                         uint32 colidx[nnz]
// Mode argument: "vev" (And with an existential element per set, Or-ed over
// the sets as in test/path_indices-test.cpp:112-150), "ve" (Link::ve of set
// 0, :48-56), or "box" (createBox ordering, graph.cpp:66-115: prints the
// endpoints of a createBox(nx,ny,nz) edge set).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "graph.h"
#include "path_expressions.h"
#include "path_indices.h"

using namespace simit;
using namespace simit::pe;

static std::unique_ptr<simit::Set> makeEdgeSet(const std::string &name,
                                              simit::Set &V, int card) {
  switch (card) {
    case 1: return std::unique_ptr<simit::Set>(new simit::Set(name, V));
    case 2: return std::unique_ptr<simit::Set>(new simit::Set(name, V, V));
    case 3: return std::unique_ptr<simit::Set>(new simit::Set(name, V, V, V));
    case 4: return std::unique_ptr<simit::Set>(new simit::Set(name, V, V, V, V));
    case 8:
      return std::unique_ptr<simit::Set>(
          new simit::Set(name, V, V, V, V, V, V, V, V));
  }
  std::cerr << "unsupported cardinality " << card << "\n";
  std::exit(2);
}

static void addEdge(simit::Set &E, const std::vector<ElementRef> &v,
                    const int32_t *ep, int card) {
  switch (card) {
    case 1: E.add(v[ep[0]]); break;
    case 2: E.add(v[ep[0]], v[ep[1]]); break;
    case 3: E.add(v[ep[0]], v[ep[1]], v[ep[2]]); break;
    case 4: E.add(v[ep[0]], v[ep[1]], v[ep[2]], v[ep[3]]); break;
    case 8:
      E.add(v[ep[0]], v[ep[1]], v[ep[2]], v[ep[3]], v[ep[4]], v[ep[5]],
            v[ep[6]], v[ep[7]]);
      break;
  }
}

static void writeIndex(const char *path, const PathIndex &pi) {
  std::ofstream out(path, std::ios::binary);
  int32_t nrows = (int32_t)pi.numElements();
  int64_t nnz = (int64_t)pi.numNeighbors();
  out.write((const char *)&nrows, 4);
  out.write((const char *)&nnz, 8);
  std::vector<uint32_t> rowptr(nrows + 1), colidx;
  uint32_t acc = 0;
  for (int32_t r = 0; r < nrows; ++r) {
    rowptr[r] = acc;
    for (auto n : pi.neighbors(r)) {
      colidx.push_back(n);
      ++acc;
    }
  }
  rowptr[nrows] = acc;
  out.write((const char *)rowptr.data(), 4 * (nrows + 1));
  out.write((const char *)colidx.data(), 4 * colidx.size());
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: ref_index vev|ve <in> <out>  |  ref_index box nx ny nz <out>\n";
    return 2;
  }
  std::string mode = argv[1];
  if (mode == "box") {
    simit::Set V, E(V, V);
    createBox(&V, &E, std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    std::ofstream out(argv[5], std::ios::binary);
    int64_t ne = E.getSize();
    out.write((const char *)&ne, 8);
    out.write((const char *)E.getEndpointsData(), 4 * 2 * ne);
    return 0;
  }
  std::ifstream in(argv[2], std::ios::binary);
  int32_t nv = 0, nsets = 0;
  in.read((char *)&nv, 4);
  in.read((char *)&nsets, 4);
  simit::Set V("V");
  std::vector<ElementRef> verts;
  for (int32_t i = 0; i < nv; ++i) verts.push_back(V.add());
  std::vector<std::unique_ptr<simit::Set>> sets;
  for (int32_t s = 0; s < nsets; ++s) {
    int64_t ne = 0;
    int32_t card = 0;
    in.read((char *)&ne, 8);
    in.read((char *)&card, 4);
    std::vector<int32_t> ep((size_t)ne * card);
    in.read((char *)ep.data(), 4 * ep.size());
    sets.push_back(makeEdgeSet("E" + std::to_string(s), V, card));
    for (int64_t e = 0; e < ne; ++e) addEdge(*sets.back(), verts, &ep[e * card], card);
  }
  PathIndexBuilder builder;
  builder.bind("V", &V);
  for (int32_t s = 0; s < nsets; ++s) builder.bind("E" + std::to_string(s), sets[s].get());

  Var vi("vi"), vj("vj");
  if (mode == "ve") {
    Var v("v", simit::pe::Set("V")), e("e", simit::pe::Set("E0"));
    PathExpression ve = Link::make(v, e, Link::ve);
    writeIndex(argv[3], builder.buildSegmented(ve, 0));
    return 0;
  }
  PathExpression acc;
  for (int32_t s = 0; s < nsets; ++s) {
    std::string en = "E" + std::to_string(s);
    Var e("e" + std::to_string(s));
    Var v1("v", simit::pe::Set("V")), e1("e", simit::pe::Set(en));
    Var e2("e", simit::pe::Set(en)), v2("v", simit::pe::Set("V"));
    PathExpression ve = Link::make(v1, e1, Link::ve);
    PathExpression ev = Link::make(e2, v2, Link::ev);
    PathExpression vev = And::make({vi, vj}, {{QuantifiedVar::Exist, e}},
                                   ve(vi, e), ev(e, vj));
    acc = s == 0 ? vev : Or::make({vi, vj}, {}, acc(vi, vj), vev(vi, vj));
  }
  writeIndex(argv[3], builder.buildSegmented(acc, 0));
  return 0;
}
