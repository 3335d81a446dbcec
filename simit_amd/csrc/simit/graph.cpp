#include "graph.h"

namespace simit {

Box createBox(Set *vertices, Set *edges, unsigned numX, unsigned numY, unsigned numZ) {
  if (numX < 1 || numY < 1 || numZ < 1) throw SimitException("createBox: empty box");
  std::vector<ElementRef> pts((size_t)numX * numY * numZ);
  auto node = [&](unsigned x, unsigned y, unsigned z) {
    return (size_t)x * numY * numZ + (size_t)y * numZ + z;
  };
  for (unsigned x = 0; x < numX; ++x)
    for (unsigned y = 0; y < numY; ++y)
      for (unsigned z = 0; z < numZ; ++z) pts[node(x, y, z)] = vertices->add();
  for (unsigned x = 0; x + 1 < numX; ++x)
    for (unsigned y = 0; y < numY; ++y)
      for (unsigned z = 0; z < numZ; ++z) edges->add(pts[node(x, y, z)], pts[node(x + 1, y, z)]);
  for (unsigned x = 0; x < numX; ++x)
    for (unsigned y = 0; y + 1 < numY; ++y)
      for (unsigned z = 0; z < numZ; ++z) edges->add(pts[node(x, y, z)], pts[node(x, y + 1, z)]);
  for (unsigned x = 0; x < numX; ++x)
    for (unsigned y = 0; y < numY; ++y)
      for (unsigned z = 0; z + 1 < numZ; ++z) edges->add(pts[node(x, y, z)], pts[node(x, y, z + 1)]);
  return Box(numX, numY, numZ, pts);
}

}  // namespace simit
