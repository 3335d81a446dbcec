#include "simit/reorder.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace simit {

uint64_t hilbertIndex3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  // inverse undo: for each bit level from the top, exchange or invert the
  // lower bits of the axes so that the curve's sub-cubes line up
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;  // invert the low bits of the first axis
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;  // exchange low bits with axis i
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  // Gray encode
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  // interleave the transposed bits: the top bit of X[0], X[1], X[2], then
  // the next level, ...
  uint64_t h = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

void reorderEdgeSetByVertexOrdering(Set &edgeSet, const std::vector<int> &vertexOrdering) {
  for (int c = 0; c < edgeSet.getCardinality(); ++c) {
    const Set *v = edgeSet.getEndpointSet(c);
    if (v->getSize() == (int)vertexOrdering.size()) {
      edgeSet.remapEndpoints(*v, vertexOrdering);
      return;  // remapEndpoints renumbers every slot that refers to v
    }
  }
  throw SimitException("reorderEdgeSetByVertexOrdering: no endpoint set of that size");
}

void reorderVertexSet(Set &edgeSet, Set &vertexSet, std::vector<int> &vertexOrdering) {
  vertexSet.permuteElements(vertexOrdering);
  edgeSet.remapEndpoints(vertexSet, vertexOrdering);
}

void reorderEdgeSet(Set &edgeSet, const std::vector<int> &edgeOrdering) {
  edgeSet.permuteElements(edgeOrdering);
}

void reorder(Set &edgeSet, Set &vertexSet, std::vector<int> &edgeOrdering,
             std::vector<int> &vertexOrdering) {
  if (!vertexSet.hasSpatialField())
    throw SimitException("reorder: the vertex set has no spatial field");
  const int n = vertexSet.getSize();
  const FieldData *xf = vertexSet.field(vertexSet.getSpatialFieldName());
  const double *x = reinterpret_cast<const double *>(xf->bytes.data());
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::fmin(lo[d], x[3 * i + d]);
      hi[d] = std::fmax(hi[d], x[3 * i + d]);
    }
  // 2^8 lattice points per axis over the bounding box, as the reference
  const int bits = 8;
  const double cells = (double)((1u << bits) - 1);
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; ++i) {
    uint32_t c[3];
    for (int d = 0; d < 3; ++d) {
      const double span = hi[d] - lo[d];
      c[d] = span > 0 ? (uint32_t)std::lround((x[3 * i + d] - lo[d]) / span * cells) : 0u;
    }
    key[i] = hilbertIndex3(c[0], c[1], c[2], bits);
  }
  std::vector<int> byKey(n);
  std::iota(byKey.begin(), byKey.end(), 0);
  std::stable_sort(byKey.begin(), byKey.end(), [&](int a, int b) { return key[a] < key[b]; });
  vertexOrdering.assign(n, 0);
  for (int r = 0; r < n; ++r) vertexOrdering[byKey[r]] = r;  // old -> new
  reorderVertexSet(edgeSet, vertexSet, vertexOrdering);
  // edges by their sorted endpoint tuples
  const int m = edgeSet.getSize(), card = edgeSet.getCardinality();
  const int *ep = edgeSet.getEndpointsData();
  std::vector<int> sorted(ep, ep + (size_t)m * card);
  for (int e = 0; e < m; ++e) std::sort(sorted.begin() + (size_t)e * card, sorted.begin() + (size_t)(e + 1) * card);
  std::vector<int> byEdge(m);
  std::iota(byEdge.begin(), byEdge.end(), 0);
  std::stable_sort(byEdge.begin(), byEdge.end(), [&](int a, int b) {
    return std::lexicographical_compare(sorted.begin() + (size_t)a * card,
                                        sorted.begin() + (size_t)(a + 1) * card,
                                        sorted.begin() + (size_t)b * card,
                                        sorted.begin() + (size_t)(b + 1) * card);
  });
  edgeOrdering.assign(m, 0);
  for (int r = 0; r < m; ++r) edgeOrdering[byEdge[r]] = r;
  reorderEdgeSet(edgeSet, edgeOrdering);
}

void reorder(Set &edgeSet, Set &vertexSet) {
  std::vector<int> e, v;
  reorder(edgeSet, vertexSet, e, v);
}

}  // namespace simit
