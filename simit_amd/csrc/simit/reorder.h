// Reordering of a vertex set and an edge set for locality: the interface of
// the reference's src/reorder.h:14-31.  An ordering maps each element's old
// position to its new one (element i moves to ordering[i]), for vertices and
// edges alike; edge endpoints follow their vertices.  (The reference moves
// an edge set's endpoints by the inverse of the map it applies to the edge
// fields; the two agree for the swaps its tests use.  Here one map moves
// both.)
#pragma once

#include <cstdint>
#include <vector>

#include "simit/graph.h"

namespace simit {

// Hilbert-curve order of the vertices (by the vertex set's spatial field,
// ties by old position), then the edges sorted by their sorted endpoint
// tuples (ties by old position).
void reorder(Set &edgeSet, Set &vertexSet);
void reorder(Set &edgeSet, Set &vertexSet, std::vector<int> &edgeOrdering,
             std::vector<int> &vertexOrdering);
// move vertex i to vertexOrdering[i] and renumber the edge set's endpoints
void reorderVertexSet(Set &edgeSet, Set &vertexSet, std::vector<int> &vertexOrdering);
// move edge i to edgeOrdering[i]
void reorderEdgeSet(Set &edgeSet, const std::vector<int> &edgeOrdering);
// renumber the endpoints after the vertices moved by vertexOrdering
void reorderEdgeSetByVertexOrdering(Set &edgeSet, const std::vector<int> &vertexOrdering);

// Index along the 3-D Hilbert curve of order `bits` of the lattice point
// (x, y, z), each coordinate in [0, 2^bits) -- Skilling's transpose
// algorithm (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc.
// 707, 2004): undo the excess work per bit level, Gray-decode, interleave.
uint64_t hilbertIndex3(uint32_t x, uint32_t y, uint32_t z, int bits);

}  // namespace simit
