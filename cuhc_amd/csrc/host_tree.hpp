// Host assembly of scipy-format merge trees from merges (edges between
// points), shared by every linkage path. The tree format and the global edge
// order follow reference src/mst.cuh:36-48 and 204-235.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <vector>

#include "extract.hpp"

namespace cuhc {

struct Edge {
  int u, v; // u < v
  double w; // FP32 distances are stored exactly; Ward heights keep FP64
};

// The one global edge order: weight, then smaller endpoint, then larger.
inline bool edgeLess(const Edge &x, const Edge &y) {
  if (x.w != y.w)
    return x.w < y.w;
  if (x.u != y.u)
    return x.u < y.u;
  return x.v < y.v;
}

// Merges that arrive in chain or round order (each merge after those that
// built its two clusters, heights possibly not monotone) into tree order:
// checks that there are n - 1 of them with finite heights, each joining two
// clusters not yet joined; raises every height to at least its children's;
// stable-sorts by height. Returns false (merges unspecified) when the check
// fails.
bool mergesToTreeOrder(std::vector<Edge> &merges, int n);

// Scipy linkage rows {min id, max id, height, size} from merges already in
// tree order; the cluster made by merge s is n + s.
void treeFromOrderedMerges(const std::vector<Edge> &merges, int n, double *tree);

// The same from FP32-weighted edges (the layout the device sort produces),
// sorted by (weight, smaller endpoint) only: each run of equal (weight, u) is
// ordered by v in place as the assembly reaches it.
struct Edge32 {
  int u, v; // u < v
  float w;
};
void treeFromOrderedEdges(Edge32 *edges, size_t m, int n, double *tree);

// Sorts by edgeLess on the host.
void sortEdges(std::vector<Edge> &edges);

} // namespace cuhc
