// Host C++ tree assembly shared by every linkage path.
//
// Edge order, union-find and the scipy-format tree follow reference
// src/mst.cuh:36-78 and 188-235; orderChainMerges follows src/nn_chain.cuh:178-188;
// condenseTree follows src/hdbscan.cuh:393-483 (sklearn's _condense_tree).
#pragma once

#include <stdint.h>
#include <vector>

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

// Disjoint sets with path halving and union by size.
class UnionFind {
public:
  explicit UnionFind(int n);
  int find(int x) {
    while (parent_[x] != x) {
      parent_[x] = parent_[parent_[x]];
      x = parent_[x];
    }
    return x;
  }
  // Joins the sets of x and y; the new root, or -1 if already joined.
  int unite(int x, int y);

private:
  std::vector<int> parent_, size_;
};

// n - 1 merges with finite heights, each joining two clusters not yet joined.
bool mergesFormTree(const std::vector<Edge> &merges, int n);

// Scipy linkage rows {min id, max id, height, size} from merges already in
// tree order; the cluster made by merge s is n + s.
void treeFromOrderedMerges(const std::vector<Edge> &merges, int n, double *tree);

// Sorts by edgeLess on the host.
void sortEdges(std::vector<Edge> &edges);

// Makes heights non-decreasing up the tree (a parent at least its children),
// then stable-sorts by height: merges in chain/round order -> tree order.
void orderChainMerges(std::vector<Edge> &merges, int n);

struct CondensedRow {
  long long parent, child;
  double lambda;
  long long size;
};

std::vector<CondensedRow> condenseTree(const double *tree, long long n, long long minClusterSize);

} // namespace cuhc
