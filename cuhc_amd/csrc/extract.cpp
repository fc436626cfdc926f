// HDBSCAN cluster extraction on the host: condensed tree, cluster
// stabilities, excess-of-mass / leaf selection, point labels and membership
// probabilities. The results are those of scikit-learn's HDBSCAN tree code
// (sklearn/cluster/_hdbscan/_tree.pyx: _condense_tree, _compute_stability,
// _get_clusters, _do_labelling, get_probabilities), which the reference
// reproduces in python/cuhc/_hdbscan.py:148-350 and src/hdbscan.cuh:393-483.
// Everything here is linear in the number of tree nodes.
#include "extract.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace cuhc {

namespace {

// Read-only view of a scipy-format merge tree: node v >= n is row v - n.
struct MergeTree {
  const double *z;
  long long n;
  long long left(long long v) const { return (long long)z[(v - n) * 4]; }
  long long right(long long v) const { return (long long)z[(v - n) * 4 + 1]; }
  double height(long long v) const { return z[(v - n) * 4 + 2]; }
  long long leaves(long long v) const { return v < n ? 1 : (long long)z[(v - n) * 4 + 3]; }
};

} // namespace

// The rows come out in the order of scikit-learn's _condense_tree, which
// walks the merge tree breadth first from the root (left child first) and,
// when a side falls out, lists that side's points in the side's own
// breadth-first order. One breadth-first pass over the whole tree gives both
// orders: restricted to a subtree, the global level order is the subtree's
// level order. Each event (a real split, or one side falling out) opens a
// block of rows; the points of a fallen side are collected in their block as
// the pass reaches them, and the blocks are concatenated in event order.
std::vector<CondensedRow> condenseTree(const double *tree, long long n, long long mcs) {
  std::vector<CondensedRow> rows;
  if (n < 2)
    return rows;
  const MergeTree t{tree, n};
  const long long nodes = 2 * n - 1, top = nodes - 1;

  struct Block {
    long long parent, child; // child < 0: a fallen side (its points follow)
    double lambda;
    long long size;
  };
  std::vector<Block> blocks;
  std::vector<long long> label(nodes, -1); // condensed cluster of a kept node
  std::vector<int> fallen(nodes, -1);      // block of a node inside a fallen side
  std::vector<std::pair<int, long long>> points; // (block, point) in pass order
  label[top] = n;
  long long nextLabel = n + 1;

  // Level by level, left to right: a node's state is set when its parent's
  // level is processed.
  std::vector<long long> frontier{top}, below;
  while (!frontier.empty()) {
    below.clear();
    for (const long long v : frontier) {
      if (v < n) {
        if (fallen[v] >= 0)
          points.emplace_back(fallen[v], v);
        continue;
      }
      const long long a = t.left(v), b = t.right(v);
      below.push_back(a);
      below.push_back(b);
      if (fallen[v] >= 0) {
        fallen[a] = fallen[b] = fallen[v];
        continue;
      }
      const double h = t.height(v);
      const double lambda = h > 0.0 ? 1.0 / h : std::numeric_limits<double>::infinity();
      const long long na = t.leaves(a), nb = t.leaves(b);
      const bool bigA = na >= mcs, bigB = nb >= mcs;
      if (bigA && bigB) {
        label[a] = nextLabel++;
        blocks.push_back({label[v], label[a], lambda, na});
        label[b] = nextLabel++;
        blocks.push_back({label[v], label[b], lambda, nb});
        continue;
      }
      if (bigA)
        label[a] = label[v];
      else {
        fallen[a] = (int)blocks.size();
        blocks.push_back({label[v], -1, lambda, 1});
      }
      if (bigB)
        label[b] = label[v];
      else {
        fallen[b] = (int)blocks.size();
        blocks.push_back({label[v], -1, lambda, 1});
      }
    }
    frontier.swap(below);
  }

  // Stable bucketing of the fallen points by block.
  std::vector<long long> start(blocks.size() + 1, 0);
  for (const auto &p : points)
    ++start[p.first + 1];
  for (size_t i = 0; i < blocks.size(); ++i)
    if (blocks[i].child >= 0)
      ++start[i + 1];
  for (size_t i = 1; i < start.size(); ++i)
    start[i] += start[i - 1];
  rows.resize(start.back());
  std::vector<long long> fillAt(start.begin(), start.end() - 1);
  for (size_t i = 0; i < blocks.size(); ++i)
    if (blocks[i].child >= 0)
      rows[fillAt[i]++] = {blocks[i].parent, blocks[i].child, blocks[i].lambda, blocks[i].size};
  for (const auto &p : points) {
    const Block &blk = blocks[p.first];
    rows[fillAt[p.first]++] = {blk.parent, p.second, blk.lambda, 1};
  }
  return rows;
}

CondensedView::CondensedView(const long long *p, const long long *c, const double *l,
                             const long long *s, long long r)
    : parent(p), child(c), lambda(l), size(s), rows(r) {
  if (rows <= 0)
    return;
  root = *std::min_element(parent, parent + rows);
  maxParent = *std::max_element(parent, parent + rows);
  long long maxChild = *std::max_element(child, child + rows);
  numIds = std::max(maxParent, maxChild) + 1;
}

// Per cluster id c (index c - root): the sum over rows with parent c, in row
// order, of (lambda - lambda at which c was born) * size; the root is born at 0.
std::vector<double> clusterStability(const CondensedView &t) {
  std::vector<double> out;
  if (t.rows <= 0)
    return out;
  std::vector<double> birth(t.numIds, std::numeric_limits<double>::quiet_NaN());
  for (long long i = 0; i < t.rows; ++i)
    birth[t.child[i]] = t.lambda[i];
  birth[t.root] = 0.0;
  out.assign(t.maxParent - t.root + 1, 0.0);
  for (long long i = 0; i < t.rows; ++i)
    out[t.parent[i] - t.root] += (t.lambda[i] - birth[t.parent[i]]) * (double)t.size[i];
  return out;
}

namespace {

// The cluster rows (size > 1) as a tree over ids root .. maxParent.
struct ClusterTree {
  long long root = 0, count = 0;
  std::vector<long long> up;          // parent of a cluster (-1: none)
  std::vector<long long> members;     // size of a cluster as a child row
  std::vector<long long> firstKid, nextKid; // children lists, row order
  explicit ClusterTree(const CondensedView &t) {
    root = t.root;
    count = t.maxParent - t.root + 1;
    long long ids = std::max(count, t.numIds - t.root);
    up.assign(ids, -1);
    members.assign(ids, 0);
    firstKid.assign(ids, -1);
    nextKid.assign(ids, -1);
    std::vector<long long> lastKid(ids, -1);
    for (long long i = 0; i < t.rows; ++i) {
      if (t.size[i] <= 1)
        continue;
      const long long c = t.child[i] - root, p = t.parent[i] - root;
      up[c] = p;
      members[c] = t.size[i];
      if (lastKid[p] < 0)
        firstKid[p] = c;
      else
        nextKid[lastKid[p]] = c;
      lastKid[p] = c;
    }
  }
  bool isLeaf(long long c) const { return firstKid[c] < 0; }
};

} // namespace

int selectClusters(const CondensedView &t, double *stability, int method, bool allowSingle,
                   long long maxClusterSize, unsigned char *selected) {
  if (t.rows <= 0)
    return 0;
  const ClusterTree ct(t);
  const long long m = ct.count;
  std::fill(selected, selected + m, 0);
  // Candidates: every id but the root's, unless a single cluster is allowed.
  const long long first = allowSingle ? 0 : 1;
  long long numPoints = 0;
  for (long long i = 0; i < t.rows; ++i)
    if (t.size[i] == 1)
      numPoints = std::max(numPoints, t.child[i] + 1);
  if (maxClusterSize <= 0)
    maxClusterSize = numPoints + 1;
  std::vector<long long> clusterSize(ct.members.begin(), ct.members.begin() + m);
  if (allowSingle) {
    long long total = 0;
    for (long long k = ct.firstKid[0]; k >= 0; k = ct.nextKid[k])
      total += ct.members[k];
    clusterSize[0] = total;
  }

  if (method == 0) {
    // Excess of mass, children before parents (larger ids first). A kept
    // cluster hides every cluster below it; a cluster is selected when it was
    // kept and no cluster above it was.
    std::vector<unsigned char> kept(m, 0);
    for (long long c = m - 1; c >= first; --c) {
      double below = 0.0;
      for (long long k = ct.firstKid[c]; k >= 0; k = ct.nextKid[k])
        below += stability[k];
      if (below > stability[c] || clusterSize[c] > maxClusterSize)
        stability[c] = below;
      else
        kept[c] = 1;
    }
    std::vector<unsigned char> hidden(m, 0);
    for (long long c = 1; c < m; ++c) { // parents have smaller ids
      const long long p = ct.up[c];
      hidden[c] = p >= 0 && (hidden[p] || kept[p]);
    }
    for (long long c = first; c < m; ++c)
      selected[c] = kept[c] && !hidden[c];
    return 0;
  }
  // Leaf selection: the leaves of the cluster tree among the candidates (none
  // when the tree has no cluster rows).
  bool anyCluster = false;
  for (long long c = 1; c < m && !anyCluster; ++c)
    anyCluster = ct.up[c] >= 0;
  if (!anyCluster)
    return 0;
  for (long long c = first; c < m; ++c)
    selected[c] = (c == 0 || ct.up[c] >= 0) && ct.isLeaf(c);
  return 0;
}

namespace {

// Output label of each selected cluster (index id - root): its rank among the
// selected ids, -1 for the others.
std::vector<long long> selectedRanks(const CondensedView &t, const unsigned char *selected,
                                     long long *count) {
  const long long m = t.maxParent - t.root + 1;
  std::vector<long long> rank(m, -1);
  long long k = 0;
  for (long long c = 0; c < m; ++c)
    if (selected[c])
      rank[c] = k++;
  *count = k;
  return rank;
}

} // namespace

void labelPoints(const CondensedView &t, const unsigned char *selected, bool allowSingle,
                 double epsilon, long long *labels) {
  const long long n = t.root; // points are 0 .. root - 1
  std::fill(labels, labels + n, -1);
  if (t.rows <= 0)
    return;
  const long long m = t.maxParent - t.root + 1;
  long long numSelected = 0;
  const std::vector<long long> rank = selectedRanks(t, selected, &numSelected);
  // owner[c]: the selected cluster a point falling out of c belongs to (the
  // nearest selected cluster at or above c), or the root. Parents precede
  // their children in row order.
  std::vector<long long> owner(m, -1);
  owner[0] = 0;
  for (long long i = 0; i < t.rows; ++i) {
    const long long c = t.child[i] - t.root;
    if (c <= 0 || c >= m || owner[c] >= 0)
      continue;
    owner[c] = selected[c] ? c : owner[t.parent[i] - t.root];
  }
  // The root's points join a single allowed root cluster only above a lambda
  // threshold: 1 / epsilon, or the largest lambda among the root's rows.
  const bool singleRoot = numSelected == 1 && allowSingle;
  double threshold = 0.0;
  if (singleRoot) {
    if (epsilon != 0.0) {
      threshold = 1.0 / epsilon;
    } else {
      threshold = -std::numeric_limits<double>::infinity();
      for (long long i = 0; i < t.rows; ++i)
        if (t.parent[i] == t.root)
          threshold = std::max(threshold, t.lambda[i]);
    }
  }
  for (long long i = 0; i < t.rows; ++i) {
    const long long p = t.child[i];
    if (p >= n)
      continue;
    const long long c = owner[t.parent[i] - t.root];
    if (c != 0)
      labels[p] = rank[c];
    else if (singleRoot && t.lambda[i] >= threshold)
      labels[p] = rank[0];
  }
}

void membershipProbabilities(const CondensedView &t, const unsigned char *selected,
                             const long long *labels, double *probabilities) {
  const long long n = t.root;
  std::fill(probabilities, probabilities + n, 0.0);
  if (t.rows <= 0)
    return;
  const long long m = t.maxParent - t.root + 1;
  // A cluster's death lambda is the largest lambda of the last run of
  // consecutive rows that has it as parent (scikit-learn's max_lambdas).
  std::vector<double> death(m, 0.0);
  for (long long i = 0; i < t.rows;) {
    long long j = i;
    double top = t.lambda[i];
    while (++j < t.rows && t.parent[j] == t.parent[i])
      top = std::max(top, t.lambda[j]);
    death[t.parent[i] - t.root] = top;
    i = j;
  }
  long long numSelected = 0;
  const std::vector<long long> rank = selectedRanks(t, selected, &numSelected);
  std::vector<long long> clusterOfRank(numSelected);
  for (long long c = 0; c < m; ++c)
    if (rank[c] >= 0)
      clusterOfRank[rank[c]] = c;
  for (long long i = 0; i < t.rows; ++i) {
    const long long p = t.child[i];
    if (p >= n || labels[p] < 0 || labels[p] >= numSelected)
      continue;
    const double top = death[clusterOfRank[labels[p]]];
    const double lam = t.lambda[i];
    probabilities[p] = (top == 0.0 || std::isinf(lam)) ? 1.0 : std::min(lam, top) / top;
  }
}

} // namespace cuhc
