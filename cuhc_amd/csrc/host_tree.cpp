#include "host_tree.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace cuhc {

UnionFind::UnionFind(int n) : parent_(n), size_(n, 1) {
  std::iota(parent_.begin(), parent_.end(), 0);
}

int UnionFind::unite(int x, int y) {
  x = find(x);
  y = find(y);
  if (x == y)
    return -1;
  if (size_[x] < size_[y])
    std::swap(x, y);
  parent_[y] = x;
  size_[x] += size_[y];
  return x;
}

bool mergesFormTree(const std::vector<Edge> &merges, int n) {
  if ((int)merges.size() != n - 1)
    return false;
  UnionFind sets(n);
  for (const Edge &m : merges)
    if (!std::isfinite(m.w) || m.u < 0 || m.u >= n || m.v < 0 || m.v >= n ||
        sets.unite(m.u, m.v) < 0)
      return false;
  return true;
}

void treeFromOrderedMerges(const std::vector<Edge> &merges, int n, double *tree) {
  // One record per point keeps parent, set size and cluster id in one cache
  // line access (the finds are random over n points).
  struct Node {
    int parent, size, id;
  };
  std::vector<Node> node(n);
  for (int i = 0; i < n; ++i)
    node[i] = {i, 1, i};
  auto find = [&](int x) {
    while (node[x].parent != x) {
      const int up = node[node[x].parent].parent;
      node[x].parent = up;
      x = up;
    }
    return x;
  };
  // The finds start at random points: the records of the endpoints 16 merges
  // ahead, and their parents 8 merges ahead, are prefetched.
  constexpr size_t kAhead = 16;
  const size_t m = merges.size();
  for (size_t s = 0; s < m; ++s) {
    if (s + kAhead < m) {
      __builtin_prefetch(&node[merges[s + kAhead].u], 1);
      __builtin_prefetch(&node[merges[s + kAhead].v], 1);
    }
    if (s + kAhead / 2 < m) {
      __builtin_prefetch(&node[node[merges[s + kAhead / 2].u].parent], 1);
      __builtin_prefetch(&node[node[merges[s + kAhead / 2].v].parent], 1);
    }
    int ru = find(merges[s].u), rv = find(merges[s].v);
    const int a = node[ru].id, b = node[rv].id;
    const int total = node[ru].size + node[rv].size;
    double *row = tree + s * 4;
    row[0] = std::min(a, b);
    row[1] = std::max(a, b);
    row[2] = merges[s].w;
    row[3] = total;
    if (node[ru].size < node[rv].size)
      std::swap(ru, rv);
    node[rv].parent = ru;
    node[ru].size = total;
    node[ru].id = n + (int)s;
  }
}

void sortEdges(std::vector<Edge> &edges) { std::sort(edges.begin(), edges.end(), edgeLess); }

void orderChainMerges(std::vector<Edge> &merges, int n) {
  UnionFind sets(n);
  std::vector<double> height(n, 0.0); // per set root
  for (Edge &m : merges) {
    const int ru = sets.find(m.u), rv = sets.find(m.v);
    m.w = std::max({m.w, height[ru], height[rv]});
    height[sets.unite(ru, rv)] = m.w;
  }
  std::stable_sort(merges.begin(), merges.end(),
                   [](const Edge &x, const Edge &y) { return x.w < y.w; });
}

// Walks the tree breadth-first from the root (left child before right):
//   both sides >= minClusterSize: a real split, two new clusters;
//   one side smaller: the big side keeps the parent's label and the small
//                     side's points fall out at this lambda;
//   both smaller: all points fall out and the parent ends.
// Rows come out in the order sklearn's _condense_tree emits them.
std::vector<CondensedRow> condenseTree(const double *tree, long long n, long long mcs) {
  std::vector<CondensedRow> rows;
  if (n < 2)
    return rows;
  const long long root = 2 * (n - 1);
  auto left = [&](long long node) { return (long long)tree[(node - n) * 4 + 0]; };
  auto right = [&](long long node) { return (long long)tree[(node - n) * 4 + 1]; };
  auto height = [&](long long node) { return tree[(node - n) * 4 + 2]; };
  auto count = [&](long long node) { return node >= n ? (long long)tree[(node - n) * 4 + 3] : 1; };

  std::vector<long long> queue;
  auto bfs = [&](long long from, std::vector<long long> &out) {
    queue.assign(1, from);
    for (size_t head = 0; head < queue.size(); ++head) {
      const long long node = queue[head];
      out.push_back(node);
      if (node >= n) {
        queue.push_back(left(node));
        queue.push_back(right(node));
      }
    }
  };

  std::vector<long long> order;
  order.reserve(root + 1);
  bfs(root, order);
  std::vector<long long> relabel(root + 1, 0);
  std::vector<char> ignore(root + 1, 0);
  relabel[root] = n;
  long long next = n + 1;
  rows.reserve(n + 2 * (n / std::max(mcs, 1LL)) + 2);

  std::vector<long long> sub;
  auto drop = [&](long long parentLabel, long long subRoot, double lambda) {
    sub.clear();
    bfs(subRoot, sub);
    for (long long s : sub) {
      if (s < n)
        rows.push_back({parentLabel, s, lambda, 1});
      ignore[s] = 1;
    }
  };

  for (long long node : order) {
    if (ignore[node] || node < n)
      continue;
    const long long l = left(node), r = right(node);
    const double h = height(node);
    const double lambda = h > 0.0 ? 1.0 / h : INFINITY;
    const long long lc = count(l), rc = count(r);
    if (lc >= mcs && rc >= mcs) {
      relabel[l] = next++;
      rows.push_back({relabel[node], relabel[l], lambda, lc});
      relabel[r] = next++;
      rows.push_back({relabel[node], relabel[r], lambda, rc});
    } else if (lc < mcs && rc < mcs) {
      drop(relabel[node], l, lambda);
      drop(relabel[node], r, lambda);
    } else if (lc < mcs) {
      relabel[r] = relabel[node];
      drop(relabel[node], l, lambda);
    } else {
      relabel[l] = relabel[node];
      drop(relabel[node], r, lambda);
    }
  }
  return rows;
}

} // namespace cuhc
