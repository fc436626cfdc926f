#include "host_tree.hpp"

#include <algorithm>
#include <cmath>

namespace cuhc {

namespace {

// Cluster records over the points: a point is its cluster's representative
// when link == itself. Each record keeps the cluster's size and scipy id and
// a value carried by the cluster (its height) next to the link, so one find
// touches one cache line per step. Finds compress the whole path (two passes).
struct Node {
  int link, size, id;
  double height;
};

// The records live in a per-thread buffer kept between calls: a fresh 24 MB
// array at 10^6 points costs its page faults on every tree.
struct Clusters {
  Node *node;
  explicit Clusters(int n) {
    static thread_local std::vector<Node> buffer;
    if (buffer.size() < (size_t)n)
      buffer.resize(n);
    node = buffer.data();
    for (int i = 0; i < n; ++i)
      node[i] = {i, 1, i, 0.0};
  }
  Clusters(const Clusters &) = delete;
  Clusters &operator=(const Clusters &) = delete;
  int find(int x) {
    int r = x;
    while (node[r].link != r)
      r = node[r].link;
    while (node[x].link != r) {
      const int next = node[x].link;
      node[x].link = r;
      x = next;
    }
    return r;
  }
  // Links the smaller cluster under the larger; returns the representative.
  int join(int ra, int rb) {
    if (node[ra].size < node[rb].size)
      std::swap(ra, rb);
    node[rb].link = ra;
    node[ra].size += node[rb].size;
    return ra;
  }
};

} // namespace

bool mergesToTreeOrder(std::vector<Edge> &merges, int n) {
  if ((int)merges.size() != n - 1)
    return false;
  Clusters sets(n);
  for (Edge &m : merges) {
    if (!std::isfinite(m.w) || m.u < 0 || m.u >= n || m.v < 0 || m.v >= n)
      return false;
    const int ra = sets.find(m.u), rb = sets.find(m.v);
    if (ra == rb)
      return false;
    m.w = std::max({m.w, sets.node[ra].height, sets.node[rb].height});
    sets.node[sets.join(ra, rb)].height = m.w;
  }
  std::stable_sort(merges.begin(), merges.end(),
                   [](const Edge &x, const Edge &y) { return x.w < y.w; });
  return true;
}

namespace {

template <class E> void assembleTree(const E *merges, size_t m, int n, double *tree) {
  Clusters sets(n);
  Node *node = sets.node;
  // The finds start at random points: the records of the endpoints 16 merges
  // ahead, and their links 8 merges ahead, are prefetched.
  constexpr size_t kAhead = 16;
  for (size_t s = 0; s < m; ++s) {
    if (s + kAhead < m) {
      __builtin_prefetch(&node[merges[s + kAhead].u], 1);
      __builtin_prefetch(&node[merges[s + kAhead].v], 1);
    }
    if (s + kAhead / 2 < m) {
      __builtin_prefetch(&node[node[merges[s + kAhead / 2].u].link], 1);
      __builtin_prefetch(&node[node[merges[s + kAhead / 2].v].link], 1);
    }
    const int ra = sets.find(merges[s].u), rb = sets.find(merges[s].v);
    const int a = node[ra].id, b = node[rb].id;
    double *row = tree + s * 4;
    row[0] = std::min(a, b);
    row[1] = std::max(a, b);
    row[2] = (double)merges[s].w;
    row[3] = node[ra].size + node[rb].size;
    node[sets.join(ra, rb)].id = n + (int)s;
  }
}

} // namespace

void treeFromOrderedMerges(const std::vector<Edge> &merges, int n, double *tree) {
  assembleTree(merges.data(), merges.size(), n, tree);
}

void treeFromOrderedEdges(const Edge32 *edges, size_t m, int n, double *tree) {
  assembleTree(edges, m, n, tree);
}

void sortEdges(std::vector<Edge> &edges) { std::sort(edges.begin(), edges.end(), edgeLess); }

} // namespace cuhc
