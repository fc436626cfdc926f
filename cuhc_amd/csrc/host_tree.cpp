#include "host_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

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

// Union-find for the ordered merges: a link per point and, per root of a
// merged cluster, its size and scipy id in one 8-byte record, both kept
// between calls: 4 bytes per point touched at random plus one record per
// root, instead of a 24-byte node, halves the assembly time at 10^6 points.
// A link is the parent (>= 0), kAlone for a point not merged yet (its own
// root: size 1, id itself, no record read) or kMerged for the root of a
// merged cluster, so only the links are initialised (one memset) and most
// merges, which take in a point still alone, read one record instead of two:
// 10^6 points after 80 ms of GPU work, 11-12 ms against 8.5-9 ms.
// Finds halve the path as they climb; the links of the endpoints 16 merges
// ahead are prefetched. With ORDER_TIES, a run of merges with equal (weight,
// u) is sorted by v when the loop reaches it (the run is read anyway).
struct RootInfo {
  int size, id;
};
constexpr int kAlone = -1, kMerged = -2;

template <bool ORDER_TIES, class E> void assembleTree(E *merges, size_t m, int n, double *tree) {
  static thread_local std::vector<int> link;
  static thread_local std::vector<RootInfo> root;
  if (link.size() < (size_t)n) {
    link.resize(n);
    root.resize(n);
  }
  int *up = link.data();
  RootInfo *info = root.data();
  static_assert(kAlone == -1, "all-ones bytes");
  std::memset(up, 0xff, (size_t)n * sizeof(int));
  auto find = [up](int x) {
    while (up[x] >= 0) {
      const int p = up[x];
      if (up[p] >= 0)
        up[x] = up[p];
      x = p;
    }
    return x;
  };
  constexpr size_t kAhead = 16;
  size_t runEnd = 0; // merges before runEnd have their ties ordered
  for (size_t s = 0; s < m; ++s) {
    if constexpr (ORDER_TIES) {
      if (s >= runEnd) {
        size_t j = s + 1;
        while (j < m && merges[j].w == merges[s].w && merges[j].u == merges[s].u)
          ++j;
        if (j - s > 1)
          std::sort(merges + s, merges + j, [](const E &x, const E &y) { return x.v < y.v; });
        runEnd = j;
      }
    }
    if (s + kAhead < m) {
      __builtin_prefetch(&up[merges[s + kAhead].u], 1);
      __builtin_prefetch(&up[merges[s + kAhead].v], 1);
    }
    int ra = find(merges[s].u), rb = find(merges[s].v);
    const RootInfo ia = up[ra] == kAlone ? RootInfo{1, ra} : info[ra];
    const RootInfo ib = up[rb] == kAlone ? RootInfo{1, rb} : info[rb];
    const int total = ia.size + ib.size;
    double *row = tree + s * 4;
    row[0] = std::min(ia.id, ib.id);
    row[1] = std::max(ia.id, ib.id);
    row[2] = (double)merges[s].w;
    row[3] = total;
    if (ia.size < ib.size)
      std::swap(ra, rb);
    up[rb] = ra;
    up[ra] = kMerged;
    info[ra] = {total, n + (int)s};
  }
}

} // namespace

void treeFromOrderedMerges(const std::vector<Edge> &merges, int n, double *tree) {
  assembleTree<false>(merges.data(), merges.size(), n, tree);
}

void treeFromOrderedEdges(Edge32 *edges, size_t m, int n, double *tree) {
  assembleTree<true>(edges, m, n, tree);
}

void sortEdges(std::vector<Edge> &edges) { std::sort(edges.begin(), edges.end(), edgeLess); }

} // namespace cuhc
