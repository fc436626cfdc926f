// Boruvka's MST over the complete graph, matrix-free (reference
// src/mst.cuh:83-186), plus the single-linkage and HDBSCAN drivers
// (src/hdbscan.cuh:353-391, src/capi.cu:68-119).
//
// Each round: the GPU finds, for every point, its nearest point in another
// component (one all-pairs pass); the host takes each component's smallest
// edge under the total order (weight, u, v) in original indices, unites, and
// relabels. Every chosen edge is the minimum edge leaving its component under
// a strict total order, so the result is the unique MST under that order:
// independent of the pass's tile order, pruning, or point permutation.
#include "mst.hpp"

#include <algorithm>
#include <numeric>

namespace cuhc {

int DevicePoints::upload(const float *hostPoints, int n, int d, bool spatial) {
  const size_t count = (size_t)n * d;
  Status s;
  s(points.alloc(count));
  if (!s.ok())
    return s.code;
  if (!spatial)
    return cuhc::upload(points, hostPoints, count * sizeof(float));
  order = spatialOrder(hostPoints, n, d);
  DevArray<float> raw;
  s(raw.alloc(count));
  s(orig.alloc(n));
  s(cuhc::upload(raw, hostPoints, count * sizeof(float)));
  s(cuhc::upload(orig, order.data(), n * sizeof(int)));
  if (s.ok())
    s(gatherRows(raw, orig, n, d, points));
  s(sync());
  return s.code;
}

int DevicePoints::adopt(const float *devicePoints, int n, int d, bool spatial) {
  if (!spatial) {
    external = devicePoints;
    return 0;
  }
  const size_t count = (size_t)n * d;
  std::vector<float> host(count);
  Status s;
  s(download(host.data(), devicePoints, count * sizeof(float)));
  if (!s.ok())
    return s.code;
  order = spatialOrder(host.data(), n, d);
  s(points.alloc(count));
  s(orig.alloc(n));
  s(cuhc::upload(orig, order.data(), n * sizeof(int)));
  if (s.ok())
    s(gatherRows(devicePoints, orig, n, d, points));
  s(sync());
  return s.code;
}

int boruvkaMst(const float *X, int n, int d, MetricSpec m, WeightSpec w, const MstOptions &opt,
               std::vector<Edge> &edges) {
  edges.clear();
  if (n < 2)
    return 0;
  edges.reserve(n - 1);
  std::vector<int> comp(n);
  std::iota(comp.begin(), comp.end(), 0);
  std::vector<unsigned long long> nearest(n);
  UnionFind sets(n); // over original indices
  auto origOf = [&](int p) { return opt.hostOrig ? opt.hostOrig[p] : p; };

  DevArray<int> dComp;
  DevArray<unsigned long long> dNearest;
  Status s;
  s(dComp.alloc(n));
  s(dNearest.alloc(n));
  MixedPass *mixed = nullptr;
  if (s.ok() && opt.mixed) {
    int st = 0;
    mixed = mixedPassCreate(X, n, d, &st);
    s(st);
  }

  int numComponents = n;
  bool pruneThisPass = opt.prune;
  std::vector<Edge> shortest;
  std::vector<int> label(n);
  while (s.ok() && numComponents > 1) {
    if (mixed) {
      s(mixedPassRun(mixed, comp.data(), numComponents, m.id == kSqEuclidean,
                     numComponents == n ? nullptr : nearest.data(), dNearest));
    } else {
      s(upload(dComp, comp.data(), n * sizeof(int)));
      s(fill(dNearest, 0xff, n * sizeof(unsigned long long)));
      if (!s.ok())
        break;
      if (pruneThisPass) {
        double share = 1.0;
        s(nearestOtherWalk(X, n, d, m, w, dComp, opt.devOrig, dNearest, &share, opt.stats));
        pruneThisPass = share <= kPruneGiveUpShare;
      } else {
        s(nearestOtherSymmetric(X, n, d, m, w, dComp, opt.devOrig, dNearest));
      }
    }
    s(download(nearest.data(), dNearest, n * sizeof(unsigned long long)));
    if (!s.ok())
      break;
    {
      PhaseTimer timer(kPhaseContract);
      // Shortest edge leaving each component, in original indices (every
      // pass packs the other endpoint's original index).
      shortest.assign(numComponents, Edge{-1, -1, 0.0});
      for (int p = 0; p < n; ++p) {
        if (nearest[p] == kEmptyKey)
          continue;
        const int a = origOf(p), b = keyIndex(nearest[p]);
        const Edge e{std::min(a, b), std::max(a, b), (double)keyWeight(nearest[p])};
        Edge &best = shortest[comp[p]];
        if (best.u < 0 || edgeLess(e, best))
          best = e;
      }
      const size_t before = edges.size();
      for (const Edge &e : shortest)
        if (e.u >= 0 && sets.unite(e.u, e.v) >= 0)
          edges.push_back(e);
      if (edges.size() == before) {
        s(kErrDisconnected);
        break;
      }
      // Relabel components 0..k-1 in order of first appearance by position.
      std::fill(label.begin(), label.end(), -1);
      numComponents = 0;
      for (int p = 0; p < n; ++p) {
        const int root = sets.find(origOf(p));
        if (label[root] < 0)
          label[root] = numComponents++;
        comp[p] = label[root];
      }
    }
  }
  if (mixed)
    mixedPassDestroy(mixed);
  return s.code;
}

int singleLinkageMst(const DevicePoints &pts, int n, int d, MetricSpec m, bool prune, bool mixed,
                     PassStats *stats, double *tree) {
  MstOptions opt;
  opt.prune = prune && pts.ordered();
  opt.hostOrig = pts.ordered() ? pts.order.data() : nullptr;
  opt.devOrig = pts.ordered() ? pts.orig.get() : nullptr;
  opt.stats = stats;
  opt.mixed = mixed && !pts.ordered() && (m.id == kEuclidean || m.id == kSqEuclidean) &&
              mixedPassSupports(d);
  std::vector<Edge> edges;
  int status = boruvkaMst(pts.data(), n, d, m, WeightSpec{}, opt, edges);
  if (status == 0)
    status = treeFromMst(std::move(edges), n, tree);
  return status;
}

int hdbscanTree(const DevicePoints &pts, int n, int d, MetricSpec m, int minSamples, float alpha,
                int coreMethod, bool prune, PassStats *stats, double *tree, float *hostCore) {
  DevArray<float> core;
  Status s;
  s(core.alloc(n));
  if (s.ok())
    s(coreDistances(pts.data(), n, d, m, minSamples, core, coreMethod,
                    prune && pts.ordered()));
  MstOptions opt;
  opt.prune = prune && pts.ordered();
  opt.hostOrig = pts.ordered() ? pts.order.data() : nullptr;
  opt.devOrig = pts.ordered() ? pts.orig.get() : nullptr;
  opt.stats = stats;
  WeightSpec w;
  w.kind = kWeightMReach;
  w.core = core;
  w.alpha = alpha;
  std::vector<Edge> edges;
  if (s.ok())
    s(boruvkaMst(pts.data(), n, d, m, w, opt, edges));
  if (s.ok() && hostCore) {
    std::vector<float> c(n);
    s(download(c.data(), core, n * sizeof(float)));
    for (int i = 0; i < n; ++i)
      hostCore[pts.ordered() ? pts.order[i] : i] = c[i];
  }
  if (s.ok())
    s(treeFromMst(std::move(edges), n, tree));
  return s.code;
}

} // namespace cuhc
