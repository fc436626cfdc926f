// Boruvka's MST over the complete graph, matrix-free (reference
// src/mst.cuh:83-186), plus the single-linkage and HDBSCAN drivers
// (src/hdbscan.cuh:353-391, src/capi.cu:68-119).
//
// Each round: the GPU finds, for every point, its nearest point in another
// component (one all-pairs pass), then contracts (contract.hip): each
// component's smallest edge under the total order (weight, u, v) in original
// indices joins it to its neighbour and the components are relabelled, all in
// device memory. Every chosen edge is the minimum edge leaving its component under
// a strict total order, so the result is the unique MST under that order:
// independent of the pass's tile order, pruning, or point permutation.
#include "mst.hpp"

#include <algorithm>
#include <numeric>

namespace cuhc {

// Both entry paths check the coordinates first (kErrNonFinite for NaN or
// infinity): the FP16 candidate kernels are compiled without NaN semantics and
// the host k-d order needs ordered comparisons.
int DevicePoints::upload(const float *hostPoints, int n, int d, bool spatial) {
  const size_t count = (size_t)n * d;
  Status s;
  s(points.alloc(count));
  if (!s.ok())
    return s.code;
  if (!spatial) {
    s(cuhc::upload(points, hostPoints, count * sizeof(float)));
    s(checkFinite(points, count));
    return s.code;
  }
  DevArray<float> raw;
  s(raw.alloc(count));
  s(orig.alloc(n));
  s(cuhc::upload(raw, hostPoints, count * sizeof(float)));
  s(checkFinite(raw, count));
  if (!s.ok())
    return s.code;
  order = spatialOrder(hostPoints, n, d);
  host = hostPoints;
  s(cuhc::upload(orig, order.data(), n * sizeof(int)));
  if (s.ok())
    s(gatherRows(raw, orig, n, d, points));
  s(sync());
  return s.code;
}

// The caller may have written devicePoints on another stream of this HIP
// runtime: the device is synchronised before the library stream reads them.
int DevicePoints::adopt(const float *devicePoints, int n, int d, bool spatial) {
  const size_t count = (size_t)n * d;
  Status s;
  s((int)hipDeviceSynchronize());
  s(checkFinite(devicePoints, count));
  if (!s.ok())
    return s.code;
  if (!spatial) {
    external = devicePoints;
    return 0;
  }
  hostCopy.resize(count);
  s(download(hostCopy.data(), devicePoints, count * sizeof(float)));
  if (!s.ok())
    return s.code;
  host = hostCopy.data();
  order = spatialOrder(host, n, d);
  s(points.alloc(count));
  s(orig.alloc(n));
  s(cuhc::upload(orig, order.data(), n * sizeof(int)));
  if (s.ok())
    s(gatherRows(devicePoints, orig, n, d, points));
  s(sync());
  return s.code;
}

int boruvkaMst(const float *X, int n, int d, MetricSpec m, WeightSpec w, const MstOptions &opt,
               Contraction &ct) {
  Status s;
  s(ct.alloc(n));
  if (!s.ok() || n < 2)
    return s.code;
  DevArray<int> dComp, dPosOf, dSizes, dSkip;
  DevArray<unsigned int> dCompBest;
  DevArray<unsigned long long> dNearest;
  s(dComp.alloc(n));
  s(dNearest.alloc(n));
  if (opt.devOrig) {
    s(dPosOf.alloc(n));
    if (s.ok())
      s(inverseOrder(opt.devOrig, n, dPosOf));
  }
  // Pruned passes bound each component's search by its best known edge, and
  // skip the rows of a component holding more than half of the points: it
  // proposes no edge in that round and joins through the others' smallest
  // edges (Boruvka only needs every proposed edge to be its component's
  // smallest; every other component still proposes).
  if (opt.prune) {
    s(dCompBest.alloc(n));
    s(dSizes.alloc(n));
    s(dSkip.alloc(1));
  }
  // Rows settled from the learned candidate lists (mutual reachability,
  // alpha = 1) are closed to the walk.
  const bool useLists = opt.lists && opt.prune && w.kind == kWeightMReach && w.alpha == 1.0f;
  // Without lists, a pruned mutual-reachability run has the first round's walk
  // keep each row's 4 smallest keys over all other points, and settles the
  // later rounds' rows from them.
  const bool keepKeys = !useLists && opt.prune && w.kind == kWeightMReach;
  DevArray<unsigned char> dClosed;
  DevArray<unsigned long long> dListCounters, dKept;
  if (useLists || keepKeys)
    s(dClosed.alloc(n));
  if (keepKeys)
    s(dKept.alloc((size_t)n * 4));
  if (useLists) {
    s(dListCounters.alloc(64 * 16));
    if (s.ok())
      s(fill(dListCounters, 0, dListCounters.bytes()));
  }
  bool keptValid = false; // dKept holds the first round's lists
  if (s.ok())
    s(resetComponents(dComp, n));
  MixedPass *mixed = nullptr;
  if (s.ok() && opt.mixed) {
    int st = 0;
    mixed = mixedPassCreate(X, n, d, &st);
    s(st);
  }
  const int *posOf = opt.devOrig ? dPosOf.get() : nullptr;
  int numComponents = n;
  bool pruneThisPass = opt.prune;
  bool haveResult = false; // dNearest holds the last round's exact result
  while (s.ok() && numComponents > 1) {
    if (mixed) {
      s(mixedPassRun(mixed, dComp, numComponents, m.id == kSqEuclidean, !haveResult, dNearest));
    } else {
      if (pruneThisPass)
        s(seedComponentBest(dNearest, dComp, posOf, n, numComponents, haveResult, dCompBest));
      s(fill(dNearest, 0xff, n * sizeof(unsigned long long)));
      if (!s.ok())
        break;
      if (pruneThisPass) {
        double share = 1.0;
        s(majorityComponent(dComp, n, numComponents, dSizes, dSkip));
        if (useLists)
          s(settleFromLists(*opt.lists, n, dComp, opt.devOrig, w.core, dSkip, dNearest,
                            dCompBest, dClosed, dListCounters));
        if (keepKeys && !keptValid) {
          s(nearestOtherWalk(X, n, d, m, w, dComp, opt.devOrig, nullptr, dNearest, &share,
                             opt.stats, 1, dSkip, nullptr, dKept));
          keptValid = true;
        } else {
          if (keepKeys)
            s(settleFromKept(dKept, n, dComp, posOf, dSkip, dNearest, dCompBest, dClosed));
          s(nearestOtherWalk(X, n, d, m, w, dComp, opt.devOrig, dCompBest, dNearest, &share,
                             opt.stats, 1, dSkip,
                             useLists || keepKeys ? dClosed.get() : nullptr));
        }
        pruneThisPass = share <= kPruneGiveUpShare;
      } else {
        s(nearestOtherSymmetric(X, n, d, m, w, dComp, opt.devOrig, dNearest));
      }
    }
    if (!s.ok())
      break;
    haveResult = true;
    PhaseTimer timer(kPhaseContract);
    int next = numComponents, added = 0;
    s(contractRound(ct, dNearest, dComp, opt.devOrig, posOf, numComponents, &next, &added));
    if (s.ok() && added == 0)
      s(kErrDisconnected);
    numComponents = next;
  }
  if (mixed)
    mixedPassDestroy(mixed);
  if (useLists && s.ok()) {
    std::vector<unsigned long long> ctr(64 * 16);
    s(download(ctr.data(), dListCounters, dListCounters.bytes()));
    LearnedStats st;
    for (int k = 0; k < 64; ++k) {
      st.settled += ctr[k * 16];
      st.considered += ctr[k * 16 + 1];
    }
    if (s.ok())
      learnedStatsAdd(st);
  }
  return s.code;
}

int singleLinkageMst(const DevicePoints &pts, int n, int d, MetricSpec m, bool prune, bool mixed,
                     PassStats *stats, double *tree) {
  MstOptions opt;
  opt.prune = prune && pts.ordered();
  opt.devOrig = pts.ordered() ? pts.orig.get() : nullptr;
  opt.stats = stats;
  opt.mixed = mixed && !pts.ordered() && (m.id == kEuclidean || m.id == kSqEuclidean) &&
              mixedPassSupports(d);
  Contraction ct;
  int status = boruvkaMst(pts.data(), n, d, m, WeightSpec{}, opt, ct);
  if (status == 0)
    status = treeFromDeviceEdges(ct.edges, n, tree);
  return status;
}

namespace {

// Spatial order stratified by core distance: the points above the 90th
// percentile of core distances (noise, in sparse regions) get their own k-d
// order after the others'. A pruned pass skips a column tile when its bound
// exceeds the largest current best of the row tile, and one sparse point's
// large mutual-reachability weight would otherwise hold that bound up for the
// dense points it shares a tile with. Returns the new order (original
// indices) and new position -> old position, or empty vectors when one
// class would be empty.
void coreStratifiedOrder(const DevicePoints &pts, int n, int d, const std::vector<float> &core,
                         std::vector<int> &order, std::vector<int> &newToOld) {
  std::vector<float> sorted(core);
  const size_t q = (size_t)(0.9 * n);
  std::nth_element(sorted.begin(), sorted.begin() + q, sorted.end());
  const float threshold = sorted[q];
  std::vector<int> low, high; // original indices
  for (int p = 0; p < n; ++p)
    (core[p] > threshold ? high : low).push_back(pts.order[p]);
  if (low.size() < 128 || high.size() < 128)
    return;
  spatialOrderIndices(pts.host, d, low.data(), (int)low.size());
  spatialOrderIndices(pts.host, d, high.data(), (int)high.size());
  order = low;
  order.insert(order.end(), high.begin(), high.end());
  std::vector<int> oldPos(n);
  for (int p = 0; p < n; ++p)
    oldPos[pts.order[p]] = p;
  newToOld.resize(n);
  for (int p = 0; p < n; ++p)
    newToOld[p] = oldPos[order[p]];
}

} // namespace

int hdbscanTree(const DevicePoints &pts, int n, int d, MetricSpec m, int minSamples, float alpha,
                int coreMethod, bool prune, PassStats *stats, double *tree, float *hostCore) {
  DevArray<float> core;
  Status s;
  s(core.alloc(n));
  // The learned method's candidate lists also settle Boruvka rows (pruned
  // passes, mutual reachability with alpha = 1, a box metric).
  RadiusLists lists;
  const bool keepLists = coreMethod == kCoreLearned && prune && pts.ordered() && alpha == 1.0f &&
                         metricHasBoxBound(m.id);
  if (s.ok()) {
    if (keepLists)
      s(coreLearned(pts.data(), n, d, m, minSamples, core, true, &lists));
    else
      s(coreDistances(pts.data(), n, d, m, minSamples, core, coreMethod,
                      prune && pts.ordered()));
  }
  // Pruned passes: reorder by core-distance class (the tree does not depend
  // on the order).
  const float *X = pts.data();
  const int *devOrig = pts.ordered() ? pts.orig.get() : nullptr;
  const std::vector<int> *order = &pts.order;
  std::vector<int> stratified, newToOld;
  DevArray<float> sPoints, sCore;
  DevArray<int> sOrig, sMap;
  if (s.ok() && prune && pts.ordered() && pts.host && n >= 1024) {
    std::vector<float> hc(n);
    s(download(hc.data(), core, n * sizeof(float)));
    if (s.ok())
      coreStratifiedOrder(pts, n, d, hc, stratified, newToOld);
    if (s.ok() && !stratified.empty()) {
      s(sPoints.alloc((size_t)n * d));
      s(sCore.alloc(n));
      s(sOrig.alloc(n));
      s(sMap.alloc(n));
      s(upload(sOrig, stratified.data(), n * sizeof(int)));
      s(upload(sMap, newToOld.data(), n * sizeof(int)));
      if (s.ok()) {
        s(gatherRows(X, sMap, n, d, sPoints));
        s(gatherRows(core, sMap, n, 1, sCore));
      }
      if (s.ok()) {
        X = sPoints;
        devOrig = sOrig;
        order = &stratified;
      }
      if (s.ok() && keepLists) {
        // List rows stay in the old order (rowMap = new -> old); partners
        // are renumbered to the new positions.
        DevArray<int> oldToNew;
        s(oldToNew.alloc(n));
        if (s.ok())
          s(inverseOrder(sMap, n, oldToNew));
        if (s.ok())
          s(remapRadiusLists(lists, n, oldToNew));
        lists.rowMap = sMap;
        s(sync());
      }
    }
  }
  MstOptions opt;
  opt.prune = prune && pts.ordered();
  opt.devOrig = devOrig;
  opt.stats = stats;
  opt.lists = keepLists && lists.pos.get() ? &lists : nullptr;
  WeightSpec w;
  w.kind = kWeightMReach;
  w.core = stratified.empty() ? core.get() : sCore.get();
  w.alpha = alpha;
  Contraction ct;
  if (s.ok())
    s(boruvkaMst(X, n, d, m, w, opt, ct));
  if (s.ok() && hostCore) {
    std::vector<float> c(n);
    s(download(c.data(), w.core, n * sizeof(float)));
    for (int i = 0; i < n; ++i)
      hostCore[pts.ordered() ? (*order)[i] : i] = c[i];
  }
  if (s.ok())
    s(treeFromDeviceEdges(ct.edges, n, tree));
  return s.code;
}

} // namespace cuhc
