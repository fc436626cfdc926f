// extern "C" boundary of libcuhc.so (declared in include/cuhc.h).
//
// Same contract as reference src/capi.cu:133-366: host arrays in and out,
// blocking, device 0, the first error wins, process-global prune mode and
// statistics (guarded by a mutex here; semantics unchanged).
#include "../../include/cuhc.h"
#include "linkage.hpp"
#include "mst.hpp"

#include <algorithm>
#include <atomic>
#include <mutex>

using namespace cuhc;

namespace {

std::atomic<int> gPruneMode{kPruneAuto};
std::atomic<int> gPrecision{0};
std::mutex gStatsMutex;
PassStats gPruneStats;

bool validLinkageAlgorithm(int linkageId, int algorithmId, int metricId) {
  return algorithmId == kAlgDense || (algorithmId == kAlgMst && linkageId == kSingle) ||
         ((algorithmId == kAlgNnChain || algorithmId == kAlgRnn) && linkageId == kWard &&
          metricId == kEuclidean);
}

// Whether the Boruvka passes use FP16 MFMA candidates + FP32 recheck.
bool useMixed(int metricId, int d) {
  if (metricId != kEuclidean && metricId != kSqEuclidean)
    return false;
  switch (gPrecision.load()) {
  case 1: return false;
  case 2: return true;
  default: return d >= 16;
  }
}

// Decides pruning for points in spatial order (auto mode samples a pruned
// pass, reference src/capi.cu:101-119; here the sampled row tiles compute
// their first 64 walk steps and count the rest against the bound reached).
int decidePruning(const DevicePoints &pts, int n, int d, MetricSpec m, bool &prune) {
  if (!prune || gPruneMode.load() == kPruneAlways)
    return 0;
  double share = 1.0;
  const int status = estimatePrunedShare(pts.data(), n, d, m, &share);
  prune = status == 0 && share <= kPruneGiveUpShare;
  return status;
}

// Points from which auto picks the learned radius for k > 64: below, the
// radix select's full rows cost less than the network's fixed training.
constexpr int kLearnedAutoMinPoints = 131072;
// Largest k auto hands to it: its lists hold at most 4,096 candidates per
// row, and a proposal needs k of them inside the radius with room to spare.
constexpr int kLearnedAutoMaxK = 128;

int linkageImpl(const float *X, bool onDevice, int n, int d, int metricId, float metricParam,
                int linkageId, int algorithmId, double *tree) {
  if (n <= 0 || d <= 0)
    return kErrInvalidShape;
  if (linkageId < kSingle || linkageId > kWard)
    return kErrInvalidLinkage;
  if (!validLinkageAlgorithm(linkageId, algorithmId, metricId))
    return kErrInvalidAlgorithm;
  if (!validMetric(metricId))
    return kErrInvalidMetric;
  if (n == 1)
    return 0;
  const MetricSpec m{metricId, metricParam};
  Status s;
  if (algorithmId == kAlgMst) {
    const bool mixed = useMixed(metricId, d);
    bool prune = !mixed && shouldPrune(gPruneMode.load(), metricId, n, d);
    DevicePoints pts;
    s(onDevice ? pts.adopt(X, n, d, prune) : pts.upload(X, n, d, prune));
    if (s.ok())
      s(decidePruning(pts, n, d, m, prune));
    PassStats stats;
    if (s.ok())
      s(singleLinkageMst(pts, n, d, m, prune, mixed, &stats, tree));
    std::lock_guard<std::mutex> lock(gStatsMutex);
    gPruneStats.tilesComputed += stats.tilesComputed;
    gPruneStats.tilesTotal += stats.tilesTotal;
    return s.code;
  }
  if (algorithmId == kAlgRnn) {
    // Ward is euclidean only; its passes work on squared distances
    // (reference src/capi.cu:210-222: pruned when the points allow it).
    const MetricSpec sq{kSqEuclidean, 2.0f};
    bool prune = shouldPrune(gPruneMode.load(), kSqEuclidean, n, d);
    DevicePoints pts;
    s(onDevice ? pts.adopt(X, n, d, prune) : pts.upload(X, n, d, prune));
    if (s.ok())
      s(decidePruning(pts, n, d, sq, prune));
    PassStats stats;
    RnnOptions opt;
    opt.prune = prune && pts.ordered();
    opt.orig = pts.ordered() ? pts.orig.get() : nullptr;
    opt.stats = &stats;
    opt.mixedFirstPass = useMixed(kEuclidean, d) && mixedPassSupports(d);
    if (s.ok())
      s(rnnWard(pts.data(), n, d, opt, tree));
    std::lock_guard<std::mutex> lock(gStatsMutex);
    gPruneStats.tilesComputed += stats.tilesComputed;
    gPruneStats.tilesTotal += stats.tilesTotal;
    return s.code;
  }
  DevicePoints pts;
  s(onDevice ? pts.adopt(X, n, d, false) : pts.upload(X, n, d, false));
  if (!s.ok())
    return s.code;
  if (algorithmId == kAlgNnChain)
    return nnChainWard(pts.data(), n, d, tree);
  return denseLinkage(pts.data(), n, d, m, linkageId, tree);
}

int hdbscanImpl(const float *X, bool onDevice, int n, int d, int metricId, float metricParam,
                int minSamples, float alpha, int coreMethod, double *tree, float *core) {
  if (n <= 0 || d <= 0 || minSamples < 1 || minSamples > n)
    return kErrInvalidShape;
  if (coreMethod < 0 || coreMethod > 3 || (coreMethod == kCoreFused && minSamples > kMaxFusedK))
    return kErrInvalidAlgorithm;
  if (!validMetric(metricId))
    return kErrInvalidMetric;
  const MetricSpec m{metricId, metricParam};
  // auto: the fused top-k up to k = 64 (reference src/hdbscan.cuh:324); above
  // it the learned radius for box metrics on large inputs (its proposals are
  // verified exactly, and it computes a small share of the pairs the
  // full-row radix select does), else the radix select. Every method gives
  // the same core distances.
  int method = coreMethod;
  if (method == kCoreAuto)
    method = minSamples <= kMaxFusedK ? kCoreFused
             : (metricHasBoxBound(metricId) && n >= kLearnedAutoMinPoints &&
                minSamples <= kLearnedAutoMaxK)
                 ? kCoreLearned
                 : kCoreRadix;
  bool prune = shouldPrune(gPruneMode.load(), metricId, n, d);
  // The learned radii are predicted from k-d tiles: spatial order pays for it
  // at any size (the results are the same in every order).
  const bool spatial = prune || (method == kCoreLearned && n > 1 &&
                                 gPruneMode.load() != kPruneNever && metricHasBoxBound(metricId));
  DevicePoints pts;
  Status s;
  s(onDevice ? pts.adopt(X, n, d, spatial) : pts.upload(X, n, d, spatial));
  if (s.ok())
    s(decidePruning(pts, n, d, m, prune));
  PassStats stats;
  if (s.ok()) {
    if (n == 1) {
      if (core)
        core[0] = 0.0f;
    } else {
      s(hdbscanTree(pts, n, d, m, minSamples, alpha, method, prune, &stats, tree, core));
    }
  }
  std::lock_guard<std::mutex> lock(gStatsMutex);
  gPruneStats.tilesComputed += stats.tilesComputed;
  gPruneStats.tilesTotal += stats.tilesTotal;
  return s.code;
}

} // namespace

int cuhc::linkageFromDevicePoints(const float *X, int n, int d, int metricId, float metricParam,
                                  int linkageId, int algorithmId, double *tree) {
  return linkageImpl(X, true, n, d, metricId, metricParam, linkageId, algorithmId, tree);
}

extern "C" {

int cuhc_pairwise_distances(const float *A, int nA, const float *B, int nB, int d, int metricId,
                            float metricParam, float *distances) {
  if (nA <= 0 || nB <= 0 || d <= 0)
    return kErrInvalidShape;
  if (!validMetric(metricId))
    return kErrInvalidMetric;
  const bool same = B == nullptr;
  DevArray<float> dA, dB, dOut;
  Status s;
  s(dA.alloc((size_t)nA * d));
  s(upload(dA, A, (size_t)nA * d * sizeof(float)));
  if (!same) {
    s(dB.alloc((size_t)nB * d));
    s(upload(dB, B, (size_t)nB * d * sizeof(float)));
  }
  s(dOut.alloc((size_t)nA * nB));
  if (s.ok())
    s(pairwiseMatrix(dA, nA, same ? dA.get() : dB.get(), nB, d, MetricSpec{metricId, metricParam},
                     dOut));
  if (s.ok())
    s(download(distances, dOut, (size_t)nA * nB * sizeof(float)));
  return s.code;
}

int cuhc_pairwise_distances_device(const float *A, int nA, const float *B, int nB, int d,
                                   int metricId, float metricParam, float *distances) {
  if (nA <= 0 || nB <= 0 || d <= 0)
    return kErrInvalidShape;
  if (!validMetric(metricId))
    return kErrInvalidMetric;
  Status s;
  s((int)hipDeviceSynchronize()); // the caller's writes may come from any stream
  if (s.ok())
    s(pairwiseMatrix(A, nA, B ? B : A, nB, d, MetricSpec{metricId, metricParam}, distances));
  s(sync());
  return s.code;
}

int cuhc_feature_ranges_device(const float *X, int n, int d, float *low, float *high) {
  if (n <= 0 || d <= 0)
    return kErrInvalidShape;
  Status s;
  s((int)hipDeviceSynchronize());
  if (s.ok())
    s(featureRanges(X, n, d, low, high));
  return s.code;
}

int cuhc_linkage(const float *X, int n, int d, int metricId, float metricParam, int linkageId,
                 int algorithmId, double *tree) {
  return linkageImpl(X, false, n, d, metricId, metricParam, linkageId, algorithmId, tree);
}

int cuhc_linkage_device(const float *deviceX, int n, int d, int metricId, float metricParam,
                        int linkageId, int algorithmId, double *tree) {
  return linkageImpl(deviceX, true, n, d, metricId, metricParam, linkageId, algorithmId, tree);
}

int cuhc_ward_rnn_with_options(const float *X, int n, int d, float minProgress, int denseBelow,
                               int allowDense, double *tree) {
  if (n <= 0 || d <= 0)
    return kErrInvalidShape;
  DevicePoints pts;
  Status s;
  s(pts.upload(X, n, d, false));
  RnnOptions opt;
  opt.minProgress = minProgress;
  opt.denseBelow = denseBelow;
  opt.allowDense = allowDense != 0;
  opt.mixedFirstPass = useMixed(kEuclidean, d) && mixedPassSupports(d);
  if (s.ok() && n > 1)
    s(rnnWard(pts.data(), n, d, opt, tree));
  return s.code;
}

int cuhc_hdbscan_tree(const float *X, int n, int d, int metricId, float metricParam,
                      int minSamples, float alpha, int coreMethod, double *tree, float *core) {
  return hdbscanImpl(X, false, n, d, metricId, metricParam, minSamples, alpha, coreMethod, tree,
                     core);
}

int cuhc_hdbscan_tree_device(const float *deviceX, int n, int d, int metricId, float metricParam,
                             int minSamples, float alpha, int coreMethod, double *tree,
                             float *core) {
  return hdbscanImpl(deviceX, true, n, d, metricId, metricParam, minSamples, alpha, coreMethod,
                     tree, core);
}

long long cuhc_condense_tree(const double *tree, long long n, long long minClusterSize,
                             long long capacity, long long *parent, long long *child,
                             double *lambda, long long *size) {
  const std::vector<CondensedRow> rows = condenseTree(tree, n, minClusterSize);
  if ((long long)rows.size() > capacity)
    return -(long long)rows.size();
  for (size_t i = 0; i < rows.size(); ++i) {
    parent[i] = rows[i].parent;
    child[i] = rows[i].child;
    lambda[i] = rows[i].lambda;
    size[i] = rows[i].size;
  }
  return (long long)rows.size();
}

int cuhc_cluster_stability(const long long *parent, const long long *child, const double *lambda,
                           const long long *size, long long rows, double *stability) {
  if (rows < 0)
    return kErrInvalidShape;
  const std::vector<double> st = clusterStability(CondensedView(parent, child, lambda, size, rows));
  std::copy(st.begin(), st.end(), stability);
  return 0;
}

int cuhc_select_clusters(const long long *parent, const long long *child, const double *lambda,
                         const long long *size, long long rows, double *stability, int method,
                         int allowSingleCluster, long long maxClusterSize,
                         unsigned char *selected) {
  if (rows < 0)
    return kErrInvalidShape;
  if (method != 0 && method != 1)
    return kErrInvalidAlgorithm;
  return selectClusters(CondensedView(parent, child, lambda, size, rows), stability, method,
                        allowSingleCluster != 0, maxClusterSize, selected);
}

int cuhc_label_points(const long long *parent, const long long *child, const double *lambda,
                      const long long *size, long long rows, const unsigned char *selected,
                      int allowSingleCluster, double epsilon, long long *labels) {
  if (rows < 0)
    return kErrInvalidShape;
  labelPoints(CondensedView(parent, child, lambda, size, rows), selected, allowSingleCluster != 0,
              epsilon, labels);
  return 0;
}

int cuhc_membership_probabilities(const long long *parent, const long long *child,
                                  const double *lambda, const long long *size, long long rows,
                                  const unsigned char *selected, const long long *labels,
                                  double *probabilities) {
  if (rows < 0)
    return kErrInvalidShape;
  membershipProbabilities(CondensedView(parent, child, lambda, size, rows), selected, labels,
                          probabilities);
  return 0;
}

int cuhc_set_pruning(int mode) {
  if (mode < 0 || mode > 2)
    return kErrInvalidAlgorithm;
  gPruneMode.store(mode);
  return 0;
}

void cuhc_prune_stats(unsigned long long *computed, unsigned long long *total) {
  std::lock_guard<std::mutex> lock(gStatsMutex);
  *computed = gPruneStats.tilesComputed;
  *total = gPruneStats.tilesTotal;
  gPruneStats = PassStats{};
}

const char *cuhc_error_string(int code) {
  switch (code) {
  case kErrInvalidMetric: return "invalid metric id";
  case kErrInvalidShape: return "array dimensions must be positive";
  case kErrInvalidLinkage: return "invalid linkage id";
  case kErrInvalidAlgorithm: return "invalid algorithm id for this linkage";
  case kErrNonFinite: return "distances overflowed float32 or are NaN; rescale the data";
  case kErrDisconnected:
    return "some points have no finite distance to the others (NaN or infinite values?)";
  case kErrWorkspace:
    return "input exceeds the fixed workspace of cuhc_linkage_device_noalloc "
           "(at most 65536 points of at most 64 features)";
  default: return hipGetErrorString((hipError_t)code);
  }
}

int cuhc_set_precision(int mode) {
  if (mode < 0 || mode > 2)
    return kErrInvalidAlgorithm;
  gPrecision.store(mode);
  return 0;
}

int cuhc_device_alloc(unsigned long long bytes, void **ptr) {
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e == hipErrorOutOfMemory) {
    // The library's cache of freed blocks may hold the memory: release it and retry.
    (void)hipGetLastError();
    releaseCachedBlocks();
    e = hipMalloc(ptr, bytes);
  }
  return (int)e;
}

void cuhc_release(void) { releaseLibraryState(); }

int cuhc_device_free(void *ptr) { return ptr ? (int)hipFree(ptr) : 0; }

int cuhc_copy_to_device(void *dst, const void *src, unsigned long long bytes) {
  return upload(dst, src, bytes);
}

int cuhc_copy_to_host(void *dst, const void *src, unsigned long long bytes) {
  return download(dst, src, bytes);
}

int cuhc_copy_device(void *dst, const void *src, unsigned long long bytes) {
  int s = (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream());
  return s ? s : sync();
}

int cuhc_set_timing(int on) {
  setTiming(on != 0);
  return 0;
}

void cuhc_read_timings(double *ms, long long *launches, long long *units, int count) {
  readTimings(ms, launches, units, count);
}

void cuhc_reset_timings(void) { resetTimings(); }

} // extern "C"
