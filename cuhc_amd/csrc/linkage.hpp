// Dense Lance-Williams linkage and Ward's matrix-free algorithms.
#pragma once

#include "host_tree.hpp"
#include "passes.hpp"

namespace cuhc {

enum LinkageId : int { kSingle = 0, kComplete = 1, kAverage = 2, kWard = 3 };
enum AlgorithmId : int { kAlgDense = 0, kAlgMst = 1, kAlgNnChain = 2, kAlgRnn = 3 };

// cuhc_linkage_device's body (capi.hip): validation, then the algorithm, on
// points already in device memory.
int linkageFromDevicePoints(const float *X, int n, int d, int metricId, float metricParam,
                            int linkageId, int algorithmId, double *tree);

// Full merge tree from the points through the condensed distance matrix
// (n (n - 1) / 2 floats in HBM) and the Lance-Williams updates. tree: host,
// (n - 1) x 4. Returns hipErrorOutOfMemory (2) if the matrix does not fit.
int denseLinkage(const float *X, int n, int d, MetricSpec m, int linkage, double *tree);

// Merge tree from a condensed matrix already on the device (overwritten).
// initialSize / initialId: per-slot starting size and tree id (device, null =
// 1 and slot); firstNewId: id of the cluster made by merge 0.
int denseLinkageOnCondensed(float *D, int n, int linkage, const int *initialSize,
                            const int *initialId, int firstNewId, double *tree);

struct RnnOptions {
  float minProgress = 0.01f;
  int denseBelow = 0;
  bool allowDense = true;
  // orig (device, nullable): position -> original index of reordered points;
  // merges are recorded with original indices. prune (points in k-d order):
  // the nearest-neighbour passes walk the cluster tiles with box pruning
  // until a pass computes more than half of the tile pairs. stats (nullable)
  // receives the passes' tile counts.
  const int *orig = nullptr;
  bool prune = false;
  PassStats *stats = nullptr;
  // The first pass (every cluster a single point, d <= 64) through the FP16
  // MFMA candidates and the FP32 recheck: the Ward weight of two points is
  // exactly their FP32 squared distance, so the keys equal the FP32 pass's.
  bool mixedFirstPass = false;
};

// Ward through rounds of reciprocal nearest neighbours (FP64 centroids and
// heights, FP32 nearest-neighbour passes), with dense or chain tails.
int rnnWard(const float *X, int n, int d, const RnnOptions &opt, double *tree);

// Ward through the nearest-neighbour chain (one cooperative kernel).
int nnChainWard(const float *X, int n, int d, double *tree);

// The chain on m starting clusters (FP64 centroids m x d, sizes): merges in
// chain order, named by slot.
int nnChainMerges(const double *centroids, const int *sizes, int m, int d,
                  std::vector<Edge> &merges);

} // namespace cuhc
