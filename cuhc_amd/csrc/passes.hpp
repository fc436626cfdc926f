// Host entry points of the all-pairs kernels (distance matrices, nearest
// other component passes, spatial pruning). All pointers are device pointers
// unless noted; everything runs on cuhc::stream().
#pragma once

#include "common.hpp"
#include "runtime.hpp"

#include <vector>

namespace cuhc {

struct MetricSpec {
  int id;
  float p; // Minkowski order
};

enum WeightKind : int { kWeightPlain = 0, kWeightMReach = 1, kWeightWard = 2 };

struct WeightSpec {
  int kind = kWeightPlain;
  const float *core = nullptr; // device, for kWeightMReach
  float alpha = 1.0f;
  const int *size = nullptr; // device, for kWeightWard
};

bool validMetric(int id);
bool metricHasBoxBound(int id);
int tilePoints(int metricId); // tile side of the metric's all-pairs kernels (128 or 64)

// out[i * nB + j] = metric(A[i], B[j]).
int pairwiseMatrix(const float *A, int nA, const float *B, int nB, int d, MetricSpec m,
                   float *out);

// Condensed upper triangle: out[i*n - i*(i+1)/2 + (j - i - 1)] = metric(X[i], X[j]), i < j.
int condensedMatrix(const float *X, int n, int d, MetricSpec m, float *out);

// nearest[i] = min over j with comp[j] != comp[i] of packKey(weight(i, j), id(j)),
// combined by atomicMin, so nearest[] must start at kEmptyKey. Each unordered
// pair is computed once (tiles I <= J). id(j) = orig[j] (device) for points in
// spatial order, else j (orig == nullptr).
int nearestOtherSymmetric(const float *X, int n, int d, MetricSpec m, WeightSpec w,
                          const int *comp, const int *orig, unsigned long long *nearest);

// ---------------------------------------------------------------------------
// Spatial pruning (reference src/spatial.cuh)

constexpr int kPruneMinPoints = 16384;
constexpr int kPruneMaxFeatures = 64;
constexpr int kPruneHardMaxFeatures = 256;
constexpr double kPruneGiveUpShare = 0.5;
enum PruneMode : int { kPruneAuto = 0, kPruneAlways = 1, kPruneNever = 2 };

bool shouldPrune(int mode, int metricId, int n, int d);

// k-d order on the host: order[i] = original index of the point at position i.
std::vector<int> spatialOrder(const float *hostPoints, int n, int d);
// The same order of the listed points (indices into hostPoints), in place.
void spatialOrderIndices(const float *hostPoints, int d, int *idx, int count);

// dst[i] = src[order[i]] (rows of d floats).
int gatherRows(const float *src, const int *order, int n, int d, float *dst);

struct PassStats {
  unsigned long long tilesComputed = 0, tilesTotal = 0;
};

// Pruned pass for points in spatial order: same contract as
// nearestOtherSymmetric, except that keys carry orig[j] (device array; null:
// positions, as for rnn cluster slots), so ties break by original index
// exactly as in an unpruned run. compBest (nullable,
// one u32 per component, seedComponentBest) holds the weight bits of the best
// edge known for each component; rows stop searching above it, so a row's
// result is exact only when it holds its component's smallest edge, which is
// all the contraction uses. Returns the share of tile pairs computed in
// *share. rowTileStride > 1 runs every stride-th row tile only (sampling).
// skipComp (nullable, device): rows of that component (-1: none) search
// nothing and keep their nearest[] entries; so do rows p with closedRows[p]
// set (nullable, device), whose exact answers are already in nearest[].
// kept (nullable, device, n x 4 keys by position; mutual reachability, no
// closed rows, no sampling): every row also keeps its 4 smallest keys over
// the other components (ascending, empty keys ~0 padded), searched without
// the component bounds, for settling rows in later rounds.
int nearestOtherWalk(const float *X, int n, int d, MetricSpec m, WeightSpec w, const int *comp,
                     const int *orig, unsigned int *compBest, unsigned long long *nearest,
                     double *share, PassStats *stats, int rowTileStride = 1,
                     const int *skipComp = nullptr, const unsigned char *closedRows = nullptr,
                     unsigned long long *kept = nullptr);

// Rows settled from the 4 keys each row kept in an earlier round (kept, by
// position; keys carry original indices, posOf maps them to positions, null:
// identity): a row whose list holds a point outside its component now takes
// the first such key as its exact nearest (the list holds the row's smallest
// keys over all other points, so every point outside the component that it
// misses weighs more), lowers its component's best known weight and is
// closed; the others, and rows of skipComp, stay open.
int settleFromKept(const unsigned long long *kept, int n, const int *comp, const int *posOf,
                   const int *skipComp, unsigned long long *nearest, unsigned int *compBest,
                   unsigned char *closed);

// *majority (device) = the component holding more than half of the n points,
// or -1. sizes: numComp ints of scratch.
int majorityComponent(const int *comp, int n, int numComp, int *sizes, int *majority);

// compBest[c] for c < numComp: "none", then (fromNearest) the weight bits of
// the previous pass's candidates whose partner is still in another component.
// posOf maps an original index to its position (null: identity).
int seedComponentBest(const unsigned long long *nearest, const int *comp, const int *posOf,
                      int n, int numComp, bool fromNearest, unsigned int *compBest);

// posOf[orig[p]] = p.
int inverseOrder(const int *orig, int n, int *posOf);

// Share of tiles a pruned pass would compute, from a sample of row tiles
// (Boruvka's first pass: plain weights, every point its own component).
int estimatePrunedShare(const float *X, int n, int d, MetricSpec m, double *share);

} // namespace cuhc
