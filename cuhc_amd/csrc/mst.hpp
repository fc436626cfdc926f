// Boruvka minimum spanning tree and the linkage/HDBSCAN drivers built on it.
#pragma once

#include "host_tree.hpp"
#include "passes.hpp"

#include <vector>

namespace cuhc {

// Points on the device, possibly in spatial order (order[i] = original index
// of position i; empty = identity).
struct DevicePoints {
  DevArray<float> points;
  DevArray<int> orig; // device copy of order (only when ordered)
  std::vector<int> order;
  bool ordered() const { return !order.empty(); }
  // Uploads hostPoints (n x d), in spatial order if spatial is set.
  int upload(const float *hostPoints, int n, int d, bool spatial);
  // Adopts device-resident points (no copy unless spatial order is wanted).
  int adopt(const float *devicePoints, int n, int d, bool spatial);
  const float *data() const { return external ? external : points.get(); }
  const float *external = nullptr;
  // Host coordinates in original order while ordered (the caller's array, or
  // hostCopy for device-resident input).
  const float *host = nullptr;
  std::vector<float> hostCopy;
};

// Edges of a spanning forest in device memory (original indices, u < v).
struct DeviceEdges {
  DevArray<int> u, v;
  DevArray<float> w;
  int count = 0;
};

// Device state of Boruvka's contraction (contract.hip).
struct Contraction {
  int n = 0;
  DevArray<unsigned long long> best1;
  DevArray<unsigned int> best2;
  DevArray<int> parent, root, flags, newId, changed, counters;
  DevArray<unsigned char> scanTemp;
  size_t scanBytes = 0;
  DeviceEdges edges;
  int alloc(int points);
};

// comp[p] = p.
int resetComponents(int *comp, int n);

// One contraction: every component's smallest outgoing edge under (weight,
// smaller endpoint, larger endpoint) from nearest[], edges appended to
// ct.edges, comp relabelled to 0..newNumComp-1. orig maps positions to
// original indices and posOf back (both null: identity). *added = edges
// appended this round.
int contractRound(Contraction &ct, const unsigned long long *nearest, int *comp, const int *orig,
                  const int *posOf, int numComp, int *newNumComp, int *added);

// Sorts device edges into the global order and writes the scipy-format tree.
int treeFromDeviceEdges(const DeviceEdges &edges, int n, double *tree);

// The learned method's candidate lists, kept for the Boruvka rounds of
// mutual reachability. List row r (the points' order at the gather) holds
// count[r] entries: every point within radius[r] of point r (itself
// included), as FP32 distances and partner positions; count[r] > cap means
// the list overflowed and is incomplete. rowMap: position in the current
// order -> list row (null: the same order).
struct RadiusLists {
  DevArray<float> dist;   // [n][cap]
  DevArray<int> pos;      // [n][cap]
  DevArray<int> count;    // [n]
  DevArray<float> radius; // [n]
  const int *rowMap = nullptr;
  int cap = 0;
};

struct MstOptions {
  bool prune = false;          // points in spatial order; use the walk pass
  const int *devOrig = nullptr;  // position -> original index (device), null = identity
  PassStats *stats = nullptr;
  bool mixed = false;          // FP16 MFMA candidate pass (plain euclidean weights only)
  // Learned-radius candidate lists (mutual reachability, alpha = 1, pruned
  // passes): rows they settle skip the walk.
  const RadiusLists *lists = nullptr;
};

// Exact MST of the complete graph over the points under the total order
// (weight, smaller original index, larger original index). Edges stay in
// device memory, in original indices. Returns 0, a HIP error, or
// kErrDisconnected.
int boruvkaMst(const float *X, int n, int d, MetricSpec m, WeightSpec w, const MstOptions &opt,
               Contraction &ct);

// Core distances: core[i] = k-th smallest distance from point i (itself = 1st).
// kCoreLearned (an extension): radius proposals from the trained network seed
// the pruned gather and a radix select; rows whose proposal fails fall back to
// the exact full search. All methods give identical values.
enum CoreMethod : int { kCoreAuto = 0, kCoreFused = 1, kCoreRadix = 2, kCoreLearned = 3 };
constexpr int kMaxFusedK = 64;
int coreDistances(const float *X, int n, int d, MetricSpec m, int k, float *core, int method,
                  bool prune);
// The k-th smallest over the first stepLimit walk steps only (own tile, then
// neighbours): a local density feature.
int coreLocal(const float *X, int n, int d, MetricSpec m, int k, int stepLimit, float *out);
// Exact core distances of the listed points (device positions) against all
// points. prune: the points are in k-d order and the list ascends, so tiles
// of listed rows get their own boxes and skip far column tiles (k <= 64).
int coreForRows(const float *X, int n, int d, MetricSpec m, int k, const int *rows, int numRows,
                float *core, bool prune = false);
// Exact core distances of every point of the listed 128-point row tiles
// (host list), with the pruned walk when k <= kMaxFusedK and the metric has a
// box bound, else through coreForRows.
int coreForTiles(const float *X, int n, int d, MetricSpec m, int k, const int *hostTiles,
                 int numTiles, float *core);
// keep (nullable): receives the candidate lists.
int coreLearned(const float *X, int n, int d, MetricSpec m, int k, float *core, bool spatial,
                RadiusLists *keep = nullptr);

// Partners renumbered for a new point order: pos = oldToNew[pos] (device).
int remapRadiusLists(RadiusLists &lists, int n, const int *oldToNew);

// Settles rows of a Boruvka round from the lists (mutual reachability with
// alpha = 1, whose weight max(core_i, core_j, d_ij) is never below d_ij):
// for row p, the smallest key (weight, orig[j]) over the listed partners j in
// other components is its exact nearest-other-component key whenever its
// weight is at most radius[p]: any partner missing from a complete list is
// farther than the radius, so its weight is larger. Such rows get nearest[p]
// = that key, their component's best known weight (compBest, nullable)
// lowered to it, and closed[p] = 1; the others closed[p] = 0. Rows of the
// skipped component (skipComp, nullable) stay open and untouched.
// counters[0] += rows settled, counters[1] += rows considered (device).
int settleFromLists(const RadiusLists &lists, int n, const int *comp, const int *orig,
                    const float *core, const int *skipComp, unsigned long long *nearest,
                    unsigned int *compBest, unsigned char *closed, unsigned long long *counters);

struct LearnedStats {
  unsigned long long proposals = 0, hits = 0, fallbacks = 0, candidates = 0, tiles = 0;
  unsigned long long settled = 0, considered = 0; // Boruvka rows answered from the lists
};
void learnedStatsAdd(const LearnedStats &s);

// The single-linkage tree of mutual reachability (host tree, host core in
// original order).
int hdbscanTree(const DevicePoints &pts, int n, int d, MetricSpec m, int minSamples, float alpha,
                int coreMethod, bool prune, PassStats *stats, double *tree, float *hostCore);

// Plain single linkage through Boruvka.
int singleLinkageMst(const DevicePoints &pts, int n, int d, MetricSpec m, bool prune, bool mixed,
                     PassStats *stats, double *tree);

// FP16 MFMA certified-candidate pass (mfma.hip): for every point of X (n x d,
// original order), the exact nearest point in another component under the
// FP32 euclidean (or squared euclidean) distance, as packKey(weight, j), the
// same value the FP32 tiles give. comp[] (device) are component labels by
// point; nearest[] holds the previous round's result unless firstRound: a
// point whose previous nearest is still in another component keeps it (the
// set of other-component points only shrinks), so only the other rows are
// searched.
struct MixedPass;
MixedPass *mixedPassCreate(const float *X, int n, int d, int *status);
void mixedPassDestroy(MixedPass *mp);
int mixedPassRun(MixedPass *mp, const int *comp, int numComponents, bool sqeuclidean,
                 bool firstRound, unsigned long long *nearest);
bool mixedPassSupports(int d);

} // namespace cuhc
