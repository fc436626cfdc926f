// HDBSCAN cluster extraction on the host (extract.cpp): condensed tree,
// stabilities, cluster selection, labels and membership probabilities.
#pragma once

#include <vector>

namespace cuhc {

struct CondensedRow {
  long long parent, child;
  double lambda;
  long long size;
};

// Condensed tree of an (n - 1) x 4 scipy-format single-linkage tree: rows
// (parent cluster, child point or cluster, lambda = 1 / height, size), the
// root cluster numbered n, new clusters numbered from n + 1.
std::vector<CondensedRow> condenseTree(const double *tree, long long n, long long minClusterSize);

// Columns of a condensed tree (rows in condenseTree's order).
struct CondensedView {
  const long long *parent, *child;
  const double *lambda;
  const long long *size;
  long long rows;
  long long root = 0, maxParent = -1, numIds = 0;
  CondensedView(const long long *parent, const long long *child, const double *lambda,
                const long long *size, long long rows);
};

// Stability of clusters root .. maxParent (index: id - root).
std::vector<double> clusterStability(const CondensedView &t);

// Excess of mass (method 0) or leaf (method 1) selection without the epsilon
// rule: selected[id - root] = 1 for the chosen clusters (maxParent - root + 1
// entries). stability (same indexing) is updated as excess of mass does.
// maxClusterSize <= 0: no limit.
int selectClusters(const CondensedView &t, double *stability, int method, bool allowSingle,
                   long long maxClusterSize, unsigned char *selected);

// Labels of the points 0 .. root - 1: -1 for noise, else the rank of the
// point's selected cluster among the selected ids.
void labelPoints(const CondensedView &t, const unsigned char *selected, bool allowSingle,
                 double epsilon, long long *labels);

// Membership probability of each labelled point: min(lambda, death) / death of
// its cluster (1 when the death lambda is 0 or the point's lambda infinite),
// 0 for noise.
void membershipProbabilities(const CondensedView &t, const unsigned char *selected,
                             const long long *labels, double *probabilities);

} // namespace cuhc
