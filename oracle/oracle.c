/* CPU oracle for the clustering hot path. TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; never by the cuhc_amd product path.
 *
 * A plain-C restatement of the reference's arithmetic and tie rules:
 *   - FP32 metric folds in feature order with explicit fmaf
 *     (reference src/distances.cuh:21-83); built with -ffp-contract=off;
 *   - edge weights: plain distance, mutual reachability
 *     max(core_i, core_j, d / alpha) (src/hdbscan.cuh:342-351);
 *   - the MST is the unique one under the total edge order
 *     (weight, smaller index, larger index) (src/mst.cuh:41-48); computed here
 *     with Prim's algorithm, which yields that same tree for any strict total
 *     order of the edges;
 *   - single linkage = MST edges sorted by that order, merged with union-find
 *     into scipy rows {min id, max id, height, size} (src/mst.cuh:204-235);
 *   - core distance: k-th smallest of {0 (self)} and the row's other distances
 *     clamped at 0 (src/hdbscan.cuh:15-17, 223);
 *   - dense Lance-Williams linkage in FP64 stored as FP32
 *     (src/linkage.cuh:36-56), merging the lexicographically smallest
 *     (distance, i, j) pair each step.
 * Parity notes: every function is pinned by tests/test_oracle.py against
 * scipy/sklearn fixtures, reference known answers and brute force; minkowski
 * is compared with a tolerance only (libm powf differs across libraries).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- metric folds ------------------------------------------------------ */

static float fold_pair(const float *a, const float *b, int d, int metric, float p) {
  switch (metric) {
  case 0:
  case 1: {
    float acc = 0.0f;
    for (int k = 0; k < d; ++k) {
      const float diff = a[k] - b[k];
      acc = fmaf(diff, diff, acc);
    }
    return metric == 0 ? sqrtf(acc) : acc;
  }
  case 2: {
    float acc = 0.0f;
    for (int k = 0; k < d; ++k)
      acc = acc + fabsf(a[k] - b[k]);
    return acc;
  }
  case 3: {
    float acc = 0.0f;
    for (int k = 0; k < d; ++k)
      acc = fmaxf(acc, fabsf(a[k] - b[k]));
    return acc;
  }
  case 4: {
    float acc = 0.0f;
    for (int k = 0; k < d; ++k) {
      const float x = fabsf(a[k] - b[k]);
      if (x > 0.0f)
        acc = acc + powf(x, p);
    }
    return powf(acc, 1.0f / p);
  }
  default: {
    float ab = 0.0f, aa = 0.0f, bb = 0.0f;
    for (int k = 0; k < d; ++k) {
      ab = fmaf(a[k], b[k], ab);
      aa = fmaf(a[k], a[k], aa);
      bb = fmaf(b[k], b[k], bb);
    }
    const float denom = sqrtf(aa * bb);
    return denom > 0.0f ? 1.0f - ab / denom : 1.0f;
  }
  }
}

float oracle_distance(const float *a, const float *b, int d, int metric, float p) {
  return fold_pair(a, b, d, metric, p);
}

/* Distances from point v to every point (out[u]), vectorisable over u for the
 * squared/euclidean metrics: XT is the feature-major copy (d x n). */
static void row_distances(const float *X, const float *XT, int n, int d, int metric, float p,
                          int v, float *out) {
  if (metric == 0 || metric == 1) {
    for (int u = 0; u < n; ++u)
      out[u] = 0.0f;
    for (int k = 0; k < d; ++k) {
      const float xv = X[(size_t)v * d + k];
      const float *col = XT + (size_t)k * n;
      for (int u = 0; u < n; ++u) {
        const float diff = xv - col[u];
        out[u] = fmaf(diff, diff, out[u]);
      }
    }
    if (metric == 0)
      for (int u = 0; u < n; ++u)
        out[u] = sqrtf(out[u]);
    return;
  }
  for (int u = 0; u < n; ++u)
    out[u] = fold_pair(X + (size_t)v * d, X + (size_t)u * d, d, metric, p);
}

static float *transpose(const float *X, int n, int d) {
  float *XT = (float *)malloc((size_t)n * d * sizeof(float));
  if (!XT)
    return NULL;
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < d; ++k)
      XT[(size_t)k * n + i] = X[(size_t)i * d + k];
  return XT;
}

int oracle_pairwise(const float *A, int nA, const float *B, int nB, int d, int metric, float p,
                    float *out) {
  for (int i = 0; i < nA; ++i)
    for (int j = 0; j < nB; ++j)
      out[(size_t)i * nB + j] = fold_pair(A + (size_t)i * d, B + (size_t)j * d, d, metric, p);
  return 0;
}

/* ---- edges and trees ----------------------------------------------------- */

typedef struct {
  float w;
  int u, v; /* u < v */
} edge_t;

static int edge_less(const edge_t *x, const edge_t *y) {
  if (x->w != y->w)
    return x->w < y->w;
  if (x->u != y->u)
    return x->u < y->u;
  return x->v < y->v;
}

static int edge_cmp(const void *a, const void *b) {
  const edge_t *x = (const edge_t *)a, *y = (const edge_t *)b;
  return edge_less(x, y) ? -1 : (edge_less(y, x) ? 1 : 0);
}

static int uf_find(int *parent, int x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

/* Scipy rows from merges in tree order (union by size, new ids n + step). */
static void tree_from_merges(const edge_t *m, int n, double *tree) {
  int *parent = (int *)malloc(n * sizeof(int));
  int *id = (int *)malloc(n * sizeof(int));
  int *size = (int *)malloc(n * sizeof(int));
  for (int i = 0; i < n; ++i) {
    parent[i] = i;
    id[i] = i;
    size[i] = 1;
  }
  for (int s = 0; s < n - 1; ++s) {
    int ru = uf_find(parent, m[s].u), rv = uf_find(parent, m[s].v);
    const int a = id[ru], b = id[rv], total = size[ru] + size[rv];
    double *row = tree + (size_t)s * 4;
    row[0] = a < b ? a : b;
    row[1] = a < b ? b : a;
    row[2] = m[s].w;
    row[3] = total;
    if (size[ru] < size[rv]) {
      const int t = ru;
      ru = rv;
      rv = t;
    }
    parent[rv] = ru;
    size[ru] = total;
    id[ru] = n + s;
  }
  free(parent);
  free(id);
  free(size);
}

static float mreach(float dist, float ci, float cj, float alpha) {
  const float c = fmaxf(ci, cj);
  return fmaxf(c, alpha == 1.0f ? dist : dist / alpha);
}

/* Prim's algorithm over the complete graph; best[u] is the smallest edge (by
 * the total order) from the tree to u. core == NULL: plain distances.
 * Non-finite (+inf) weights never count, as in the GPU pass. */
int oracle_mst(const float *X, int n, int d, int metric, float p, const float *core, float alpha,
               int *eu, int *ev, float *ew) {
  if (n < 2)
    return 0;
  float *XT = transpose(X, n, d);
  float *row = (float *)malloc(n * sizeof(float));
  edge_t *best = (edge_t *)malloc(n * sizeof(edge_t));
  unsigned char *in = (unsigned char *)calloc(n, 1);
  if (!XT || !row || !best || !in)
    return -1;
  for (int u = 0; u < n; ++u)
    best[u] = (edge_t){INFINITY, -1, -1};
  int v = 0, count = 0;
  in[0] = 1;
  for (int step = 0; step < n - 1; ++step) {
    row_distances(X, XT, n, d, metric, p, v, row);
    for (int u = 0; u < n; ++u) {
      if (in[u])
        continue;
      float w = row[u];
      if (core)
        w = mreach(w, core[v], core[u], alpha);
      if (!(w != INFINITY))
        continue;
      if (!(w >= 0.0f))
        w = 0.0f; /* negatives and NaN pack as 0 on the GPU */
      const edge_t e = {w, v < u ? v : u, v < u ? u : v};
      if (best[u].u < 0 || edge_less(&e, &best[u]))
        best[u] = e;
    }
    int pick = -1;
    for (int u = 0; u < n; ++u)
      if (!in[u] && best[u].u >= 0 && (pick < 0 || edge_less(&best[u], &best[pick])))
        pick = u;
    if (pick < 0)
      break; /* disconnected */
    in[pick] = 1;
    eu[count] = best[pick].u;
    ev[count] = best[pick].v;
    ew[count] = best[pick].w;
    ++count;
    v = pick;
  }
  free(XT);
  free(row);
  free(best);
  free(in);
  return count == n - 1 ? 0 : -5;
}

int oracle_tree_from_mst(const int *eu, const int *ev, const float *ew, int n, double *tree) {
  if (n < 2)
    return 0;
  edge_t *e = (edge_t *)malloc((size_t)(n - 1) * sizeof(edge_t));
  for (int i = 0; i < n - 1; ++i)
    e[i] = (edge_t){ew[i], eu[i], ev[i]};
  qsort(e, n - 1, sizeof(edge_t), edge_cmp);
  tree_from_merges(e, n, tree);
  free(e);
  return 0;
}

int oracle_single_linkage(const float *X, int n, int d, int metric, float p, double *tree) {
  if (n < 2)
    return 0;
  int *eu = (int *)malloc(n * sizeof(int)), *ev = (int *)malloc(n * sizeof(int));
  float *ew = (float *)malloc(n * sizeof(float));
  int status = oracle_mst(X, n, d, metric, p, NULL, 1.0f, eu, ev, ew);
  if (status == 0)
    oracle_tree_from_mst(eu, ev, ew, n, tree);
  free(eu);
  free(ev);
  free(ew);
  return status;
}

/* k-th smallest of v[0..n) (k is 1-based) by quickselect on a copy. */
static float kth_smallest(float *v, int n, int k) {
  int lo = 0, hi = n - 1;
  const int target = k - 1;
  while (lo < hi) {
    const float pivot = v[lo + (hi - lo) / 2];
    int i = lo, j = hi;
    while (i <= j) {
      while (v[i] < pivot)
        ++i;
      while (v[j] > pivot)
        --j;
      if (i <= j) {
        const float t = v[i];
        v[i] = v[j];
        v[j] = t;
        ++i;
        --j;
      }
    }
    if (target <= j)
      hi = j;
    else if (target >= i)
      lo = i;
    else
      return v[target];
  }
  return v[target];
}

int oracle_core_distances(const float *X, int n, int d, int metric, float p, int k, float *core) {
  float *XT = transpose(X, n, d);
  float *row = (float *)malloc(n * sizeof(float));
  if (!XT || !row)
    return -1;
  for (int i = 0; i < n; ++i) {
    row_distances(X, XT, n, d, metric, p, i, row);
    for (int j = 0; j < n; ++j)
      row[j] = j == i ? 0.0f : (row[j] > 0.0f ? row[j] : 0.0f);
    core[i] = kth_smallest(row, n, k);
  }
  free(XT);
  free(row);
  return 0;
}

int oracle_hdbscan_tree(const float *X, int n, int d, int metric, float p, int k, float alpha,
                        double *tree, float *core) {
  int status = oracle_core_distances(X, n, d, metric, p, k, core);
  if (status || n < 2)
    return status;
  int *eu = (int *)malloc(n * sizeof(int)), *ev = (int *)malloc(n * sizeof(int));
  float *ew = (float *)malloc(n * sizeof(float));
  status = oracle_mst(X, n, d, metric, p, core, alpha, eu, ev, ew);
  if (status == 0)
    oracle_tree_from_mst(eu, ev, ew, n, tree);
  free(eu);
  free(ev);
  free(ew);
  return status;
}

/* ---- dense Lance-Williams -------------------------------------------------- */

static double lance_williams(int linkage, double dak, double dbk, double dab, double na, double nb,
                             double nk) {
  switch (linkage) {
  case 0: return fmin(dak, dbk);
  case 1: return fmax(dak, dbk);
  case 2: return (na * dak + nb * dbk) / (na + nb);
  default: {
    const double total = na + nb + nk;
    const double sq = ((na + nk) * dak * dak + (nb + nk) * dbk * dbk - nk * dab * dab) / total;
    return sqrt(fmax(sq, 0.0));
  }
  }
}

/* Merges the lexicographically smallest (distance, i, j), i < j, of the active
 * slots each step; the merged cluster keeps slot i. D is the full n x n FP32
 * matrix (overwritten). Linear scans; for test sizes. */
int oracle_dense_on_matrix(float *D, int n, int linkage, double *tree) {
  if (n < 2)
    return 0;
  int *size = (int *)malloc(n * sizeof(int)), *id = (int *)malloc(n * sizeof(int));
  unsigned char *active = (unsigned char *)malloc(n);
  int *fn = (int *)malloc(n * sizeof(int)); /* forward nearest: argmin over j > i */
  for (int i = 0; i < n; ++i) {
    size[i] = 1;
    id[i] = i;
    active[i] = 1;
  }
#define DIJ(i, j) D[(size_t)(i) * n + (j)]
  for (int i = 0; i < n; ++i) {
    fn[i] = -1;
    for (int j = i + 1; j < n; ++j)
      if (fn[i] < 0 || DIJ(i, j) < DIJ(i, fn[i]))
        fn[i] = j;
  }
  for (int step = 0; step < n - 1; ++step) {
    int a = -1;
    for (int i = 0; i < n; ++i)
      if (active[i] && fn[i] >= 0 && (a < 0 || DIJ(i, fn[i]) < DIJ(a, fn[a])))
        a = i;
    const int b = fn[a];
    const float dab = DIJ(a, b);
    const int na = size[a], nb = size[b];
    double *row = tree + (size_t)step * 4;
    row[0] = id[a] < id[b] ? id[a] : id[b];
    row[1] = id[a] < id[b] ? id[b] : id[a];
    row[2] = dab;
    row[3] = na + nb;
    active[b] = 0;
    for (int k = 0; k < n; ++k) {
      if (!active[k] || k == a)
        continue;
      const float v = (float)lance_williams(linkage, DIJ(a, k), DIJ(b, k), dab, na, nb, size[k]);
      DIJ(a, k) = v;
      DIJ(k, a) = v;
    }
    size[a] = na + nb;
    id[a] = n + step;
    /* Only rows i <= a (entry (i, a) changed) and a < i < b (entry (i, b)
       retired) can have a new forward nearest. Row a and rows whose nearest
       was a or b are rescanned; a row i < a otherwise only compares its
       current nearest with the new (i, a). */
    for (int i = 0; i < b; ++i) {
      if (!active[i])
        continue;
      if (i == a || fn[i] == a || fn[i] == b) {
        fn[i] = -1;
        for (int j = i + 1; j < n; ++j)
          if (active[j] && (fn[i] < 0 || DIJ(i, j) < DIJ(i, fn[i])))
            fn[i] = j;
      } else if (i < a && (DIJ(i, a) < DIJ(i, fn[i]) ||
                           (DIJ(i, a) == DIJ(i, fn[i]) && a < fn[i]))) {
        fn[i] = a;
      }
    }
  }
#undef DIJ
  free(size);
  free(id);
  free(active);
  free(fn);
  return 0;
}

/* ---- radius-proposal network ------------------------------------------------
 * 4 inputs -> 32 tanh -> 32 tanh -> 1; parameters W1[32][4], b1, W2[32][32],
 * b2, W3[32], b3 (the layout of cuhc_mlp_*). Loss: mean of 0.5 (y - t)^2.
 * Forward and backward in FP64 by the chain rule, written out directly. */
#define MLP_F 4
#define MLP_H 32

static void mlp_forward1(const float *p, const float *x, double *h1, double *h2, double *y) {
  const float *W1 = p, *b1 = W1 + MLP_H * MLP_F, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
  const float *W3 = b2 + MLP_H, *b3 = W3 + MLP_H;
  for (int h = 0; h < MLP_H; ++h) {
    double z = b1[h];
    for (int f = 0; f < MLP_F; ++f)
      z += (double)W1[h * MLP_F + f] * x[f];
    h1[h] = tanh(z);
  }
  for (int h = 0; h < MLP_H; ++h) {
    double z = b2[h];
    for (int g = 0; g < MLP_H; ++g)
      z += (double)W2[h * MLP_H + g] * h1[g];
    h2[h] = tanh(z);
  }
  double out = b3[0];
  for (int h = 0; h < MLP_H; ++h)
    out += (double)W3[h] * h2[h];
  *y = out;
}

int oracle_mlp_forward(const float *p, const float *X, int m, double *y) {
  double h1[MLP_H], h2[MLP_H];
  for (int s = 0; s < m; ++s)
    mlp_forward1(p, X + (size_t)s * MLP_F, h1, h2, &y[s]);
  return 0;
}

int oracle_mlp_gradient(const float *p, const float *X, const float *t, int m, double *grad,
                        double *loss) {
  const int P = MLP_H * MLP_F + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;
  const float *W2 = p + MLP_H * MLP_F + MLP_H, *W3 = W2 + MLP_H * MLP_H + MLP_H;
  double *gW1 = grad, *gb1 = gW1 + MLP_H * MLP_F, *gW2 = gb1 + MLP_H, *gb2 = gW2 + MLP_H * MLP_H;
  double *gW3 = gb2 + MLP_H, *gb3 = gW3 + MLP_H;
  for (int q = 0; q < P; ++q)
    grad[q] = 0.0;
  double total = 0.0;
  for (int s = 0; s < m; ++s) {
    const float *x = X + (size_t)s * MLP_F;
    double h1[MLP_H], h2[MLP_H], y, dz2[MLP_H], dz1[MLP_H];
    mlp_forward1(p, x, h1, h2, &y);
    const double dy = y - t[s];
    total += 0.5 * dy * dy;
    for (int h = 0; h < MLP_H; ++h)
      dz2[h] = dy * W3[h] * (1.0 - h2[h] * h2[h]);
    for (int g = 0; g < MLP_H; ++g) {
      double back = 0.0;
      for (int h = 0; h < MLP_H; ++h)
        back += (double)W2[h * MLP_H + g] * dz2[h];
      dz1[g] = back * (1.0 - h1[g] * h1[g]);
    }
    for (int h = 0; h < MLP_H; ++h) {
      for (int f = 0; f < MLP_F; ++f)
        gW1[h * MLP_F + f] += dz1[h] * x[f];
      gb1[h] += dz1[h];
      for (int g = 0; g < MLP_H; ++g)
        gW2[h * MLP_H + g] += dz2[h] * h1[g];
      gb2[h] += dz2[h];
      gW3[h] += dy * h2[h];
    }
    gb3[0] += dy;
  }
  for (int q = 0; q < P; ++q)
    grad[q] /= m;
  if (loss)
    *loss = total / m;
  return 0;
}

int oracle_dense_linkage(const float *X, int n, int d, int metric, float p, int linkage,
                         double *tree) {
  if (n < 2)
    return 0;
  float *D = (float *)malloc((size_t)n * n * sizeof(float));
  if (!D)
    return -1;
  oracle_pairwise(X, n, X, n, d, metric, p, D);
  const int status = oracle_dense_on_matrix(D, n, linkage, tree);
  free(D);
  return status;
}
