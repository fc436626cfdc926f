// All-pairs FP32 kernels: distance matrices (full and condensed), the
// symmetric nearest-other-component pass, and the spatially pruned walk pass.
//
// Reference counterparts: pairwiseDistanceKernel and
// nearestOtherComponentSymmetricKernel (src/distance_matrix.cuh:179-200,
// 287-390), buildTileBoxesKernel / nearestOtherComponentWalkKernel /
// estimatePrunedShare (src/spatial.cuh:136-343, 424-469). The tile fold is
// tile.hpp; argmins carry (weight, index) as one u64 key.
#include "passes.hpp"
#include "primitives.hpp"
#include "tile.hpp"

#include <algorithm>
#include <atomic>
#include <thread>
#include <numeric>
#include <type_traits>

namespace cuhc {

bool validMetric(int id) { return id >= kEuclidean && id <= kCosine; }
bool metricHasBoxBound(int id) { return validMetric(id) && id != kCosine; }
int tilePoints(int metricId) { return metricId == kCosine ? 64 : 128; }

namespace {

template <class F> int withMetric(MetricSpec m, F &&f) {
  switch (m.id) {
  case kEuclidean: return f(Euclidean{});
  case kSqEuclidean: return f(SqEuclidean{});
  case kManhattan: return f(Manhattan{});
  case kChebyshev: return f(Chebyshev{});
  case kMinkowski: return f(Minkowski{m.p});
  case kCosine: return f(Cosine{});
  default: return kErrInvalidMetric;
  }
}

// LDS staging width: 8 features for d <= 8, else 32 (a fill stages only the
// features present, rounded up to 8).
template <class F> int withFill(int d, F &&f) {
  if (d <= 8)
    return f(std::integral_constant<int, 8>{});
  return f(std::integral_constant<int, 32>{});
}

int persistentBlocks(long long work, int perCU) {
  const long long cap = (long long)numCUs() * perCU;
  return (int)std::max<long long>(1, std::min<long long>(work, cap));
}

// ---------------------------------------------------------------------------
// Distance matrices

template <class T, class Metric>
__global__ void __launch_bounds__(256)
    pairwiseMatrixKernel(const float *__restrict__ A, int nA, const float *__restrict__ B, int nB,
                         int d, Metric metric, float *__restrict__ out, int tilesB) {
  __shared__ __align__(16) TileSmem<T> sm;
  const int tileI = blockIdx.x / tilesB, tileJ = blockIdx.x - tileI * tilesB;
  const int row0 = tileI * T::DIM, col0 = tileJ * T::DIM;
  float v[T::TR][T::TR];
  computeTile<T>(sm, A, nA, B, nB, d, metric, row0, col0, v);
#pragma unroll
  for (int r = 0; r < T::TR; ++r) {
    const int row = row0 + T::rowInTile(r);
    if (row >= nA)
      continue;
#pragma unroll
    for (int c = 0; c < T::TR; ++c) {
      const int col = col0 + T::colInTile(c);
      if (col < nB)
        out[(size_t)row * nB + col] = v[r][c];
    }
  }
}

template <class T, class Metric>
__global__ void __launch_bounds__(256)
    condensedMatrixKernel(const float *__restrict__ X, int n, int d, Metric metric,
                          float *__restrict__ out, long long numPairs) {
  __shared__ __align__(16) TileSmem<T> sm;
  for (long long t = blockIdx.x; t < numPairs; t += gridDim.x) {
    int I, J;
    decodeTrianglePair(t, I, J);
    const int row0 = I * T::DIM, col0 = J * T::DIM;
    float v[T::TR][T::TR];
    computeTile<T>(sm, X, n, X, n, d, metric, row0, col0, v);
#pragma unroll
    for (int r = 0; r < T::TR; ++r) {
      const long long i = row0 + T::rowInTile(r);
      if (i >= n)
        continue;
      const long long base = i * n - i * (i + 1) / 2 - i - 1;
#pragma unroll
      for (int c = 0; c < T::TR; ++c) {
        const long long j = col0 + T::colInTile(c);
        if (j > i && j < n)
          out[base + j] = v[r][c];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Symmetric nearest-other-component pass

// Reduces a key over the 16 lanes of a tile row (xor 8, 4, 2, 1).
DEV unsigned long long rowGroupMin(unsigned long long k) {
#pragma unroll
  for (int m = 8; m > 0; m >>= 1)
    k = minKey(k, shflXorKey(k, m));
  return k;
}

// Pair filter of the symmetric pass. A point's limit comes from the weight of
// its current nearest (+inf while it has none; the key's high word, read
// alone: nearest only decreases, so a stale read is a looser limit). A pair
// passes when its filter value is within the larger of its two points'
// limits; every pair whose weight is within a point's nearest weight passes.
//   metrics with an accumulator bound (plain and mutual-reachability
//   weights): value = the raw accumulator, limit = accBound(distBound(w));
//   Ward: 2 si sj / (si + sj) >= min(si, sj), so the weight is at least
//   sq min(si, sj) up to the rounding of its four operations (2^-22):
//   value = sq min(si, sj), limit = w (1 + 2^-20) + FLT_MIN.
template <class Metric, class Weight> struct PairFilter {
  static constexpr bool kWard = std::is_same<Weight, WardWeight>::value;
  static constexpr bool kOn = (Metric::kAccFilter && Weight::kDistBound) || kWard;
  DEV static float limit(const Metric &metric, const Weight &weight,
                         const unsigned long long *nearest, int p) {
    return limitOf(metric, weight,
                   reinterpret_cast<const unsigned int *>(nearest)[2 * (size_t)p + 1]);
  }
  DEV static float limitOf(const Metric &metric, const Weight &weight, unsigned int bits) {
    if (bits == 0xffffffffu)
      return INFINITY;
    const float w = floatOf(bits);
    if constexpr (kWard)
      return w * (1.0f + 0x1p-20f) + FLT_MIN;
    else
      return metric.accBound(weight.distBound(w));
  }
  DEV static float value(float acc, typename Weight::Value vi, typename Weight::Value vj) {
    if constexpr (kWard)
      return acc * fminf(vi, vj);
    else
      return acc;
  }
};

// Row and column data of a symmetric tile staged in LDS with the tile's first
// fill: threads [0, DIM) load column col0 + t, threads [DIM, 2 DIM) row
// row0 + t - DIM (component, weight value, original index and, when
// filtering, the weight bits of the current nearest, turned into a limit).
// Points past n get component -1 (columns) or -2 (rows) and limit -inf.
template <class Metric, class Weight, bool kFilter> struct PairStage {
  using WV = typename Weight::Value;
  struct Value {
    int comp;
    WV val;
    unsigned int id, bits;
  };
  const int *comp;
  const int *orig;
  const unsigned long long *nearest;
  Metric metric;
  Weight weight;
  int row0, col0, n, dim;
  int *sComp;      // [2 dim]: columns, then rows
  WV *sVal;        // [2 dim]
  unsigned *sId;   // [2 dim]
  float *sLimit;   // [2 dim]
  DEV int point() const {
    const int t = (int)threadIdx.x;
    return t < dim ? col0 + t : row0 + (t - dim);
  }
  DEV Value load() const {
    Value v{};
    if ((int)threadIdx.x < 2 * dim) {
      const int pl = min(point(), n - 1);
      v.comp = comp[pl];
      v.val = weight.value(pl);
      v.id = orig ? (unsigned)orig[pl] : (unsigned)pl;
      if (kFilter)
        v.bits = reinterpret_cast<const unsigned int *>(nearest)[2 * (size_t)pl + 1];
    }
    return v;
  }
  DEV void store(const Value &v) const {
    const int t = (int)threadIdx.x;
    if (t < 2 * dim) {
      const int p = point();
      const bool ok = p < n;
      sComp[t] = ok ? v.comp : (t < dim ? -1 : -2);
      sVal[t] = v.val;
      sId[t] = ok ? v.id : (unsigned)p;
      if (kFilter)
        sLimit[t] = ok ? PairFilter<Metric, Weight>::limitOf(metric, weight, v.bits) : -INFINITY;
    }
  }
};

template <class T, class Metric, class Weight>
__global__ void __launch_bounds__(256, 3) // three waves per SIMD (at most 168 VGPRs)
    nearestOtherSymmetricKernel(const float *__restrict__ X, int n, int d, Metric metric,
                                Weight weight, const int *__restrict__ comp,
                                const int *__restrict__ orig,
                                unsigned long long *__restrict__ nearest, long long numPairs) {
  constexpr int TR = T::TR;
  // With an accumulator bound, a pair is finished (sqrt, weight, keys) only
  // when its raw accumulator is within the limit of its row point or of its
  // column point (from their current nearest) and it crosses components.
  // Every point's nearest pair is within its own limit, so every point's
  // result stays exact.
  using Filter = PairFilter<Metric, Weight>;
  constexpr bool kFilter = Filter::kOn;
  __shared__ __align__(16) TileSmem<T> sm;
  __shared__ unsigned long long colBest[T::DIM];
  // Point data of the tile: columns [0, DIM), rows [DIM, 2 DIM).
  __shared__ int sComp[2 * T::DIM];
  __shared__ typename Weight::Value sVal[2 * T::DIM];
  __shared__ unsigned int sId[2 * T::DIM];
  __shared__ float sLimit[2 * T::DIM];
  static_assert(2 * T::DIM <= T::THREADS, "one point per thread");

  for (long long t = blockIdx.x; t < numPairs; t += gridDim.x) {
    int I, J;
    decodeTrianglePair(t, I, J);
    const int row0 = I * T::DIM, col0 = J * T::DIM;
    for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
      colBest[i] = kEmptyKey; // visible after computeTile's barriers

    float w[TR][TR]; // raw accumulators when filtering
    // The point data arrives with the tile's first fill (read after
    // computeTile's barriers; rewritten only after this tile's last barrier).
    const PairStage<Metric, Weight, kFilter> stage{comp, orig, nearest, metric, weight, row0,
                                                   col0, n, T::DIM, sComp, sVal, sId, sLimit};
    computeTile<T, Metric, !kFilter>(sm, X, n, X, n, d, metric, row0, col0, w, stage);

    int colComp[TR];
    unsigned int colId[TR];
    typename Weight::Value colVal[TR];
#pragma unroll
    for (int c = 0; c < TR; ++c) {
      const int i = T::colInTile(c);
      colComp[c] = sComp[i];
      colVal[c] = sVal[i];
      colId[c] = sId[i];
    }
    // A pair counts if both points exist, lie in different components and
    // have a finite weight (+inf never counts, as in the reference's
    // symmetric pass; NaN counts and packs as 0).
    unsigned long long rowKey[TR];
    unsigned long long countMask = 0;
    float colLimit[TR];
    if constexpr (kFilter) {
#pragma unroll
      for (int c = 0; c < TR; ++c)
        colLimit[c] = sLimit[T::colInTile(c)];
    }
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      const int ri = T::DIM + T::rowInTile(r);
      const int rc = sComp[ri];
      const typename Weight::Value rv = sVal[ri];
      unsigned long long best = kEmptyKey;
      if constexpr (kFilter) {
        const float rowLimit = sLimit[ri];
#pragma unroll
        for (int c = 0; c < TR; ++c) {
          const float acc = w[r][c];
          if (!(Filter::value(acc, rv, colVal[c]) > fmaxf(rowLimit, colLimit[c])) &&
              colComp[c] != rc) {
            const float ww = weight(metric.finish(acc), rv, colVal[c]);
            w[r][c] = ww;
            if (colComp[c] >= 0 && rc >= 0 && ww != INFINITY) {
              countMask |= 1ull << (r * TR + c);
              best = minKey(best, packKey(ww, colId[c]));
            }
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < TR; ++c) {
          const float ww = weight(w[r][c], rv, colVal[c]);
          const bool counts = colComp[c] >= 0 && rc >= 0 && colComp[c] != rc && ww != INFINITY;
          w[r][c] = ww;
          if (counts) {
            countMask |= 1ull << (r * TR + c);
            best = minKey(best, packKey(ww, colId[c]));
          }
        }
      }
      // Wave-uniform: skip the reduction when no lane has a candidate.
      rowKey[r] = __ballot(best != kEmptyKey) != 0 ? rowGroupMin(best) : kEmptyKey;
    }
    if (T::tc() == 0) {
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int row = row0 + T::rowInTile(r);
        if (row < n && rowKey[r] != kEmptyKey)
          atomicMin(&nearest[row], rowKey[r]);
      }
    }
    if (I != J) {
      // Columns: this thread's rows, then lanes tc+16, +32, +48 of the wave,
      // then the 4 waves through LDS (skipped by a wave with no counted pair).
      if (__ballot(countMask != 0) != 0) {
        unsigned int rowId[TR];
#pragma unroll
        for (int r = 0; r < TR; ++r)
          rowId[r] = sId[T::DIM + T::rowInTile(r)];
#pragma unroll
        for (int c = 0; c < TR; ++c) {
          unsigned long long best = kEmptyKey;
#pragma unroll
          for (int r = 0; r < TR; ++r)
            if (countMask & (1ull << (r * TR + c)))
              best = minKey(best, packKey(w[r][c], rowId[r]));
          best = minKey(best, shflXorKey(best, 16));
          best = minKey(best, shflXorKey(best, 32));
          if ((threadIdx.x & 63) < 16 && best != kEmptyKey)
            atomicMin(&colBest[T::colInTile(c)], best);
        }
      }
      __syncthreads();
      for (int i = threadIdx.x; i < T::DIM; i += T::THREADS) {
        const int col = col0 + i;
        if (col < n && colBest[i] != kEmptyKey)
          atomicMin(&nearest[col], colBest[i]);
      }
    }
    __syncthreads(); // colBest reused by the next tile
  }
}

// ---------------------------------------------------------------------------
// Spatial pruning: boxes and the walk pass

struct Boxes {
  float *low = nullptr, *high = nullptr; // tiles x d
  float *minValue = nullptr;             // smallest per-point weight value in the tile
  int *minComp = nullptr, *maxComp = nullptr;
  int numTiles = 0;
};

// Boxes of the tiles of X (n points of d features, tileDim per tile). The
// weight values and components of point p are those of position pos[p]
// (pos null: p), for rows gathered out of the full order.
template <class Weight>
__global__ void buildBoxesKernel(const float *__restrict__ X, int n, int d, int tileDim,
                                 Weight weight, const int *__restrict__ comp,
                                 const int *__restrict__ pos, Boxes boxes) {
  const int tile = blockIdx.x;
  const int first = tile * tileDim, last = min(n, first + tileDim);
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    float lo = INFINITY, hi = -INFINITY;
    for (int p = first; p < last; ++p) {
      const float v = X[(size_t)p * d + f];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    boxes.low[(size_t)tile * d + f] = lo;
    boxes.high[(size_t)tile * d + f] = hi;
  }
  // Per-point values and component range: a wave-wide min/max.
  if (threadIdx.x < 64) {
    float mv = INFINITY;
    int cmin = 0x7fffffff, cmax = -1;
    for (int p = first + (int)threadIdx.x; p < last; p += 64) {
      const int q = pos ? pos[p] : p;
      mv = fminf(mv, (float)weight.value(q));
      const int c = comp ? comp[q] : 0;
      cmin = min(cmin, c);
      cmax = max(cmax, c);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      mv = fminf(mv, __shfl_xor(mv, m, 64));
      cmin = min(cmin, __shfl_xor(cmin, m, 64));
      cmax = max(cmax, __shfl_xor(cmax, m, 64));
    }
    if (threadIdx.x == 0) {
      boxes.minValue[tile] = Weight::kHasValue ? mv : 0.0f;
      boxes.minComp[tile] = cmin;
      boxes.maxComp[tile] = cmax;
    }
  }
}

// Relative slack of the box bound (reference src/spatial.cuh:183-185): far
// larger than the rounding of a pair's distance or of the bound, far too
// small to cost pruning.
HD float boxSlack(int d) { return 1e-5f + (d + 8) * 5e-7f; }

template <class Metric>
DEV float boxLowerBound(const Metric &metric, const float *lowA, const float *highA,
                        const float *lowB, const float *highB, int d) {
  typename Metric::Acc acc = metric.init();
  for (int k = 0; k < d; ++k) {
    const float gap = fmaxf(0.0f, fmaxf(lowB[k] - highA[k], lowA[k] - highB[k]));
    acc = metric.fold(acc, gap, 0.0f);
  }
  return metric.finish(acc);
}

// Column tile visited at `step` from row tile I: I, I+1, I-1, I+2, ... (-1 if outside).
DEV int walkTile(int rowTile, int step, int numTiles) {
  const int off = (step + 1) >> 1;
  const int tile = (step & 1) ? rowTile + off : rowTile - off;
  return (tile >= 0 && tile < numTiles) ? tile : -1;
}

// Column data of a walk tile staged in LDS with the tile's first fill:
// thread t < DIM loads column col0 + t's component, weight value and
// original index (from a clamped index; -1 marks columns past n).
template <class Weight> struct ColumnStage {
  using WV = typename Weight::Value;
  struct Value {
    int comp;
    WV val;
    unsigned int idx;
  };
  const int *comp;
  const int *orig;
  Weight weight;
  int col0, n, dim;
  int *sComp;
  WV *sVal;
  unsigned int *sIdx;
  DEV Value load() const {
    Value v{};
    if ((int)threadIdx.x < dim) {
      const int col = col0 + (int)threadIdx.x;
      const int cl = min(col, n - 1);
      v.comp = comp[cl];
      v.val = weight.value(cl);
      v.idx = orig ? (unsigned)orig[cl] : (unsigned)cl; // null: keys carry positions
    }
    return v;
  }
  DEV void store(const Value &v) const {
    if ((int)threadIdx.x < dim) {
      const bool ok = col0 + (int)threadIdx.x < n;
      sComp[threadIdx.x] = ok ? v.comp : -1;
      sVal[threadIdx.x] = v.val;
      sIdx[threadIdx.x] = v.idx;
    }
  }
};

// One workgroup per row tile walks the column tiles outward from its own
// (walkTile order). Lower bounds of the weight between the two boxes are
// computed 64 steps at a time by wave 0, one step per lane, into LDS; a step
// is skipped when its bound exceeds every row's need. A row needs candidates
// below min(its current best, its component's best known edge): Boruvka only
// uses each component's smallest outgoing edge, so once any point of the
// component has found an edge of weight b, no row of that component needs a
// candidate above b (compBest, shared by all workgroups through atomicMin on
// the weight bits; null disables it). The row that holds the component's
// smallest edge always finds it exactly; other rows may stop early, and their
// results are then only upper bounds, which the contraction never prefers.
// The same need filters single pairs: wave 0 publishes each row's need in
// LDS, and for metrics with an accumulator bound a pair is finished (sqrt,
// weight, key) only when its raw accumulator is within the row's need and
// it crosses components; the pairs filtered out weigh more than the need.
// The rows of a compact walk: n rows gathered into X (d features each) from
// positions pos[0 .. n) of the full order (ascending), in tiles with their
// own boxes; center[t] is the full-order tile around which row tile t walks.
struct RowSet {
  const float *X = nullptr;
  int n = 0;
  const int *pos = nullptr;
  Boxes boxes;
  const int *center = nullptr;
};

// KEEP: every row keeps its kKeptKeys smallest keys over all other
// components, written to kept[position][0 ..] (ascending, kEmptyKey padded),
// and its need is the last of them, so that later rounds can settle the row
// from the list (compBest must be null: a component's best would cut the
// search at the row's first key).
constexpr int kKeptKeys = 4;

// The keeping walk runs the first round only; three waves per SIMD hide its
// waits better than two despite 5 spilled registers (C3's four Borůvka walks:
// 298 ms with two waves, 265 ms with three).
template <class T, class Metric, class Weight, bool COMPACT = false, bool KEEP = false>
__global__ void __launch_bounds__(256, (KEEP ? 3 : 1))
    nearestOtherWalkKernel(const float *__restrict__ X, int n, int d, Metric metric,
                           Weight weight, const int *__restrict__ comp,
                           const int *__restrict__ orig, Boxes boxes,
                           unsigned int *__restrict__ compBest,
                           unsigned long long *__restrict__ nearest,
                           unsigned long long *tilesComputed, int rowTileStride,
                           int computeSteps, const int *__restrict__ skipComp,
                           const unsigned char *__restrict__ closedRows, RowSet rows,
                           unsigned long long *__restrict__ kept) {
  constexpr int TR = T::TR;
  __shared__ unsigned long long rowTop[KEEP ? T::DIM * kKeptKeys : 1];
  // Rows: the full order's tiles, or (COMPACT) the gathered rows, whose
  // position in the full order indexes comp, the weight values, closedRows
  // and nearest.
  const float *__restrict__ rowX = COMPACT ? rows.X : X;
  const int nRows = COMPACT ? rows.n : n;
  const Boxes &rowBoxes = COMPACT ? rows.boxes : boxes;
  auto posOf = [&](int row) { return COMPACT ? rows.pos[row] : row; };
  constexpr int kBatch = 64;
  __shared__ __align__(16) TileSmem<T> sm;
  __shared__ unsigned long long rowBest[T::DIM];
  __shared__ float rowNeed[T::DIM]; // -1: no row
  __shared__ float lowerBound[kBatch];
  __shared__ float bound;
  __shared__ int tileImproved; // a row's best improved in the current tile
  __shared__ int sColComp[T::DIM];
  __shared__ typename Weight::Value sColVal[T::DIM];
  __shared__ unsigned int sColIdx[T::DIM];
  extern __shared__ float rowBox[]; // low[d], high[d]
  constexpr bool kFilter = Metric::kAccFilter && Weight::kDistBound;

  const int rowTile = blockIdx.x * rowTileStride, numTiles = boxes.numTiles;
  const int row0 = rowTile * T::DIM;
  const int center = COMPACT ? rows.center[rowTile] : rowTile;
  // Rows of the skipped component are closed; a tile of only such rows has
  // nothing to do (uniform: every thread reads the same values).
  const int skip = skipComp ? *skipComp : -1;
  if (skip >= 0 && rowBoxes.minComp[rowTile] == skip && rowBoxes.maxComp[rowTile] == skip)
    return;
  // A row is open unless it lies past the end, in the skipped component, or
  // is closed (its exact answer is already in nearest[]).
  auto isOpen = [&](int row) {
    if (row >= nRows)
      return false;
    const int p = posOf(row);
    return comp[p] != skip && !(closedRows && closedRows[p]);
  };
  int anyOpen = 0;
  if (KEEP)
    for (int i = threadIdx.x; i < T::DIM * kKeptKeys; i += T::THREADS)
      rowTop[i] = kEmptyKey;
  for (int i = threadIdx.x; i < T::DIM; i += T::THREADS) {
    rowBest[i] = kEmptyKey;
    const bool open = isOpen(row0 + i);
    rowNeed[i] = open ? INFINITY : -1.0f;
    anyOpen |= open;
  }
  if (!__syncthreads_or(anyOpen))
    return; // uniform: every row of the tile is settled
  for (int k = threadIdx.x; k < d; k += T::THREADS) {
    rowBox[k] = rowBoxes.low[(size_t)rowTile * d + k];
    rowBox[d + k] = rowBoxes.high[(size_t)rowTile * d + k];
  }
  // Wave 0 owns rows lane and lane + 64 for the bound: their components and
  // the components' best known edges (refreshed once per batch).
  const int ra = row0 + (int)(threadIdx.x & 63), rb = ra + 64;
  const bool open0 = threadIdx.x < 64 && isOpen(ra);
  const bool open1 = threadIdx.x < 64 && isOpen(rb);
  const int comp0 = open0 ? comp[posOf(ra)] : 0, comp1 = open1 ? comp[posOf(rb)] : 0;
  float cap0 = INFINITY, cap1 = INFINITY;
  auto capOf = [&](unsigned int bits) { return bits == 0xffffffffu ? INFINITY : floatOf(bits); };
  if (threadIdx.x == 0) {
    bound = INFINITY;
    tileImproved = 0;
  }
  const bool rowTilePure = rowBoxes.minComp[rowTile] == rowBoxes.maxComp[rowTile];
  const int rowCompTile = rowBoxes.minComp[rowTile];
  const float rowMinValue = rowBoxes.minValue[rowTile];
  const float slack = 1.0f - boxSlack(d);

  int rowComp[TR];
  typename Weight::Value rowVal[TR];
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int row = row0 + T::rowInTile(r);
    rowComp[r] = isOpen(row) ? comp[posOf(row)] : -2;
    rowVal[r] = weight.value(row < nRows ? posOf(row) : 0);
  }
  unsigned long long computed = 0;
  // The 8-feature tiles serve d <= 8 only (one fill): the row slice is
  // staged once for every tile.
  constexpr bool kRowsResident = T::FILL == 8;
  if (kRowsResident)
    stageRows<T>(sm, rowX, nRows, d, row0);
  __syncthreads();

  // The bound: the largest need over the open rows (inf while a row of a
  // component without a known edge has no candidate yet). Wave 0 only; it
  // also publishes each row's need.
  auto newBound = [&]() {
    float need = 0.0f;
    if (open0) {
      const unsigned long long k =
          KEEP ? rowTop[threadIdx.x * kKeptKeys + kKeptKeys - 1] : rowBest[threadIdx.x];
      const float w = k == kEmptyKey ? INFINITY : keyWeight(k);
      if (compBest && w < cap0) {
        atomicMin(&compBest[comp0], bitsOf(w));
        cap0 = w;
      }
      need = fminf(w, cap0);
      rowNeed[threadIdx.x] = need;
    }
    if (open1) {
      const unsigned long long k = KEEP ? rowTop[(threadIdx.x + 64) * kKeptKeys + kKeptKeys - 1]
                                        : rowBest[threadIdx.x + 64];
      const float w = k == kEmptyKey ? INFINITY : keyWeight(k);
      if (compBest && w < cap1) {
        atomicMin(&compBest[comp1], bitsOf(w));
        cap1 = w;
      }
      const float need1 = fminf(w, cap1);
      rowNeed[threadIdx.x + 64] = need1;
      need = fmaxf(need, need1);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
      need = fmaxf(need, __shfl_xor(need, m, 64));
    return need;
  };

  const int steps = 2 * numTiles;
  for (int base = 0; base < steps; base += kBatch) {
    if (threadIdx.x < kBatch) {
      if (compBest) {
        // Other workgroups may have found better edges for these components.
        if (open0)
          cap0 = fminf(cap0, capOf(compBest[comp0]));
        if (open1)
          cap1 = fminf(cap1, capOf(compBest[comp1]));
        const float b = newBound();
        if (threadIdx.x == 0)
          bound = b;
      }
      // -1: never computed (outside the grid, or both tiles wholly one component).
      const int step = base + (int)threadIdx.x;
      const int colTile = step < steps ? walkTile(center, step, numTiles) : -1;
      float lb = -1.0f;
      if (colTile >= 0 &&
          !(rowTilePure && boxes.minComp[colTile] == boxes.maxComp[colTile] &&
            boxes.minComp[colTile] == rowCompTile)) {
        const float db = boxLowerBound(metric, rowBox, rowBox + d, boxes.low + (size_t)colTile * d,
                                       boxes.high + (size_t)colTile * d, d);
        lb = weight(db, rowMinValue, boxes.minValue[colTile]) * slack;
      }
      lowerBound[threadIdx.x] = lb;
    }
    __syncthreads();
    for (int j = 0; j < kBatch; ++j) {
      const float lb = lowerBound[j];
      if (lb < 0.0f || lb > bound)
        continue; // uniform: lowerBound and bound are shared
      const int colTile = walkTile(center, base + j, numTiles);
      ++computed;
      if (base + j >= computeSteps)
        continue; // estimate: counted against the bound reached so far, not computed
      const int col0 = colTile * T::DIM;
      float w[TR][TR]; // raw accumulators when filtering
      // The column data arrives with the tile's first fill (LDS, read after
      // computeTile's barriers; rewritten only after this tile's last barrier).
      const ColumnStage<Weight> stage{comp, orig, weight, col0, n, T::DIM,
                                      sColComp, sColVal, sColIdx};
      computeTile<T, Metric, !kFilter, kRowsResident>(sm, rowX, nRows, X, n, d, metric, row0,
                                                      col0, w, stage);
      int colComp[TR];
      typename Weight::Value colVal[TR];
      unsigned int colIdx[TR];
#pragma unroll
      for (int c = 0; c < TR; ++c) {
        const int i = T::colInTile(c);
        colComp[c] = sColComp[i];
        colVal[c] = sColVal[i];
        colIdx[c] = sColIdx[i];
      }
      if constexpr (KEEP) {
        const int lane = threadIdx.x & 63;
        const unsigned long long groupMask = 0xffffull << (lane & 48);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          const int i = T::rowInTile(r);
          unsigned long long *top = rowTop + i * kKeptKeys;
          // Pairs that can enter the row's list (a stale last key only adds
          // pairs that the insertion then rejects).
          const float need = rowNeed[i];
          const float limit = need < 0.0f ? -INFINITY
                              : kFilter   ? metric.accBound(weight.distBound(need))
                                          : INFINITY;
          const unsigned long long last = top[kKeptKeys - 1];
          unsigned int pending = 0;
#pragma unroll
          for (int c = 0; c < TR; ++c) {
            const float acc = w[r][c];
            if (!(acc > limit) && colComp[c] != rowComp[r] && rowComp[r] >= 0 && colComp[c] >= 0) {
              const float ww = weight(kFilter ? metric.finish(acc) : acc, rowVal[r], colVal[c]);
              w[r][c] = ww;
              if (ww != INFINITY && packKey(ww, colIdx[c]) < last)
                pending |= 1u << c;
            }
          }
          // One insertion per row group at a time; the lanes stay converged.
          while (true) {
            const unsigned long long work = __ballot(pending != 0);
            if (work == 0)
              break;
            const unsigned long long mine = work & groupMask;
            if (mine != 0 && lane == __ffsll((long long)mine) - 1) {
              const int c = __ffs(pending) - 1;
              pending &= pending - 1;
              float wv = 0.0f;
              unsigned int idx = 0;
#pragma unroll
              for (int cc = 0; cc < TR; ++cc)
                if (cc == c) {
                  wv = w[r][cc];
                  idx = colIdx[cc];
                }
              unsigned long long key = packKey(wv, idx);
              if (key < top[kKeptKeys - 1]) {
#pragma unroll
                for (int q = 0; q < kKeptKeys; ++q) {
                  const unsigned long long cur = top[q];
                  top[q] = key < cur ? key : cur;
                  key = key < cur ? cur : key;
                }
                tileImproved = 1;
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          }
        }
      } else {
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        unsigned long long best = kEmptyKey;
        if constexpr (kFilter) {
          // rowNeed was published before this tile's barriers.
          const float need = rowNeed[T::rowInTile(r)];
          const float limit = need < 0.0f ? -INFINITY : metric.accBound(weight.distBound(need));
#pragma unroll
          for (int c = 0; c < TR; ++c) {
            const float acc = w[r][c];
            if (!(acc > limit) && colComp[c] != rowComp[r]) {
              const float ww = weight(metric.finish(acc), rowVal[r], colVal[c]);
              if (rowComp[r] >= 0 && colComp[c] >= 0 && ww != INFINITY)
                best = minKey(best, packKey(ww, colIdx[c]));
            }
          }
        } else {
#pragma unroll
          for (int c = 0; c < TR; ++c) {
            const float ww = weight(w[r][c], rowVal[r], colVal[c]);
            const bool counts = rowComp[r] >= 0 && colComp[c] >= 0 &&
                                colComp[c] != rowComp[r] && ww != INFINITY;
            if (counts)
              best = minKey(best, packKey(ww, colIdx[c]));
          }
        }
        // Wave-uniform: most rows of most tiles have no candidate.
        if (__ballot(best != kEmptyKey) != 0) {
          best = rowGroupMin(best);
          if (T::tc() == 0 && best != kEmptyKey) {
            const int i = T::rowInTile(r);
            if (best < rowBest[i]) {
              rowBest[i] = best;
              tileImproved = 1;
            }
          }
        }
      }
      }
      __syncthreads();
      // The needs and the bound change only with a row's best: a tile that
      // improved none skips the update and its barrier. Every wave reads the
      // flag after the barrier above; it is cleared only after the barrier
      // below (all reads done) and set again only after the next tile's
      // first barrier, so every wave takes the same branch.
      if (tileImproved) {
        if (threadIdx.x < 64) {
          const float b = newBound();
          if (threadIdx.x == 0)
            bound = b;
        }
        __syncthreads();
        if (threadIdx.x == 0)
          tileImproved = 0;
      }
    }
    __syncthreads(); // lowerBound is rewritten by the next batch
  }
  if (KEEP) {
    for (int e = threadIdx.x; e < T::DIM * kKeptKeys; e += T::THREADS) {
      const int i = e / kKeptKeys;
      if (row0 + i < nRows)
        kept[(size_t)posOf(row0 + i) * kKeptKeys + e % kKeptKeys] = rowTop[e];
    }
    for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
      rowBest[i] = rowTop[i * kKeptKeys];
  }
  for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
    if (row0 + i < nRows && rowBest[i] != kEmptyKey)
      atomicMin(&nearest[posOf(row0 + i)], rowBest[i]);
  if (tilesComputed && threadIdx.x == 0)
    atomicAdd(tilesComputed, computed);
}

// Seeds each component's best known edge weight for the next pass from the
// previous pass's results: a point whose previous candidate lies in another
// component still has that edge. With few components the workgroup first
// combines its points in LDS (one global atomic per component per workgroup).
constexpr int kSeedLds = 4096;

__global__ void __launch_bounds__(256)
    seedComponentBestKernel(const unsigned long long *__restrict__ nearest,
                            const int *__restrict__ comp, const int *__restrict__ posOf, int n,
                            int numComp, unsigned int *compBest) {
  __shared__ unsigned int local[kSeedLds];
  const bool useLds = numComp <= kSeedLds;
  if (useLds) {
    for (int c = threadIdx.x; c < numComp; c += blockDim.x)
      local[c] = 0xffffffffu;
    __syncthreads();
  }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const unsigned long long key = nearest[p];
    if (key == kEmptyKey)
      continue;
    const int q = (int)keyIndex(key);
    const int c = comp[p];
    if (comp[posOf ? posOf[q] : q] == c)
      continue; // the partner joined this component
    const unsigned int bits = (unsigned int)(key >> 32);
    if (useLds)
      atomicMin(&local[c], bits);
    else
      atomicMin(&compBest[c], bits);
  }
  if (useLds) {
    __syncthreads();
    for (int c = threadIdx.x; c < numComp; c += blockDim.x)
      if (local[c] != 0xffffffffu)
        atomicMin(&compBest[c], local[c]);
  }
}

__global__ void inverseOrderKernel(const int *__restrict__ orig, int n, int *__restrict__ posOf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n)
    posOf[orig[p]] = p;
}

__global__ void gatherRowsKernel(const float *__restrict__ src, const int *__restrict__ order,
                                 int n, int d, float *__restrict__ dst) {
  const size_t total = (size_t)n * d;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t row = e / d, f = e - row * d;
    dst[e] = src[(size_t)order[row] * d + f];
  }
}

__global__ void iotaKernel(int *out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = i;
}

template <class Metric, class F> int withWeightFor(const Metric &, WeightSpec w, F &&f) {
  switch (w.kind) {
  case kWeightPlain: return f(PlainWeight{});
  case kWeightMReach:
    return w.alpha == 1.0f ? f(MReachWeightT<false>{w.core, 1.0f})
                           : f(MReachWeightT<true>{w.core, w.alpha});
  default: return kErrInvalidAlgorithm;
  }
}

} // namespace

// ---------------------------------------------------------------------------
// Host entry points

int pairwiseMatrix(const float *A, int nA, const float *B, int nB, int d, MetricSpec m,
                   float *out) {
  return withMetric(m, [&](auto metric) {
    using Metric = decltype(metric);
    return withFill(d, [&](auto fill) {
      using T = typename TileFor<Metric, decltype(fill)::value>::type;
      const int tilesA = (nA + T::DIM - 1) / T::DIM, tilesB = (nB + T::DIM - 1) / T::DIM;
      const long long blocks = (long long)tilesA * tilesB;
      if (blocks > 0x7fffffffLL)
        return (int)hipErrorInvalidConfiguration;
      hipLaunchKernelGGL((pairwiseMatrixKernel<T, Metric>), dim3((unsigned)blocks), dim3(256), 0,
                         stream(), A, nA, B, nB, d, metric, out, tilesB);
      return (int)hipGetLastError();
    });
  });
}

int condensedMatrix(const float *X, int n, int d, MetricSpec m, float *out) {
  return withMetric(m, [&](auto metric) {
    using Metric = decltype(metric);
    return withFill(d, [&](auto fill) {
      using T = typename TileFor<Metric, decltype(fill)::value>::type;
      const long long tiles = (n + T::DIM - 1) / T::DIM;
      const long long pairs = tiles * (tiles + 1) / 2;
      hipLaunchKernelGGL((condensedMatrixKernel<T, Metric>), dim3(persistentBlocks(pairs, 8)),
                         dim3(256), 0, stream(), X, n, d, metric, out, pairs);
      return (int)hipGetLastError();
    });
  });
}

int nearestOtherSymmetric(const float *X, int n, int d, MetricSpec m, WeightSpec w,
                          const int *comp, const int *orig, unsigned long long *nearest) {
  if (w.kind == kWeightWard) {
    if (m.id != kSqEuclidean)
      return kErrInvalidMetric;
    WardWeight ward{w.size};
    return withFill(d, [&](auto fill) {
      using T = Tile<8, decltype(fill)::value>;
      const long long tiles = (n + T::DIM - 1) / T::DIM, pairs = tiles * (tiles + 1) / 2;
      hipLaunchKernelGGL((nearestOtherSymmetricKernel<T, SqEuclidean, WardWeight>),
                         dim3(persistentBlocks(pairs, 8)), dim3(256), 0, stream(), X, n, d,
                         SqEuclidean{}, ward, comp, orig, nearest, pairs);
      return (int)hipGetLastError();
    });
  }
  return withMetric(m, [&](auto metric) {
    using Metric = decltype(metric);
    return withWeightFor(metric, w, [&](auto weight) {
      using Weight = decltype(weight);
      return withFill(d, [&](auto fill) {
        using T = typename TileFor<Metric, decltype(fill)::value>::type;
        const long long tiles = (n + T::DIM - 1) / T::DIM, pairs = tiles * (tiles + 1) / 2;
        PhaseTimer timer(kPhaseDistanceTiles);
        addUnits(kPhaseDistanceTiles, (long long)n * (n + 1) / 2);
        hipLaunchKernelGGL((nearestOtherSymmetricKernel<T, Metric, Weight>),
                           dim3(persistentBlocks(pairs, 8)), dim3(256), 0, stream(), X, n, d,
                           metric, weight, comp, orig, nearest, pairs);
        return (int)hipGetLastError();
      });
    });
  });
}

bool shouldPrune(int mode, int metricId, int n, int d) {
  const int tiles = (n + 127) / 128;
  if (!metricHasBoxBound(metricId) || mode == kPruneNever || tiles < 2 ||
      d > kPruneHardMaxFeatures)
    return false;
  return mode == kPruneAlways || (n >= kPruneMinPoints && d <= kPruneMaxFeatures);
}

// k-d order: split each range at a tile boundary near its middle along the
// feature with the widest spread (estimated from a sample), until each range
// is one tile (the behaviour of reference src/spatial.cuh:48-91). The order
// inside the partitions of nth_element is implementation-defined; ties between
// equal weights always break by original index, so results do not depend on it.
// The top levels run on one thread until there are enough ranges, then the
// ranges are finished on up to 16 threads (disjoint ranges: the same order).
void spatialOrderIndices(const float *pts, int d, int *idx, int count) {
  constexpr int kTile = 128, kSample = 64;
  struct Range {
    int begin, end;
  };
  auto split = [&](Range range, std::vector<Range> &out) {
    const int length = range.end - range.begin;
    if (length <= kTile)
      return;
    int feature = 0;
    float widest = -1.0f;
    const int step = std::max(1, length / kSample);
    for (int f = 0; f < d; ++f) {
      float lo = INFINITY, hi = -INFINITY;
      for (int i = range.begin; i < range.end; i += step) {
        const float v = pts[(size_t)idx[i] * d + f];
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
      if (hi - lo > widest) {
        widest = hi - lo;
        feature = f;
      }
    }
    const int tiles = (length + kTile - 1) / kTile;
    const int middle = range.begin + tiles / 2 * kTile;
    std::nth_element(idx + range.begin, idx + middle, idx + range.end, [&](int x, int y) {
      return pts[(size_t)x * d + feature] < pts[(size_t)y * d + feature];
    });
    out.push_back({range.begin, middle});
    out.push_back({middle, range.end});
  };
  auto finish = [&](Range range) {
    std::vector<Range> stack{range};
    while (!stack.empty()) {
      const Range r = stack.back();
      stack.pop_back();
      split(r, stack);
    }
  };
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // Runs job(i) for i < count on the pool (disjoint ranges of idx).
  auto parallel = [&](size_t jobs, auto &&job) {
    std::atomic<size_t> cursor{0};
    auto worker = [&]() {
      for (size_t i = cursor++; i < jobs; i = cursor++)
        job(i);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < std::min<size_t>(threads, jobs); ++t)
      pool.emplace_back(worker);
    worker();
    for (auto &t : pool)
      t.join();
  };
  // Breadth first, one level at a time, each level's ranges in parallel,
  // until there are enough ranges to finish them depth first in parallel.
  std::vector<Range> level{{0, count}};
  while (level.size() < 4 * threads) {
    std::vector<std::vector<Range>> children(level.size());
    parallel(level.size(), [&](size_t i) { split(level[i], children[i]); });
    std::vector<Range> next;
    for (const auto &c : children)
      next.insert(next.end(), c.begin(), c.end());
    if (next.empty())
      return; // every range is one tile (ranges of one tile are final)
    level.swap(next);
  }
  parallel(level.size(), [&](size_t i) { finish(level[i]); });
}

std::vector<int> spatialOrder(const float *pts, int n, int d) {
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  spatialOrderIndices(pts, d, order.data(), n);
  return order;
}

int gatherRows(const float *src, const int *order, int n, int d, float *dst) {
  hipLaunchKernelGGL(gatherRowsKernel, dim3(std::max(1, std::min(4096, (n * d + 255) / 256))),
                     dim3(256), 0, stream(), src, order, n, d, dst);
  return (int)hipGetLastError();
}

namespace {

struct DeviceBoxes {
  DevArray<float> low, high, minValue;
  DevArray<int> minComp, maxComp;
  Boxes view;
  int alloc(int tiles, int d) {
    Status s;
    s(low.alloc((size_t)tiles * d));
    s(high.alloc((size_t)tiles * d));
    s(minValue.alloc(tiles));
    s(minComp.alloc(tiles));
    s(maxComp.alloc(tiles));
    view = Boxes{low, high, minValue, minComp, maxComp, tiles};
    return s.code;
  }
};

// Walk steps computed per row tile by the sampled estimate (stride > 1); later
// steps are only counted against the bound the computed ones reached (an
// overestimate of the pruned share: the bound would keep shrinking).
constexpr int kEstimateSteps = 64;

template <class T, class Metric, class Weight>
int runWalk(const float *X, int n, int d, const Metric &metric, const Weight &weight,
            const int *comp, const int *orig, unsigned int *compBest,
            unsigned long long *nearest, double *share, PassStats *stats, int stride,
            const int *skipComp, const unsigned char *closedRows,
            unsigned long long *kept = nullptr) {
  const int tiles = (n + T::DIM - 1) / T::DIM;
  DeviceBoxes boxes;
  DevArray<unsigned long long> counter;
  Status s;
  s(boxes.alloc(tiles, d));
  s(counter.alloc(1));
  if (!s.ok())
    return s.code;
  s(fill(counter, 0, sizeof(unsigned long long)));
  hipLaunchKernelGGL((buildBoxesKernel<Weight>), dim3(tiles), dim3(64), 0, stream(), X, n, d,
                     T::DIM, weight, comp, nullptr, boxes.view);
  s((int)hipGetLastError());
  const int sampled = (tiles + stride - 1) / stride;
  {
    PhaseTimer timer(kPhaseDistanceTiles);
    bool launched = false;
    if constexpr (Weight::kHasValue) {
      if (kept) {
        hipLaunchKernelGGL((nearestOtherWalkKernel<T, Metric, Weight, false, true>),
                           dim3(sampled), dim3(256), 2 * d * sizeof(float), stream(), X, n, d,
                           metric, weight, comp, orig, boxes.view, nullptr, nearest,
                           counter.get(), stride, stride > 1 ? kEstimateSteps : 0x7fffffff,
                           skipComp, closedRows, RowSet{}, kept);
        launched = true;
      }
    }
    if (!launched)
      hipLaunchKernelGGL((nearestOtherWalkKernel<T, Metric, Weight>), dim3(sampled), dim3(256),
                         2 * d * sizeof(float), stream(), X, n, d, metric, weight, comp, orig,
                         boxes.view, compBest, nearest, counter.get(), stride,
                         stride > 1 ? kEstimateSteps : 0x7fffffff, skipComp, closedRows,
                         RowSet{}, nullptr);
    s((int)hipGetLastError());
  }
  unsigned long long computed = 0;
  s(download(&computed, counter, sizeof(computed)));
  const unsigned long long total = (unsigned long long)sampled * tiles;
  *share = total ? (double)computed / (double)total : 1.0;
  addUnits(kPhaseDistanceTiles, (long long)computed * T::DIM * T::DIM);
  if (stats && s.ok()) {
    stats->tilesComputed += computed;
    stats->tilesTotal += total;
  }
  return s.code;
}

// Open rows of a walk: not closed and not in the skipped component.
__global__ void openRowFlagsKernel(const int *__restrict__ comp, const int *__restrict__ skipComp,
                                   const unsigned char *__restrict__ closedRows, int n,
                                   int *flags) {
  const int skip = skipComp ? *skipComp : -1;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= n; p += gridDim.x * blockDim.x)
    flags[p] = p < n && comp[p] != skip && !closedRows[p];
}

__global__ void openRowListKernel(const int *__restrict__ flags, const int *__restrict__ index,
                                  int n, int *rowPos) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    if (flags[p])
      rowPos[index[p]] = p;
}

// The full-order tile of each compact row tile's middle row.
__global__ void rowTileCentersKernel(const int *__restrict__ rowPos, int nOpen, int tileDim,
                                     int rowTiles, int *center) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < rowTiles)
    center[t] = rowPos[min(t * tileDim + tileDim / 2, nOpen - 1)] / tileDim;
}

// The walk over the open rows only, when closed rows are given: those rows
// (in the full order, so spatially coherent) are gathered into compact row
// tiles with their own boxes, each walking the column tiles around the full
// order's tile of its middle row. With most rows closed, a tile of the full
// order usually keeps a few open rows and walks as far as they need; the
// compact tiles walk for open rows only. The share reported is of the full
// grid's tile pairs.
template <class T, class Metric, class Weight>
int runWalkOpen(const float *X, int n, int d, const Metric &metric, const Weight &weight,
                const int *comp, const int *orig, unsigned int *compBest,
                unsigned long long *nearest, double *share, PassStats *stats,
                const int *skipComp, const unsigned char *closedRows) {
  const int tiles = (n + T::DIM - 1) / T::DIM;
  Status s;
  DevArray<int> flags, index, rowPos;
  DevArray<unsigned char> temp;
  s(flags.alloc((size_t)n + 1));
  s(index.alloc((size_t)n + 1));
  s(rowPos.alloc(n));
  s(temp.alloc(scanTempBytes(n + 1)));
  if (!s.ok())
    return s.code;
  const int blocks = std::max(1, std::min(4 * numCUs(), (n + 256) / 256));
  hipLaunchKernelGGL(openRowFlagsKernel, dim3(blocks), dim3(256), 0, stream(), comp, skipComp,
                     closedRows, n, flags.get());
  s(exclusiveScan(flags, index, n + 1, temp));
  int nOpen = 0;
  s(download(&nOpen, index.get() + n, sizeof(int)));
  if (!s.ok())
    return s.code;
  if (4LL * nOpen > 3LL * n) // most rows open: the full order's tiles
    return runWalk<T>(X, n, d, metric, weight, comp, orig, compBest, nearest, share, stats, 1,
                      skipComp, closedRows);
  *share = 0.0;
  if (nOpen == 0)
    return 0;
  hipLaunchKernelGGL(openRowListKernel, dim3(blocks), dim3(256), 0, stream(), flags.get(),
                     index.get(), n, rowPos.get());
  const int rowTiles = (nOpen + T::DIM - 1) / T::DIM;
  DeviceBoxes colBoxes, rowBoxes;
  DevArray<float> rowX;
  DevArray<int> center;
  DevArray<unsigned long long> counter;
  s(colBoxes.alloc(tiles, d));
  s(rowBoxes.alloc(rowTiles, d));
  s(rowX.alloc((size_t)nOpen * d));
  s(center.alloc(rowTiles));
  s(counter.alloc(1));
  if (!s.ok())
    return s.code;
  s(fill(counter, 0, sizeof(unsigned long long)));
  s(gatherRows(X, rowPos, nOpen, d, rowX));
  hipLaunchKernelGGL((buildBoxesKernel<Weight>), dim3(tiles), dim3(64), 0, stream(), X, n, d,
                     T::DIM, weight, comp, nullptr, colBoxes.view);
  hipLaunchKernelGGL((buildBoxesKernel<Weight>), dim3(rowTiles), dim3(64), 0, stream(),
                     rowX.get(), nOpen, d, T::DIM, weight, comp, rowPos.get(), rowBoxes.view);
  hipLaunchKernelGGL(rowTileCentersKernel, dim3((rowTiles + 255) / 256), dim3(256), 0, stream(),
                     rowPos.get(), nOpen, T::DIM, rowTiles, center.get());
  s((int)hipGetLastError());
  RowSet rows;
  rows.X = rowX;
  rows.n = nOpen;
  rows.pos = rowPos;
  rows.boxes = rowBoxes.view;
  rows.center = center;
  {
    PhaseTimer timer(kPhaseDistanceTiles);
    hipLaunchKernelGGL((nearestOtherWalkKernel<T, Metric, Weight, true>), dim3(rowTiles),
                       dim3(256), 2 * d * sizeof(float), stream(), X, n, d, metric, weight, comp,
                       orig, colBoxes.view, compBest, nearest, counter.get(), 1, 0x7fffffff,
                       skipComp, closedRows, rows, nullptr);
    s((int)hipGetLastError());
  }
  unsigned long long computed = 0;
  s(download(&computed, counter, sizeof(computed)));
  const unsigned long long total = (unsigned long long)tiles * tiles;
  *share = (double)computed / (double)total;
  addUnits(kPhaseDistanceTiles, (long long)computed * T::DIM * T::DIM);
  if (stats && s.ok()) {
    stats->tilesComputed += computed;
    stats->tilesTotal += total;
  }
  return s.code;
}

} // namespace

int nearestOtherWalk(const float *X, int n, int d, MetricSpec m, WeightSpec w, const int *comp,
                     const int *orig, unsigned int *compBest, unsigned long long *nearest,
                     double *share, PassStats *stats, int stride, const int *skipComp,
                     const unsigned char *closedRows, unsigned long long *kept) {
  if (!metricHasBoxBound(m.id))
    return kErrInvalidMetric;
  if (kept && (w.kind != kWeightMReach || closedRows || stride != 1))
    return kErrInvalidAlgorithm;
  if (w.kind == kWeightWard) {
    if (m.id != kSqEuclidean)
      return kErrInvalidMetric;
    return withFill(d, [&](auto fill) {
      using T = Tile<8, decltype(fill)::value>;
      return runWalk<T>(X, n, d, SqEuclidean{}, WardWeight{w.size}, comp, orig, compBest,
                        nearest, share, stats, stride, skipComp, closedRows);
    });
  }
  return withMetric(m, [&](auto metric) {
    using Metric = decltype(metric);
    if constexpr (!Metric::kBoxBound) {
      return kErrInvalidMetric;
    } else {
      return withWeightFor(metric, w, [&](auto weight) {
        return withFill(d, [&](auto fill) {
          using T = Tile<8, decltype(fill)::value>;
          // Closed rows come with mutual-reachability weights (their answers
          // settled from candidate lists): walk the open rows only.
          if constexpr (decltype(weight)::kHasValue) {
            if (closedRows && stride == 1)
              return runWalkOpen<T>(X, n, d, metric, weight, comp, orig, compBest, nearest,
                                    share, stats, skipComp, closedRows);
          }
          return runWalk<T>(X, n, d, metric, weight, comp, orig, compBest, nearest, share,
                            stats, stride, skipComp, closedRows, kept);
        });
      });
    }
  });
}

namespace {

__global__ void settleFromKeptKernel(const unsigned long long *__restrict__ kept, int n,
                                     const int *__restrict__ comp, const int *__restrict__ posOf,
                                     const int *__restrict__ skipComp,
                                     unsigned long long *__restrict__ nearest,
                                     unsigned int *__restrict__ compBest,
                                     unsigned char *__restrict__ closed) {
  const int skip = skipComp ? *skipComp : -1;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int c = comp[p];
    bool settled = false;
    if (c != skip) {
#pragma unroll
      for (int q = 0; q < kKeptKeys && !settled; ++q) {
        const unsigned long long key = kept[(size_t)p * kKeptKeys + q];
        if (key == kEmptyKey)
          break;
        const int j = posOf ? posOf[keyIndex(key)] : (int)keyIndex(key);
        if (comp[j] != c) {
          nearest[p] = key;
          if (compBest)
            atomicMin(&compBest[c], (unsigned int)(key >> 32));
          settled = true;
        }
      }
    }
    closed[p] = settled ? 1 : 0;
  }
}

} // namespace

int settleFromKept(const unsigned long long *kept, int n, const int *comp, const int *posOf,
                   const int *skipComp, unsigned long long *nearest, unsigned int *compBest,
                   unsigned char *closed) {
  const int blocks = std::max(1, std::min(4 * numCUs(), (n + 255) / 256));
  hipLaunchKernelGGL(settleFromKeptKernel, dim3(blocks), dim3(256), 0, stream(), kept, n, comp,
                     posOf, skipComp, nearest, compBest, closed);
  return (int)hipGetLastError();
}

int seedComponentBest(const unsigned long long *nearest, const int *comp, const int *posOf,
                      int n, int numComp, bool fromNearest, unsigned int *compBest) {
  Status st;
  st(fill(compBest, 0xff, (size_t)numComp * sizeof(unsigned int)));
  if (fromNearest) {
    const int blocks = std::max(1, std::min(4 * numCUs(), (n + 255) / 256));
    hipLaunchKernelGGL(seedComponentBestKernel, dim3(blocks), dim3(256), 0, stream(), nearest,
                       comp, posOf, n, numComp, compBest);
    st((int)hipGetLastError());
  }
  return st.code;
}

int inverseOrder(const int *orig, int n, int *posOf) {
  hipLaunchKernelGGL(inverseOrderKernel, dim3((n + 255) / 256), dim3(256), 0, stream(), orig, n,
                     posOf);
  return (int)hipGetLastError();
}

namespace {

// Component sizes; with few components each workgroup counts in LDS first
// (same-address atomics from every point would serialise).
constexpr int kSizeLds = 4096;

__global__ void __launch_bounds__(256)
    componentSizesKernel(const int *__restrict__ comp, int n, int numComp, int *sizes) {
  __shared__ int local[kSizeLds];
  const bool useLds = numComp <= kSizeLds;
  if (useLds) {
    for (int c = threadIdx.x; c < numComp; c += blockDim.x)
      local[c] = 0;
    __syncthreads();
  }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    if (useLds)
      atomicAdd(&local[comp[p]], 1);
    else
      atomicAdd(&sizes[comp[p]], 1);
  }
  if (useLds) {
    __syncthreads();
    for (int c = threadIdx.x; c < numComp; c += blockDim.x)
      if (local[c])
        atomicAdd(&sizes[c], local[c]);
  }
}

// At most one component holds more than half of the points.
__global__ void majorityKernel(const int *__restrict__ sizes, int numComp, int n,
                               int *majority) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < numComp && 2 * (long long)sizes[c] > (long long)n)
    *majority = c;
}

} // namespace

int majorityComponent(const int *comp, int n, int numComp, int *sizes, int *majority) {
  Status s;
  s(fill(majority, 0xff, sizeof(int))); // -1
  if (2 * (long long)numComp > (long long)n + 1) // every component has at most n - numComp + 1 points
    return s.code;
  s(fill(sizes, 0, (size_t)numComp * sizeof(int)));
  const int blocks = std::max(1, std::min(1024, (n + 255) / 256));
  hipLaunchKernelGGL(componentSizesKernel, dim3(blocks), dim3(256), 0, stream(), comp, n, numComp,
                     sizes);
  hipLaunchKernelGGL(majorityKernel, dim3((numComp + 255) / 256), dim3(256), 0, stream(), sizes,
                     numComp, n, majority);
  s((int)hipGetLastError());
  return s.code;
}

int estimatePrunedShare(const float *X, int n, int d, MetricSpec m, double *share) {
  constexpr int kSampleTiles = 32;
  const int tiles = (n + 127) / 128;
  const int stride = std::max(1, tiles / kSampleTiles);
  DevArray<int> comp, orig;
  DevArray<unsigned long long> nearest;
  Status s;
  s(comp.alloc(n));
  s(orig.alloc(n));
  s(nearest.alloc(n));
  if (!s.ok())
    return s.code;
  hipLaunchKernelGGL(iotaKernel, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, stream(),
                     comp.get(), n);
  hipLaunchKernelGGL(iotaKernel, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, stream(),
                     orig.get(), n);
  s(fill(nearest, 0xff, nearest.bytes()));
  s(nearestOtherWalk(X, n, d, m, WeightSpec{}, comp, orig, nullptr, nearest, share, nullptr,
                     stride));
  return s.code;
}

} // namespace cuhc
