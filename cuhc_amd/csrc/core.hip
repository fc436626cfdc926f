// Core distances: core[i] = the k-th smallest distance from point i to the
// points, the point itself counted as the 1st (distance 0), values clamped at
// 0 (reference src/hdbscan.cuh:35-338).
//
// fused (k <= 64): one workgroup per row tile walks the column tiles nearest
//   first and keeps each row's k smallest distances sorted in LDS. After each
//   tile the 16 lanes of a row mark the distances that beat the row's k-th,
//   and the row's lowest lane with work inserts one value per step (64-bit
//   ballot over the wave: 4 rows insert in parallel). Optional box pruning for
//   points in spatial order.
// radix: distances of a block of rows into a bounded buffer, then per row a
//   4-pass MSB-first byte radix select over the FP32 bit patterns with
//   per-wave LDS histograms.
// Both are exact and bit-identical (the k-th smallest value is unique).
#include "mst.hpp"
#include "tile.hpp"

#include <algorithm>
#include <type_traits>

namespace cuhc {

namespace {

// Column tile boxes (low/high per tile, numTiles > 0 turns pruning on), and
// for gathered rows the boxes of their row tiles (rowLow/rowHigh; null: the
// rows are the points themselves and share the column tiles' boxes).
struct CoreBoxes {
  const float *low = nullptr, *high = nullptr;
  int numTiles = 0;
  const float *rowLow = nullptr, *rowHigh = nullptr;
};

DEV int walkTileCore(int rowTile, int step, int numTiles) {
  const int off = (step + 1) >> 1;
  const int tile = (step & 1) ? rowTile + off : rowTile - off;
  return (tile >= 0 && tile < numTiles) ? tile : -1;
}

HD float coreBoxSlack(int d) { return 1e-5f + (d + 8) * 5e-7f; }

template <class Metric>
DEV float coreBoxBound(const Metric &metric, const float *lowA, const float *highA,
                       const float *lowB, const float *highB, int d) {
  typename Metric::Acc acc = metric.init();
  for (int k = 0; k < d; ++k) {
    const float gap = fmaxf(0.0f, fmaxf(lowB[k] - highA[k], lowA[k] - highB[k]));
    acc = metric.fold(acc, gap, 0.0f);
  }
  return metric.finish(acc);
}

// Inserts v into the ascending list[0..k), dropping the largest, if v beats list[k-1].
DEV void insertSorted(float *list, int k, float v) {
  if (!(v < list[k - 1]))
    return;
  int pos = k - 1;
  while (pos > 0 && list[pos - 1] > v) {
    list[pos] = list[pos - 1];
    --pos;
  }
  list[pos] = v;
}

// Rows are the points themselves (rowsX null: row tile blockIdx.x of X, or
// tile tileList[blockIdx.x] with a list) or listed points (rowsX: numRows
// gathered rows, rowPos their positions in X, for the exact search of
// selected points).
// Four waves per SIMD for the one-fill tiles (d <= 8) of the squared-sum
// metrics: the pass waits on tile loads, and the extra wave hides more of them
// than its few spills cost (1M points, d = 8: 195 to 168 ms). The other folds
// would spill far more at 128 VGPRs.
template <class T, class Metric> constexpr int coreMinWaves() {
  return T::FILL <= 8 &&
                 (std::is_same<Metric, Euclidean>::value || std::is_same<Metric, SqEuclidean>::value)
             ? 4
             : 1;
}

template <class T, class Metric>
__global__ void __launch_bounds__(256, (coreMinWaves<T, Metric>()))
    kthNearestFusedKernel(const float *__restrict__ X, int n, int d, Metric metric, int k,
                          float *__restrict__ core, CoreBoxes boxes, int stepLimit,
                          const float *__restrict__ rowsX, int numRows,
                          const int *__restrict__ rowPos, const int *__restrict__ tileList) {
  constexpr int TR = T::TR;
  constexpr int kBatch = 64; // walk steps whose box bounds wave 0 computes at once
  __shared__ __align__(16) TileSmem<T> sm;
  __shared__ float lowerBound[kBatch];
  __shared__ float largestKth; // the largest k-th distance over the tile's rows
  __shared__ int inserted;     // a row's list changed in the current tile
  extern __shared__ float lists[]; // [DIM][k], then the row tile's box (2d) if pruning

  const int rowTile = tileList ? tileList[blockIdx.x] : (int)blockIdx.x;
  const int row0 = rowTile * T::DIM;
  const float *A = rowsX ? rowsX : X;
  const int nA = rowsX ? numRows : n;
  for (int i = threadIdx.x; i < T::DIM * k; i += T::THREADS)
    lists[i] = INFINITY;
  const bool prune = boxes.numTiles > 0;
  float *rowBox = lists + T::DIM * k;
  if (prune) {
    const float *rl = boxes.rowLow ? boxes.rowLow : boxes.low;
    const float *rh = boxes.rowHigh ? boxes.rowHigh : boxes.high;
    for (int f = threadIdx.x; f < d; f += T::THREADS) {
      rowBox[f] = rl[(size_t)rowTile * d + f];
      rowBox[d + f] = rh[(size_t)rowTile * d + f];
    }
  }
  if (threadIdx.x == 0) {
    largestKth = INFINITY; // every list starts empty
    inserted = 0;
  }
  // The 8-feature tiles serve d <= 8 only (one fill): the row slice is
  // staged once for every tile.
  constexpr bool kRowsResident = T::FILL == 8;
  if (kRowsResident)
    stageRows<T>(sm, A, nA, d, row0);
  __syncthreads();
  const float slack = 1.0f - coreBoxSlack(d);
  const int numTiles = (n + T::DIM - 1) / T::DIM;
  const int lane = threadIdx.x & 63;
  const unsigned long long groupMask = 0xffffull << (lane & 48);
  // Listed rows have no home tile: the walk starts anywhere and covers all;
  // with pruning (rows listed in point order) it starts at the column tile of
  // the tile's first row, which is near the others.
  const int homeTile =
      rowsX ? (prune ? rowPos[row0] / T::DIM : (int)(blockIdx.x % numTiles)) : rowTile;

  // A step is skipped when the box bound between the row tile and its column
  // tile exceeds the largest k-th distance of the tile's rows: wave 0 computes
  // the bounds of 64 steps at once into LDS, and the largest k-th distance
  // is refreshed only after a tile that changed a list, so skipped steps and
  // unchanged tiles cost no barrier.
  const int steps = min(2 * numTiles, stepLimit);
  for (int base = 0; base < steps; base += kBatch) {
    if (prune) {
      if (threadIdx.x < kBatch) {
        const int step = base + (int)threadIdx.x;
        const int colTile = step < steps ? walkTileCore(homeTile, step, numTiles) : -1;
        lowerBound[threadIdx.x] =
            colTile >= 0 ? coreBoxBound(metric, rowBox, rowBox + d, boxes.low + (size_t)colTile * d,
                                        boxes.high + (size_t)colTile * d, d) * slack
                         : 0.0f;
      }
      __syncthreads();
    }
    const int batchEnd = min(steps, base + kBatch);
    for (int step = base; step < batchEnd; ++step) {
      const int colTile = walkTileCore(homeTile, step, numTiles);
      if (colTile < 0)
        continue;
      if (prune) {
        const float largest = largestKth; // uniform: LDS, written only between barriers
        if (largest != INFINITY && lowerBound[step - base] > largest)
          continue;
      }
      const int col0 = colTile * T::DIM;
      // With an accumulator bound the tile holds raw accumulators, and only the
      // pairs within the bound of the row's current k-th distance are finished
      // (most pairs of a tile are not, which skips their sqrt).
      constexpr bool kFilter = Metric::kAccFilter;
      float v[TR][TR];
      computeTile<T, Metric, !kFilter, kRowsResident>(sm, A, nA, X, n, d, metric, row0, col0, v);
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int rit = T::rowInTile(r);
        const int row = row0 + rit;
        float *list = lists + rit * k;
        unsigned int pending = 0;
        if (row < nA) {
          const float kth = list[k - 1];
          const int self = rowsX ? rowPos[row] : row;
          if constexpr (kFilter) {
            const float limit = metric.accBound(kth);
#pragma unroll
            for (int c = 0; c < TR; ++c) {
              const int col = col0 + T::colInTile(c);
              const float acc = v[r][c];
              if (col < n && !(acc > limit)) {
                const float value = col == self ? 0.0f : fmaxf(metric.finish(acc), 0.0f);
                v[r][c] = value;
                if (value < kth)
                  pending |= 1u << c;
              }
            }
          } else {
#pragma unroll
            for (int c = 0; c < TR; ++c) {
              const int col = col0 + T::colInTile(c);
              // The point itself is distance 0 (drops float noise such as cosine's 1e-7).
              v[r][c] = col == self ? 0.0f : fmaxf(v[r][c], 0.0f);
              if (col < n && v[r][c] < kth)
                pending |= 1u << c;
            }
          }
        }
        // One insertion per row per step until no lane of the wave has work;
        // all lanes stay converged for the ballot.
        while (true) {
          const unsigned long long work = __ballot(pending != 0);
          if (work == 0)
            break;
          const unsigned long long mine = work & groupMask;
          if (mine != 0 && lane == __ffsll((long long)mine) - 1) {
            const int c = __ffs(pending) - 1;
            pending &= pending - 1;
            float value = 0.0f;
#pragma unroll
            for (int cc = 0; cc < TR; ++cc)
              if (cc == c)
                value = v[r][cc];
            insertSorted(list, k, value);
            if (prune)
              inserted = 1;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      }
      if (prune) {
        __syncthreads(); // insertions done
        // Every wave reads the flag after the barrier above; it is cleared
        // after the barrier below and set again only after the next tile's
        // first barrier, so every wave takes the same branch.
        if (inserted) {
          if (threadIdx.x < 64) {
            float largest = 0.0f;
            for (int i = threadIdx.x; i < T::DIM; i += 64)
              if (row0 + i < nA)
                largest = fmaxf(largest, lists[i * k + k - 1]);
#pragma unroll
            for (int m = 32; m > 0; m >>= 1)
              largest = fmaxf(largest, __shfl_xor(largest, m, 64));
            if (threadIdx.x == 0)
              largestKth = largest;
          }
          __syncthreads();
          if (threadIdx.x == 0)
            inserted = 0;
        }
      }
    }
    if (prune)
      __syncthreads(); // lowerBound is rewritten by the next batch
  }
  __syncthreads();
  for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
    if (row0 + i < nA)
      core[rowsX ? rowPos[row0 + i] : row0 + i] = lists[i * k + k - 1];
}

__global__ void coreBoxesKernel(const float *__restrict__ X, int n, int d, int tileDim,
                                float *low, float *high) {
  const int tile = blockIdx.x;
  const int first = tile * tileDim, last = min(n, first + tileDim);
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    float lo = INFINITY, hi = -INFINITY;
    for (int p = first; p < last; ++p) {
      const float v = X[(size_t)p * d + f];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    low[(size_t)tile * d + f] = lo;
    high[(size_t)tile * d + f] = hi;
  }
}

// One workgroup (256 threads) per row of the distance block. The block's row r
// is point rowList[firstRow + r] (or firstRow + r without a list); its core
// distance goes to core[that point].
__global__ void __launch_bounds__(256)
    kthSmallestRowKernel(const float *__restrict__ dist, int numCols, int firstRow, int k,
                         const int *__restrict__ rowList, float *__restrict__ core) {
  __shared__ unsigned int hist[4][256];
  __shared__ unsigned int prefix, prefixMask, remaining;
  const int row = blockIdx.x;
  const int self = rowList ? rowList[firstRow + row] : firstRow + row;
  const float *rd = dist + (size_t)row * numCols;
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    prefix = 0;
    prefixMask = 0;
    remaining = (unsigned)k;
  }
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 4 * 256; i += 256)
      (&hist[0][0])[i] = 0;
    __syncthreads();
    const unsigned int pre = prefix, mask = prefixMask;
    for (int col = threadIdx.x; col < numCols; col += 256) {
      const float value = col == self ? 0.0f : fmaxf(rd[col], 0.0f);
      const unsigned int bits = __float_as_uint(value);
      if ((bits & mask) == pre)
        atomicAdd(&hist[wave][(bits >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned int below = 0, bin = 0;
      while (true) {
        const unsigned int count = hist[0][bin] + hist[1][bin] + hist[2][bin] + hist[3][bin];
        if (below + count >= remaining || bin == 255)
          break;
        below += count;
        ++bin;
      }
      remaining -= below;
      prefix |= bin << shift;
      prefixMask |= 0xffu << shift;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    core[self] = __uint_as_float(prefix);
}

template <class F> int withMetricCore(MetricSpec m, F &&f) {
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

// tileList (device, numListed entries): only those row tiles.
int coreFused(const float *X, int n, int d, MetricSpec m, int k, float *core, bool prune,
              int stepLimit = 0x7fffffff, const int *tileList = nullptr, int numListed = 0) {
  return withMetricCore(m, [&](auto metric) {
    using Metric = decltype(metric);
    auto run = [&](auto fill) {
      using T = typename TileFor<Metric, decltype(fill)::value>::type;
      const int tiles = (n + T::DIM - 1) / T::DIM;
      size_t smem = (size_t)T::DIM * k * sizeof(float);
      CoreBoxes boxes;
      DevArray<float> low, high;
      Status s;
      if (prune && Metric::kBoxBound) {
        s(low.alloc((size_t)tiles * d));
        s(high.alloc((size_t)tiles * d));
        if (!s.ok())
          return s.code;
        hipLaunchKernelGGL(coreBoxesKernel, dim3(tiles), dim3(64), 0, stream(), X, n, d, T::DIM,
                           low.get(), high.get());
        boxes = CoreBoxes{low, high, tiles};
        smem += 2 * d * sizeof(float);
      }
      PhaseTimer timer(kPhaseCore);
      hipLaunchKernelGGL((kthNearestFusedKernel<T, Metric>), dim3(tileList ? numListed : tiles),
                         dim3(256), smem, stream(), X, n, d, metric, k, core, boxes, stepLimit,
                         nullptr, 0, nullptr, tileList);
      s((int)hipGetLastError());
      s(sync());
      return s.code;
    };
    if (d <= 8)
      return run(std::integral_constant<int, 8>{});
    return run(std::integral_constant<int, 32>{});
  });
}

// Distance rows per radix-select launch: a quarter of the free device memory
// (at most 64 GiB; at n = 10^6 that is 16,384 rows of 4 MB, so one launch
// fills the chip many times over), at least one row, a multiple of the tile
// rows when possible, never more than wanted.
int radixChunkRows(int n, int wanted, int tileRows) {
  size_t freeBytes = 0, totalBytes = 0;
  if (hipMemGetInfo(&freeBytes, &totalBytes) != hipSuccess) {
    (void)hipGetLastError();
    freeBytes = (size_t)4 << 30;
  }
  const size_t budget = std::min(freeBytes / 4, (size_t)64 << 30);
  const size_t rowBytes = (size_t)n * sizeof(float);
  size_t rows = std::max<size_t>(1, budget / rowBytes);
  if (rows >= (size_t)tileRows)
    rows = rows / tileRows * tileRows;
  return (int)std::min<size_t>(rows, (size_t)wanted);
}

int coreRadix(const float *X, int n, int d, MetricSpec m, int k, float *core) {
  const int rows = radixChunkRows(n, n, tilePoints(m.id));
  DevArray<float> chunk;
  Status s;
  s(chunk.alloc((size_t)rows * n));
  PhaseTimer timer(kPhaseCore);
  for (int first = 0; s.ok() && first < n; first += rows) {
    const int count = std::min(rows, n - first);
    s(pairwiseMatrix(X + (size_t)first * d, count, X, n, d, m, chunk));
    hipLaunchKernelGGL(kthSmallestRowKernel, dim3(count), dim3(256), 0, stream(), chunk.get(), n,
                       first, k, nullptr, core);
    s((int)hipGetLastError());
  }
  s(sync());
  return s.code;
}

} // namespace

int coreLocal(const float *X, int n, int d, MetricSpec m, int k, int stepLimit, float *out) {
  return coreFused(X, n, d, m, std::min(k, kMaxFusedK), out, false, stepLimit);
}

namespace {

// Exact core distances of listed rows with the fused top-k lists (k <= 64):
// the rows are gathered. Without pruning every column tile is visited; with
// it (points in k-d order, rows listed in ascending position, a box metric)
// each tile of 128 gathered rows gets its own box and skips the column tiles
// whose bound exceeds its largest k-th distance, as the tiles of the points
// themselves do.
int coreForRowsFused(const float *X, int n, int d, MetricSpec m, int k, const int *rows,
                     int numRows, float *core, bool prune) {
  DevArray<float> rowData;
  Status s;
  s(rowData.alloc((size_t)numRows * d));
  if (s.ok())
    s(gatherRows(X, rows, numRows, d, rowData));
  if (!s.ok())
    return s.code;
  return withMetricCore(m, [&](auto metric) {
    using Metric = decltype(metric);
    auto run = [&](auto fill) {
      using T = typename TileFor<Metric, decltype(fill)::value>::type;
      const int rowTiles = (numRows + T::DIM - 1) / T::DIM;
      size_t smem = (size_t)T::DIM * k * sizeof(float);
      CoreBoxes boxes;
      DevArray<float> low, high, rowLow, rowHigh;
      Status st;
      if (prune && Metric::kBoxBound) {
        const int tiles = (n + T::DIM - 1) / T::DIM;
        st(low.alloc((size_t)tiles * d));
        st(high.alloc((size_t)tiles * d));
        st(rowLow.alloc((size_t)rowTiles * d));
        st(rowHigh.alloc((size_t)rowTiles * d));
        if (!st.ok())
          return st.code;
        hipLaunchKernelGGL(coreBoxesKernel, dim3(tiles), dim3(64), 0, stream(), X, n, d, T::DIM,
                           low.get(), high.get());
        hipLaunchKernelGGL(coreBoxesKernel, dim3(rowTiles), dim3(64), 0, stream(),
                           rowData.get(), numRows, d, T::DIM, rowLow.get(), rowHigh.get());
        boxes = CoreBoxes{low, high, tiles, rowLow, rowHigh};
        smem += 2 * d * sizeof(float);
      }
      PhaseTimer timer(kPhaseCore);
      hipLaunchKernelGGL((kthNearestFusedKernel<T, Metric>), dim3(rowTiles), dim3(256), smem,
                         stream(), X, n, d, metric, k, core, boxes, 0x7fffffff,
                         rowData.get(), numRows, rows, nullptr);
      st((int)hipGetLastError());
      st(sync());
      return st.code;
    };
    if (d <= 8)
      return run(std::integral_constant<int, 8>{});
    return run(std::integral_constant<int, 32>{});
  });
}

} // namespace

int coreForRows(const float *X, int n, int d, MetricSpec m, int k, const int *rows, int numRows,
                float *core, bool prune) {
  if (numRows <= 0)
    return 0;
  if (k <= kMaxFusedK)
    return coreForRowsFused(X, n, d, m, k, rows, numRows, core, prune);
  const int chunkRows = radixChunkRows(n, numRows, 1);
  DevArray<float> chunk, rowData;
  Status s;
  s(chunk.alloc((size_t)chunkRows * n));
  s(rowData.alloc((size_t)chunkRows * d));
  for (int first = 0; s.ok() && first < numRows; first += chunkRows) {
    const int count = std::min(chunkRows, numRows - first);
    s(gatherRows(X, rows + first, count, d, rowData));
    s(pairwiseMatrix(rowData, count, X, n, d, m, chunk));
    hipLaunchKernelGGL(kthSmallestRowKernel, dim3(count), dim3(256), 0, stream(), chunk.get(), n,
                       first, k, rows, core);
    s((int)hipGetLastError());
  }
  s(sync());
  return s.code;
}

int coreForTiles(const float *X, int n, int d, MetricSpec m, int k, const int *hostTiles,
                 int numTiles, float *core) {
  constexpr int kDim = 128; // the fused kernel's row tile for box metrics
  if (numTiles <= 0)
    return 0;
  Status s;
  if (k <= kMaxFusedK && metricHasBoxBound(m.id)) {
    DevArray<int> list;
    s(list.alloc(numTiles));
    s(upload(list, hostTiles, numTiles * sizeof(int)));
    if (s.ok())
      s(coreFused(X, n, d, m, k, core, true, 0x7fffffff, list, numTiles));
    return s.code;
  }
  std::vector<int> rows;
  for (int t = 0; t < numTiles; ++t)
    for (int p = hostTiles[t] * kDim; p < std::min(n, (hostTiles[t] + 1) * kDim); ++p)
      rows.push_back(p);
  DevArray<int> dRows;
  s(dRows.alloc(rows.size()));
  s(upload(dRows, rows.data(), rows.size() * sizeof(int)));
  if (s.ok())
    s(coreForRows(X, n, d, m, k, dRows, (int)rows.size(), core));
  return s.code;
}

int coreDistances(const float *X, int n, int d, MetricSpec m, int k, float *core, int method,
                  bool prune) {
  if (k < 1 || k > n)
    return kErrInvalidShape;
  if (method == kCoreFused && k > kMaxFusedK)
    return kErrInvalidAlgorithm;
  if (method == kCoreLearned)
    return coreLearned(X, n, d, m, k, core, prune);
  if (method == kCoreFused || (method == kCoreAuto && k <= kMaxFusedK))
    return coreFused(X, n, d, m, k, core, prune);
  return coreRadix(X, n, d, m, k, core);
}

} // namespace cuhc
