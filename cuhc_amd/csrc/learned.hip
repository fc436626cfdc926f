// Learned-radius core distances. No reference counterpart; the values equal
// the fused and radix methods bit for bit.
//
// 1. Features per point, from the k-d tiles (points in spatial order): the
//    log of the k-th smallest distance within its own tile and within its tile
//    and the two neighbouring tiles, the log of the tile's box diagonal and of
//    the point's distance to the tile centroid.
// 2. A training sample of 64 row tiles (8,192 points) spread evenly over
//    the k-d order gets its exact core distances (the pruned fused walk of
//    those tiles; full rows when k > 64); the network (mlp.hip) is trained on
//    the GPU to predict their log, and the proposal margin is the 99.5%
//    residual quantile on the sample plus 0.02.
// 3. Every point gets a radius r_i = exp(prediction + margin). One workgroup
//    per row tile walks the column tiles and skips those whose box lower
//    bound (with the pruning slack) exceeds every row's radius; every
//    distance <= r_i (the point itself as 0) is appended to the row's
//    candidate list (512 slots per row for k <= 64, up to 4,096 above).
// 4. A wave per row radix-selects the k-th smallest candidate (32 one-bit
//    passes with 64-bit ballots). The list holds every point within r_i, so
//    when it has at least k entries (and did not overflow) the k-th smallest
//    is the exact core distance. Other rows (the proposal was too small or too
//    large) are recomputed by the exact full search.
#include "mlp.hpp"
#include "mst.hpp"
#include "tile.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

namespace cuhc {

namespace {

// Candidate slots per row: 512 up to k = 64 (the fused top-k's range), 32 k
// rounded up to a power of two above (at most 4,096). The radius carries a
// margin over the predicted core distance, and in d dimensions the points
// inside grow like the margin's d-th power (at d = 8 a 22% margin holds about
// five times k points, and the margin grows with the spread of the residuals).
int candSlotsFor(int k) {
  int cap = 512;
  while (k > 64 && cap < 32 * k && cap < 4096)
    cap *= 2;
  return cap;
}
constexpr int kTrainSamples = 8192;
std::atomic<int> gTrainSteps{300};
constexpr float kTrainRate = 0.01f;

std::mutex gLearnedMutex;
LearnedStats gLearned;

__global__ void tileStatsKernel(const float *__restrict__ X, int n, int d, int tileDim,
                                float *low, float *high, float *centroid, float *diag) {
  const int tile = blockIdx.x;
  const int first = tile * tileDim, last = min(n, first + tileDim);
  __shared__ float span2[64];
  float acc = 0.0f;
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    float lo = INFINITY, hi = -INFINITY;
    double sum = 0.0;
    for (int p = first; p < last; ++p) {
      const float v = X[(size_t)p * d + f];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
      sum += v;
    }
    low[(size_t)tile * d + f] = lo;
    high[(size_t)tile * d + f] = hi;
    centroid[(size_t)tile * d + f] = (float)(sum / (last - first));
    acc += (hi - lo) * (hi - lo);
  }
  span2[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float total = 0.0f;
    for (int i = 0; i < (int)blockDim.x; ++i)
      total += span2[i];
    diag[tile] = sqrtf(total);
  }
}

__global__ void featuresKernel(const float *__restrict__ X, int n, int d, int tileDim,
                               const float *local1, const float *local3, const float *centroid,
                               const float *diag, float *feats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int tile = i / tileDim;
  float c2 = 0.0f;
  for (int f = 0; f < d; ++f) {
    const float diff = X[(size_t)i * d + f] - centroid[(size_t)tile * d + f];
    c2 = fmaf(diff, diff, c2);
  }
  const float tiny = 1e-30f;
  // A tile with fewer than k points (the last one) has no own-tile k-th
  // distance; the three-tile value stands in (finite whenever k <= n).
  const float own = local1[i] == INFINITY ? local3[i] : local1[i];
  feats[(size_t)i * 4 + 0] = logf(own + tiny);
  feats[(size_t)i * 4 + 1] = logf(local3[i] + tiny);
  feats[(size_t)i * 4 + 2] = logf(diag[tile] + tiny);
  feats[(size_t)i * 4 + 3] = logf(sqrtf(c2) + tiny);
}

__global__ void normalizeKernel(float *feats, int n, const float *mean, const float *inv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * 4)
    return;
  feats[e] = (feats[e] - mean[e & 3]) * inv[e & 3];
}

__global__ void radiusKernel(const float *pred, int n, float scale, float shift, float *radius) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  // r = exp(pred * sigma + mu + margin), never below 0
  radius[i] = expf(fmaf(pred[i], scale, shift));
}

HD float learnedSlack(int d) { return 1e-5f + (d + 8) * 5e-7f; }

template <class Metric>
DEV float boxBound(const Metric &metric, const float *lowA, const float *highA, const float *lowB,
                   const float *highB, int d) {
  typename Metric::Acc acc = metric.init();
  for (int k = 0; k < d; ++k) {
    const float gap = fmaxf(0.0f, fmaxf(lowB[k] - highA[k], lowA[k] - highB[k]));
    acc = metric.fold(acc, gap, 0.0f);
  }
  return metric.finish(acc);
}

DEV int walkStep(int rowTile, int step, int numTiles) {
  const int off = (step + 1) >> 1;
  const int tile = (step & 1) ? rowTile + off : rowTile - off;
  return (tile >= 0 && tile < numTiles) ? tile : -1;
}

template <class T, class Metric>
__global__ void __launch_bounds__(256)
    radiusGatherKernel(const float *__restrict__ X, int n, int d, Metric metric,
                       const float *__restrict__ radius, const float *__restrict__ low,
                       const float *__restrict__ high, float *__restrict__ cand,
                       int *__restrict__ candPos, int *__restrict__ count, int cap,
                       unsigned long long *tilesComputed) {
  constexpr int TR = T::TR;
  __shared__ __align__(16) TileSmem<T> sm;
  __shared__ int cnt[T::DIM];
  __shared__ float rmaxWave[4];
  __shared__ bool skipStep[64];
  extern __shared__ float rowBox[]; // low[d], high[d]
  const int numTiles = (n + T::DIM - 1) / T::DIM;
  const int rowTile = blockIdx.x, row0 = rowTile * T::DIM;
  for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
    cnt[i] = 0;
  for (int k = threadIdx.x; k < d; k += T::THREADS) {
    rowBox[k] = low[(size_t)rowTile * d + k];
    rowBox[d + k] = high[(size_t)rowTile * d + k];
  }
  float rRow[TR];
  float rmax = 0.0f;
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int row = row0 + T::rowInTile(r);
    rRow[r] = row < n ? radius[row] : -1.0f;
    rmax = fmaxf(rmax, rRow[r]);
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    rmax = fmaxf(rmax, __shfl_xor(rmax, m, 64));
  if ((threadIdx.x & 63) == 0)
    rmaxWave[threadIdx.x >> 6] = rmax;
  __syncthreads();
  rmax = fmaxf(fmaxf(rmaxWave[0], rmaxWave[1]), fmaxf(rmaxWave[2], rmaxWave[3]));
  const float slack = 1.0f - learnedSlack(d);
  unsigned long long computed = 0;

  // The radii are fixed, so wave 0 computes the skip decisions of 64 walk
  // steps at once (one barrier per 64 steps, none per skipped step).
  const int steps = 2 * numTiles;
  for (int base = 0; base < steps; base += 64) {
    if (threadIdx.x < 64) {
      const int step = base + (int)threadIdx.x;
      const int colTile = step < steps ? walkStep(rowTile, step, numTiles) : -1;
      skipStep[threadIdx.x] =
          colTile < 0 || boxBound(metric, rowBox, rowBox + d, low + (size_t)colTile * d,
                                  high + (size_t)colTile * d, d) * slack > rmax;
    }
    __syncthreads();
    const int batchEnd = min(steps, base + 64);
    for (int step = base; step < batchEnd; ++step) {
      if (skipStep[step - base])
        continue;
      const int colTile = walkStep(rowTile, step, numTiles);
      ++computed;
      const int col0 = colTile * T::DIM;
      // Raw accumulators: only the pairs within the bound of the row's radius
      // are finished (sqrt) and tested exactly.
      float v[TR][TR];
      computeTile<T, Metric, false>(sm, X, n, X, n, d, metric, row0, col0, v);
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int rit = T::rowInTile(r), row = row0 + rit;
        if (row >= n)
          continue;
        const float limit = metric.accBound(rRow[r]);
#pragma unroll
        for (int c = 0; c < TR; ++c) {
          const int col = col0 + T::colInTile(c);
          if (col >= n || v[r][c] > limit)
            continue;
          const float val = col == row ? 0.0f : fmaxf(metric.finish(v[r][c]), 0.0f);
          if (val <= rRow[r]) {
            const int pos = atomicAdd(&cnt[rit], 1);
            if (pos < cap) {
              cand[(size_t)row * cap + pos] = val;
              if (candPos)
                candPos[(size_t)row * cap + pos] = col;
            }
          }
        }
      }
    }
    __syncthreads(); // skipStep is rewritten by the next batch
  }
  __syncthreads();
  for (int i = threadIdx.x; i < T::DIM; i += T::THREADS)
    if (row0 + i < n)
      count[row0 + i] = cnt[i];
  if (threadIdx.x == 0)
    atomicAdd(tilesComputed, computed);
}

// One wave per row: the k-th smallest of the row's candidates by a binary
// radix select over the (non-negative) FP32 bit patterns; rows that cannot be
// answered from their list go to the fallback list.
template <int PER> // candidate slots per lane (the row's cap / 64)
__global__ void __launch_bounds__(256)
    candidateSelectKernel(const float *__restrict__ cand, const int *__restrict__ count, int n,
                          int k, float *__restrict__ core, int *fallback, int *numFallback,
                          unsigned long long *candidates) {
  constexpr int kCap = 64 * PER;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n)
    return;
  const int c = count[row];
  if (lane == 0) // striped over 64 cache lines: one address would serialise a million waves
    atomicAdd(&candidates[(row & 63) * 16], (unsigned long long)c);
  if (c < k || c > kCap) {
    if (lane == 0)
      fallback[atomicAdd(numFallback, 1)] = row;
    return;
  }
  unsigned int bits[PER];
  bool valid[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int e = lane + 64 * j;
    valid[j] = e < c;
    bits[j] = valid[j] ? __float_as_uint(cand[(size_t)row * kCap + e]) : 0u;
  }
  unsigned int prefix = 0, mask = 0;
  int remaining = k;
  for (int b = 31; b >= 0; --b) {
    const unsigned int bit = 1u << b;
    int zeros = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const bool match = valid[j] && (bits[j] & mask) == prefix && (bits[j] & bit) == 0u;
      zeros += __popcll(__ballot(match));
    }
    if (zeros < remaining) {
      prefix |= bit;
      remaining -= zeros;
    }
    mask |= bit;
  }
  if (lane == 0)
    core[row] = __uint_as_float(prefix);
}

__global__ void remapListsKernel(int *__restrict__ pos, const int *__restrict__ count, int n,
                                 int cap, const int *__restrict__ oldToNew) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * cap)
    return;
  const int row = (int)(e / cap), slot = (int)(e - (size_t)row * cap);
  if (slot < min(count[row], cap))
    pos[e] = oldToNew[pos[e]];
}

// One wave per row (see settleFromLists).
__global__ void __launch_bounds__(256)
    settleFromListsKernel(const float *__restrict__ dist, const int *__restrict__ pos,
                          const int *__restrict__ count, const float *__restrict__ radius,
                          const int *__restrict__ rowMap, int cap, int n,
                          const int *__restrict__ comp, const int *__restrict__ orig,
                          const float *__restrict__ core, const int *__restrict__ skipComp,
                          unsigned long long *__restrict__ nearest,
                          unsigned int *__restrict__ compBest, unsigned char *__restrict__ closed,
                          unsigned long long *counters) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (p >= n)
    return;
  const int c = comp[p];
  const int skip = skipComp ? *skipComp : -1;
  bool settled = false;
  if (c != skip) {
    const int row = rowMap ? rowMap[p] : p;
    const int cnt = count[row];
    if (cnt <= cap) {
      const MReachWeightT<false> weight{core, 1.0f};
      const float ci = core[p];
      unsigned long long best = kEmptyKey;
      for (int e = lane; e < cnt; e += 64) {
        const int j = pos[(size_t)row * cap + e];
        if (comp[j] == c)
          continue;
        const float w = weight(dist[(size_t)row * cap + e], ci, core[j]);
        if (w != INFINITY)
          best = minKey(best, packKey(w, orig ? (unsigned int)orig[j] : (unsigned int)j));
      }
#pragma unroll
      for (int m = 32; m > 0; m >>= 1)
        best = minKey(best, shflXorKey(best, m));
      settled = best != kEmptyKey && keyWeight(best) <= radius[row];
      if (settled && lane == 0) {
        nearest[p] = best;
        if (compBest)
          atomicMin(&compBest[c], (unsigned int)(best >> 32));
      }
    }
    if (lane == 0) // striped over 64 lines: one address would serialise every row
      atomicAdd(&counters[(p & 63) * 16 + 1], 1ull);
  }
  if (lane == 0) {
    closed[p] = settled ? 1 : 0;
    if (settled)
      atomicAdd(&counters[(p & 63) * 16], 1ull);
  }
}

template <class F> int withBoxMetric(MetricSpec m, F &&f) {
  switch (m.id) {
  case kEuclidean: return f(Euclidean{});
  case kSqEuclidean: return f(SqEuclidean{});
  case kManhattan: return f(Manhattan{});
  case kChebyshev: return f(Chebyshev{});
  case kMinkowski: return f(Minkowski{m.p});
  default: return kErrInvalidMetric;
  }
}

} // namespace

void learnedStatsAdd(const LearnedStats &s) {
  std::lock_guard<std::mutex> lock(gLearnedMutex);
  gLearned.proposals += s.proposals;
  gLearned.hits += s.hits;
  gLearned.fallbacks += s.fallbacks;
  gLearned.candidates += s.candidates;
  gLearned.tiles += s.tiles;
  gLearned.settled += s.settled;
  gLearned.considered += s.considered;
}

int remapRadiusLists(RadiusLists &lists, int n, const int *oldToNew) {
  if (!lists.pos.get() || lists.cap == 0)
    return 0;
  const size_t total = (size_t)n * lists.cap;
  hipLaunchKernelGGL(remapListsKernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     stream(), lists.pos.get(), lists.count.get(), n, lists.cap, oldToNew);
  return (int)hipGetLastError();
}

int settleFromLists(const RadiusLists &lists, int n, const int *comp, const int *orig,
                    const float *core, const int *skipComp, unsigned long long *nearest,
                    unsigned int *compBest, unsigned char *closed, unsigned long long *counters) {
  hipLaunchKernelGGL(settleFromListsKernel, dim3((n + 3) / 4), dim3(256), 0, stream(),
                     lists.dist.get(), lists.pos.get(), lists.count.get(), lists.radius.get(),
                     lists.rowMap, lists.cap, n, comp, orig, core, skipComp, nearest, compBest,
                     closed, counters);
  return (int)hipGetLastError();
}

int coreLearned(const float *X, int n, int d, MetricSpec m, int k, float *core, bool spatial,
                RadiusLists *keep) {
  // Exact in any order; the proposals are only sharp, and the fallback search
  // only pruned, for points in k-d order (spatial).
  if (!metricHasBoxBound(m.id))
    return coreDistances(X, n, d, m, k, core, kCoreAuto, false);
  constexpr int DIM = 128;
  const int tiles = (n + DIM - 1) / DIM;
  Status s;
  DevArray<float> local1, local3, low, high, centroid, diag, feats, params, pred;
  DevArray<int> fallback, numFallback;
  RadiusLists own;
  RadiusLists &lists = keep ? *keep : own;
  DevArray<float> &radius = lists.radius, &cand = lists.dist;
  DevArray<int> &count = lists.count;
  DevArray<unsigned long long> counters;
  s(local1.alloc(n));
  s(local3.alloc(n));
  s(low.alloc((size_t)tiles * d));
  s(high.alloc((size_t)tiles * d));
  s(centroid.alloc((size_t)tiles * d));
  s(diag.alloc(tiles));
  s(feats.alloc((size_t)n * 4));
  s(params.alloc(kMlpParams));
  s(radius.alloc(n));
  s(pred.alloc(n));
  s(count.alloc(n));
  s(fallback.alloc(n));
  s(numFallback.alloc(1));
  s(counters.alloc(1 + 64 * 16)); // [0] tiles computed, then 64 striped candidate counts
  if (!s.ok())
    return s.code;
  PhaseTimer timer(kPhaseCore);
  // 1. Features.
  s(coreLocal(X, n, d, m, k, 1, local1));
  s(coreLocal(X, n, d, m, k, 3, local3));
  hipLaunchKernelGGL(tileStatsKernel, dim3(tiles), dim3(64), 0, stream(), X, n, d, DIM,
                     low.get(), high.get(), centroid.get(), diag.get());
  hipLaunchKernelGGL(featuresKernel, dim3((n + 255) / 256), dim3(256), 0, stream(), X, n, d, DIM,
                     local1.get(), local3.get(), centroid.get(), diag.get(), feats.get());
  s((int)hipGetLastError());
  // 2. Training sample with exact core distances (written into core[], which
  //    is overwritten below): whole row tiles spread evenly over the order, so
  //    that the exact search is the pruned walk of those tiles.
  const int sampleTiles = std::min(tiles, kTrainSamples / DIM);
  std::vector<int> tileList(sampleTiles), rows;
  for (int i = 0; i < sampleTiles; ++i) {
    tileList[i] = (int)(((long long)i * tiles) / sampleTiles);
    for (int p = tileList[i] * DIM; p < std::min(n, (tileList[i] + 1) * DIM); ++p)
      rows.push_back(p);
  }
  const int ms = (int)rows.size();
  s(coreForTiles(X, n, d, m, k, tileList.data(), sampleTiles, core));
  std::vector<float> hf((size_t)n * 4), hcore(n);
  s(download(hf.data(), feats, hf.size() * sizeof(float)));
  s(download(hcore.data(), core, n * sizeof(float)));
  if (!s.ok())
    return s.code;
  // Normalisation from the sample.
  std::vector<float> xs((size_t)ms * 4), ts(ms);
  double mean[4] = {0, 0, 0, 0}, var[4] = {0, 0, 0, 0}, tm = 0, tv = 0;
  for (int i = 0; i < ms; ++i) {
    for (int f = 0; f < 4; ++f)
      mean[f] += hf[(size_t)rows[i] * 4 + f];
    tm += std::log((double)hcore[rows[i]] + 1e-30);
  }
  for (int f = 0; f < 4; ++f)
    mean[f] /= ms;
  tm /= ms;
  for (int i = 0; i < ms; ++i) {
    for (int f = 0; f < 4; ++f) {
      const double z = hf[(size_t)rows[i] * 4 + f] - mean[f];
      var[f] += z * z;
    }
    const double z = std::log((double)hcore[rows[i]] + 1e-30) - tm;
    tv += z * z;
  }
  float hmean[4], hinv[4];
  for (int f = 0; f < 4; ++f) {
    hmean[f] = (float)mean[f];
    const double sd = std::sqrt(var[f] / ms);
    hinv[f] = sd > 1e-12 ? (float)(1.0 / sd) : 1.0f;
  }
  const double tsd = std::max(std::sqrt(tv / ms), 1e-6);
  for (int i = 0; i < ms; ++i) {
    for (int f = 0; f < 4; ++f)
      xs[(size_t)i * 4 + f] = (hf[(size_t)rows[i] * 4 + f] - hmean[f]) * hinv[f];
    ts[i] = (float)((std::log((double)hcore[rows[i]] + 1e-30) - tm) / tsd);
  }
  DevArray<float> dxs, dts, dmean, dinv, dpredS;
  s(dxs.alloc(xs.size()));
  s(dts.alloc(ms));
  s(dmean.alloc(4));
  s(dinv.alloc(4));
  s(dpredS.alloc(ms));
  s(upload(dxs, xs.data(), xs.size() * sizeof(float)));
  s(upload(dts, ts.data(), ms * sizeof(float)));
  s(upload(dmean, hmean, sizeof(hmean)));
  s(upload(dinv, hinv, sizeof(hinv)));
  // 3. Train on the GPU; margin from the sample residuals.
  std::vector<float> p0(kMlpParams);
  mlpInitHost(p0.data(), 12345u);
  s(upload(params, p0.data(), kMlpParams * sizeof(float)));
  s(mlpTrain(params, dxs, dts, ms, gTrainSteps.load(), kTrainRate));
  s(mlpForward(params, dxs, ms, dpredS));
  std::vector<float> ps(ms);
  s(download(ps.data(), dpredS, ms * sizeof(float)));
  if (!s.ok())
    return s.code;
  std::vector<double> res(ms);
  for (int i = 0; i < ms; ++i)
    res[i] = ((double)ts[i] - ps[i]) * tsd;
  std::sort(res.begin(), res.end());
  const double margin = std::max(0.0, res[std::min(ms - 1, (int)(0.995 * ms))]) + 0.02;
  // 4. Radii for every point, gather, select, fallback.
  hipLaunchKernelGGL(normalizeKernel, dim3((n * 4 + 255) / 256), dim3(256), 0, stream(),
                     feats.get(), n, dmean.get(), dinv.get());
  s(mlpForward(params, feats, n, pred));
  hipLaunchKernelGGL(radiusKernel, dim3((n + 255) / 256), dim3(256), 0, stream(), pred.get(), n,
                     (float)tsd, (float)(tm + margin), radius.get());
  const int cap = candSlotsFor(k);
  s(cand.alloc((size_t)n * cap));
  if (keep)
    s(lists.pos.alloc((size_t)n * cap));
  lists.cap = cap;
  s(fill(numFallback, 0, sizeof(int)));
  s(fill(counters, 0, counters.bytes()));
  if (!s.ok())
    return s.code;
  s(withBoxMetric(m, [&](auto metric) {
    using Metric = decltype(metric);
    auto run = [&](auto fill) {
      using T = Tile<8, decltype(fill)::value>;
      hipLaunchKernelGGL((radiusGatherKernel<T, Metric>), dim3(tiles), dim3(256),
                         2 * d * sizeof(float), stream(), X, n, d, metric, radius.get(),
                         low.get(), high.get(), cand.get(), lists.pos.get(), count.get(), cap,
                         counters.get());
      return (int)hipGetLastError();
    };
    return d <= 8 ? run(std::integral_constant<int, 8>{}) : run(std::integral_constant<int, 32>{});
  }));
  auto select = [&](auto per) {
    hipLaunchKernelGGL((candidateSelectKernel<decltype(per)::value>), dim3((n + 3) / 4),
                       dim3(256), 0, stream(), cand.get(), count.get(), n, k, core,
                       fallback.get(), numFallback.get(), counters.get() + 1);
  };
  switch (cap) {
  case 512: select(std::integral_constant<int, 8>{}); break;
  case 1024: select(std::integral_constant<int, 16>{}); break;
  case 2048: select(std::integral_constant<int, 32>{}); break;
  default: select(std::integral_constant<int, 64>{}); break;
  }
  s((int)hipGetLastError());
  int nf = 0;
  s(download(&nf, numFallback, sizeof(int)));
  if (s.ok() && nf > 0) {
    // The failed rows in ascending position (k-d order): tiles of them are
    // spatially compact enough for the pruned search.
    std::vector<int> rowsFailed(nf);
    s(download(rowsFailed.data(), fallback, nf * sizeof(int)));
    std::sort(rowsFailed.begin(), rowsFailed.end());
    s(upload(fallback, rowsFailed.data(), nf * sizeof(int)));
    if (s.ok())
      s(coreForRows(X, n, d, m, k, fallback, nf, core, spatial));
  }
  std::vector<unsigned long long> ctr(counters.count);
  s(download(ctr.data(), counters, counters.bytes()));
  LearnedStats st;
  st.proposals = (unsigned long long)n;
  st.fallbacks = (unsigned long long)nf;
  st.hits = (unsigned long long)(n - nf);
  for (int k = 0; k < 64; ++k)
    st.candidates += ctr[1 + k * 16];
  st.tiles = ctr[0];
  if (s.ok())
    learnedStatsAdd(st);
  s(sync());
  return s.code;
}

} // namespace cuhc

extern "C" void cuhc_learned_stats(unsigned long long *proposals, unsigned long long *hits,
                                   unsigned long long *fallbacks, unsigned long long *candidates,
                                   unsigned long long *tiles, unsigned long long *settled,
                                   unsigned long long *considered) {
  std::lock_guard<std::mutex> lock(cuhc::gLearnedMutex);
  *proposals = cuhc::gLearned.proposals;
  *hits = cuhc::gLearned.hits;
  *fallbacks = cuhc::gLearned.fallbacks;
  *candidates = cuhc::gLearned.candidates;
  *tiles = cuhc::gLearned.tiles;
  if (settled)
    *settled = cuhc::gLearned.settled;
  if (considered)
    *considered = cuhc::gLearned.considered;
  cuhc::gLearned = cuhc::LearnedStats{};
}

extern "C" int cuhc_set_learned_training(int steps) {
  if (steps < 0)
    return cuhc::kErrInvalidShape;
  cuhc::gTrainSteps.store(steps);
  return 0;
}

extern "C" int cuhc_mlp_gradient(const float *params, const float *X, const float *t, int m,
                                 float *grad, float *loss) {
  using namespace cuhc;
  if (!params || !X || !t || !grad || m < 1)
    return kErrInvalidShape;
  DevArray<float> dp, dx, dt, work, dg, dl;
  Status s;
  s(dp.alloc(kMlpParams));
  s(dx.alloc((size_t)m * kMlpIn));
  s(dt.alloc(m));
  s(work.alloc(mlpWorkFloats(m)));
  s(dg.alloc(kMlpParams));
  s(dl.alloc(1));
  s(upload(dp, params, kMlpParams * sizeof(float)));
  s(upload(dx, X, (size_t)m * kMlpIn * sizeof(float)));
  s(upload(dt, t, m * sizeof(float)));
  if (s.ok())
    s(mlpGradient(dp, dx, dt, m, work, dg, dl));
  s(download(grad, dg, kMlpParams * sizeof(float)));
  if (loss)
    s(download(loss, dl, sizeof(float)));
  return s.code;
}

extern "C" int cuhc_mlp_forward(const float *params, const float *X, int m, float *y) {
  using namespace cuhc;
  if (!params || !X || !y || m < 1)
    return kErrInvalidShape;
  DevArray<float> dp, dx, dy;
  Status s;
  s(dp.alloc(kMlpParams));
  s(dx.alloc((size_t)m * kMlpIn));
  s(dy.alloc(m));
  s(upload(dp, params, kMlpParams * sizeof(float)));
  s(upload(dx, X, (size_t)m * kMlpIn * sizeof(float)));
  if (s.ok())
    s(mlpForward(dp, dx, m, dy));
  s(download(y, dy, m * sizeof(float)));
  return s.code;
}

extern "C" int cuhc_mlp_train(float *params, const float *X, const float *t, int m, int steps,
                              float lr) {
  using namespace cuhc;
  if (!params || !X || !t || m < 1 || steps < 0)
    return kErrInvalidShape;
  DevArray<float> dp, dx, dt;
  Status s;
  s(dp.alloc(kMlpParams));
  s(dx.alloc((size_t)m * kMlpIn));
  s(dt.alloc(m));
  s(upload(dp, params, kMlpParams * sizeof(float)));
  s(upload(dx, X, (size_t)m * kMlpIn * sizeof(float)));
  s(upload(dt, t, m * sizeof(float)));
  if (s.ok())
    s(mlpTrain(dp, dx, dt, m, steps, lr));
  s(download(params, dp, kMlpParams * sizeof(float)));
  return s.code;
}

extern "C" int cuhc_mlp_param_count(void) { return cuhc::kMlpParams; }
