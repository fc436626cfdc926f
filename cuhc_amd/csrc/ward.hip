// Exact Ward linkage without the n x n matrix: rounds of reciprocal nearest
// neighbours (reference src/rnn_ward.cuh) and the nearest-neighbour chain
// (src/nn_chain.cuh).
//
// Precision: cluster centroids live in FP64 on the GPU and every merge updates
// them in FP64 (c_a + (c_b - c_a) * n_b / (n_a + n_b), exact for duplicate
// centroids); merge heights are accumulated in FP64,
// sqrt(2 n_a n_b / (n_a + n_b) * sum (c_a - c_b)^2). The RNN rounds choose
// nearest neighbours with the FP32 all-pairs pass on an FP32 copy of the
// centroids and Ward's FP32 weight (symmetric bit for bit, so the closest
// pair always sees each other and every round makes progress).
#include "linkage.hpp"
#include "mst.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <numeric>

namespace cuhc {

namespace {

namespace cg = cooperative_groups;

DEV int reciprocalPartner(const unsigned long long *nearest, int i) {
  const unsigned long long k = nearest[i];
  if (k == kEmptyKey)
    return -1;
  const int j = keyIndex(k);
  return keyIndex(nearest[j]) == i && nearest[j] != kEmptyKey ? j : -1;
}

__global__ void survivorsKernel(const unsigned long long *nearest, int m, int *survives) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m)
    return;
  const int j = reciprocalPartner(nearest, i);
  survives[i] = !(j >= 0 && j < i);
}

// Exclusive scan, three launches: per-block scan of 1024 values, a one-block
// scan of the block totals, and the offsets added back.
constexpr int kScanBlock = 1024;

__global__ void __launch_bounds__(kScanBlock) scanBlocksKernel(const int *in, int m, int *out,
                                                               int *blockTotals) {
  __shared__ int waveTotals[kScanBlock / 64];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  const int v = i < m ? in[i] : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v; // inclusive scan inside the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    waveTotals[wave] = x;
  __syncthreads();
  if (wave == 0) {
    int t = lane < kScanBlock / 64 ? waveTotals[lane] : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(t, off, 64);
      if (lane >= off)
        t += y;
    }
    if (lane < kScanBlock / 64)
      waveTotals[lane] = t; // inclusive over waves
  }
  __syncthreads();
  const int before = wave > 0 ? waveTotals[wave - 1] : 0;
  if (i < m)
    out[i] = before + x - v;
  if (threadIdx.x == kScanBlock - 1)
    blockTotals[blockIdx.x] = before + x;
}

__global__ void __launch_bounds__(kScanBlock) scanTotalsKernel(int *totals, int numBlocks) {
  // Sequential over chunks of 1024 (numBlocks is small: m / 1024).
  __shared__ int carry;
  __shared__ int waveTotals[kScanBlock / 64];
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  for (int base = 0; base < numBlocks; base += kScanBlock) {
    const int i = base + threadIdx.x;
    const int v = i < numBlocks ? totals[i] : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off)
        x += y;
    }
    if (lane == 63)
      waveTotals[wave] = x;
    __syncthreads();
    if (wave == 0) {
      int t = lane < kScanBlock / 64 ? waveTotals[lane] : 0;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(t, off, 64);
        if (lane >= off)
          t += y;
      }
      if (lane < kScanBlock / 64)
        waveTotals[lane] = t;
    }
    __syncthreads();
    const int before = (wave > 0 ? waveTotals[wave - 1] : 0) + carry;
    if (i < numBlocks)
      totals[i] = before + x - v; // exclusive
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1)
      carry = before + x;
    __syncthreads();
  }
}

__global__ void addOffsetsKernel(int *out, int m, const int *offsets) {
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  if (i < m)
    out[i] += offsets[blockIdx.x];
}

struct Clusters {
  double *c64;
  float *c32;
  int *size;
  int *rep;
};

struct DevMerge {
  int u, v;
  double h;
};

// One thread per (slot, feature): surviving slot i goes to newIndex[i]; if it
// pairs with a higher slot j, the merged centroid goes there and the merge is
// recorded (heights in FP64).
__global__ void rnnMergeKernel(Clusters cur, Clusters next, const unsigned long long *nearest,
                               const int *newIndex, int m, int d, DevMerge *merges,
                               int *numMerges) {
  const size_t total = (size_t)m * d;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / d), k = (int)(e - (size_t)i * d);
    const int j = reciprocalPartner(nearest, i);
    if (j >= 0 && j < i)
      continue;
    const int to = newIndex[i];
    const double ci = cur.c64[e];
    double c = ci;
    if (j >= 0) {
      const double si = cur.size[i], sj = cur.size[j];
      c = fma(cur.c64[(size_t)j * d + k] - ci, sj / (si + sj), ci);
    }
    next.c64[(size_t)to * d + k] = c;
    next.c32[(size_t)to * d + k] = (float)c;
    if (k == 0) {
      next.size[to] = cur.size[i] + (j >= 0 ? cur.size[j] : 0);
      next.rep[to] = cur.rep[i];
      if (j >= 0) {
        const double *a = cur.c64 + (size_t)i * d, *b = cur.c64 + (size_t)j * d;
        double sq = 0.0;
        for (int f = 0; f < d; ++f) {
          const double diff = a[f] - b[f];
          sq = fma(diff, diff, sq);
        }
        const double si = cur.size[i], sj = cur.size[j];
        const int u = cur.rep[i], v = cur.rep[j];
        merges[atomicAdd(numMerges, 1)] =
            DevMerge{min(u, v), max(u, v), sqrt(2.0 * (si * sj) / (si + sj) * sq)};
      }
    }
  }
}

// Every point its own cluster; rep = its original index (orig, null: the
// position), comp = its slot.
__global__ void initClustersKernel(const float *X, int n, int d, Clusters c, int *comp,
                                   const int *orig) {
  const size_t total = (size_t)n * d;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    c.c64[e] = (double)X[e];
    c.c32[e] = X[e];
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    c.size[i] = 1;
    c.rep[i] = orig ? orig[i] : i;
    comp[i] = i;
  }
}

// Condensed Ward distances between cluster centroids, in FP64 from the FP64
// centroids, stored as FP32.
__global__ void wardCondensedKernel(const double *c, const int *size, int m, int d, float *D) {
  const long long total = (long long)m * (m - 1) / 2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    // row i from t: largest i with i*m - i*(i+1)/2 <= t
    const double mm = m;
    long long i = (long long)((2.0 * mm - 1.0 - sqrt((2.0 * mm - 1.0) * (2.0 * mm - 1.0) -
                                                     8.0 * (double)t)) * 0.5);
    if (i < 0)
      i = 0;
    while (i > 0 && i * m - i * (i + 1) / 2 > t)
      --i;
    while ((i + 1) * m - (i + 1) * (i + 2) / 2 <= t)
      ++i;
    const long long j = t - (i * m - i * (i + 1) / 2) + i + 1;
    double sq = 0.0;
    for (int f = 0; f < d; ++f) {
      const double diff = c[i * d + f] - c[j * d + f];
      sq = fma(diff, diff, sq);
    }
    const double si = size[i], sj = size[j];
    D[t] = (float)sqrt(2.0 * (si * sj) / (si + sj) * sq);
  }
}

bool wardMatrixFits(int m) {
  size_t freeBytes = 0, totalBytes = 0;
  if (hipMemGetInfo(&freeBytes, &totalBytes) != hipSuccess)
    return false;
  freeBytes += cachedDeviceBytes();
  return 4.0 * ((double)m * (m - 1) / 2) <= 0.9 * (double)freeBytes;
}

int denseTail(const Clusters &c, int m, int d, int n, std::vector<Edge> &merges) {
  DevArray<float> D;
  Status s;
  s(D.alloc((size_t)m * (m - 1) / 2));
  if (!s.ok())
    return s.code;
  hipLaunchKernelGGL(wardCondensedKernel, dim3(2048), dim3(256), 0, stream(), c.c64, c.size, m,
                     d, D.get());
  s((int)hipGetLastError());
  std::vector<double> tree((size_t)(m - 1) * 4);
  if (s.ok())
    s(denseLinkageOnCondensed(D, m, kWard, c.size, c.rep, n, tree.data()));
  if (!s.ok())
    return s.code;
  // Tree ids -> representative points: a new cluster is represented by the
  // representative of its first child.
  std::vector<int> repOfStep(m - 1);
  auto repOf = [&](double id, int step) {
    const int cl = (id >= 0 && id < n + step) ? (int)id : -1;
    return cl < n ? cl : repOfStep[cl - n];
  };
  for (int step = 0; step < m - 1; ++step) {
    const double *row = &tree[(size_t)step * 4];
    const int u = repOf(row[0], step), v = repOf(row[1], step);
    merges.push_back(Edge{std::min(u, v), std::max(u, v), row[2]});
    repOfStep[step] = u;
  }
  return 0;
}

int chainTail(const Clusters &c, int m, int d, std::vector<Edge> &merges) {
  std::vector<Edge> tail;
  int status = nnChainMerges(c.c64, c.size, m, d, tail);
  std::vector<int> rep(m);
  if (status == 0)
    status = download(rep.data(), c.rep, m * sizeof(int));
  if (status == 0)
    for (const Edge &e : tail) {
      const int u = rep[e.u], v = rep[e.v];
      merges.push_back(Edge{std::min(u, v), std::max(u, v), e.w});
    }
  return status;
}

// ---------------------------------------------------------------------------
// Nearest-neighbour chain (cooperative kernel)

struct ChainState {
  double *c; // m x d, slot s = centroid of the cluster in slot s
  int *size; // 0 = retired
  int *chain;
  int *ctl;  // [0] chain length, [1] current slot (-1 = done), [2] next start, [3] merges
  double *blockDist;
  int *blockIdx_;
  DevMerge *merges;
  int m, d;
  long long maxSteps;
};

struct DistIdx {
  double dist;
  int idx;
};

DEV bool lessDI(double d1, int i1, double d2, int i2) {
  return d1 < d2 || (d1 == d2 && i1 < i2);
}

template <int BS> DEV DistIdx blockMinDI(DistIdx v, DistIdx *red) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    const double od = __shfl_xor(v.dist, m, 64);
    const int oi = __shfl_xor(v.idx, m, 64);
    if (lessDI(od, oi, v.dist, v.idx)) {
      v.dist = od;
      v.idx = oi;
    }
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  DistIdx r = red[0];
  for (int w = 1; w < BS / 64; ++w)
    if (lessDI(red[w].dist, red[w].idx, r.dist, r.idx))
      r = red[w];
  __syncthreads();
  return r;
}

constexpr int kChainBlock = 256;

__global__ void __launch_bounds__(kChainBlock) nnChainKernel(ChainState st) {
  cg::grid_group grid = cg::this_grid();
  extern __shared__ double ca[]; // d doubles: the current slot's centroid
  __shared__ DistIdx red[kChainBlock / 64];
  __shared__ int shA, shPrev, shMerge;
  const int m = st.m, d = st.d;
  const int per = (m + gridDim.x - 1) / gridDim.x;
  const int first = blockIdx.x * per, last = min(m, first + per);

  for (long long step = 0; step < st.maxSteps; ++step) {
    if (threadIdx.x == 0) {
      shA = st.ctl[1];
      const int len = st.ctl[0];
      shPrev = len >= 2 ? st.chain[len - 2] : -1;
    }
    __syncthreads();
    const int a = shA, prev = shPrev;
    if (a < 0)
      return; // uniform: every workgroup read the same value after the barrier
    for (int k = threadIdx.x; k < d; k += kChainBlock)
      ca[k] = st.c[(size_t)a * d + k];
    __syncthreads();
    const double sa = st.size[a];
    DistIdx best{INFINITY, 0x7fffffff};
    for (int s = first + (int)threadIdx.x; s < last; s += kChainBlock) {
      const int ss = st.size[s];
      if (ss == 0 || s == a)
        continue;
      const double *cs = st.c + (size_t)s * d;
      double sq = 0.0;
      for (int f = 0; f < d; ++f) {
        const double diff = ca[f] - cs[f];
        sq = fma(diff, diff, sq);
      }
      const double w = 2.0 * (sa * ss) / (sa + ss) * sq;
      const int idx = s == prev ? -1 : s; // the previous chain element wins ties
      if (lessDI(w, idx, best.dist, best.idx)) {
        best.dist = w;
        best.idx = idx;
      }
    }
    best = blockMinDI<kChainBlock>(best, red);
    if (threadIdx.x == 0) {
      st.blockDist[blockIdx.x] = best.dist;
      st.blockIdx_[blockIdx.x] = best.idx;
    }
    grid.sync();
    if (blockIdx.x == 0) {
      DistIdx all{INFINITY, 0x7fffffff};
      for (int b = threadIdx.x; b < (int)gridDim.x; b += kChainBlock)
        if (lessDI(st.blockDist[b], st.blockIdx_[b], all.dist, all.idx)) {
          all.dist = st.blockDist[b];
          all.idx = st.blockIdx_[b];
        }
      all = blockMinDI<kChainBlock>(all, red);
      if (threadIdx.x == 0) {
        const int nearest = all.idx == -1 ? prev : all.idx;
        if (nearest == prev || nearest == 0x7fffffff) {
          shMerge = nearest == 0x7fffffff ? -2 : prev;
        } else {
          shMerge = -1;
          st.chain[st.ctl[0]++] = nearest;
          st.ctl[1] = nearest;
        }
      }
      __syncthreads();
      if (shMerge == -2) {
        if (threadIdx.x == 0)
          st.ctl[1] = -1; // no candidate (non-finite data): stop; the host reports it
      } else if (shMerge >= 0) {
        const int b = shMerge;
        const int keep = min(a, b), retire = max(a, b);
        const double sb = st.size[b];
        const double wb = sb / (sa + sb);
        for (int k = threadIdx.x; k < d; k += kChainBlock)
          st.c[(size_t)keep * d + k] = fma(st.c[(size_t)b * d + k] - ca[k], wb, ca[k]);
        if (threadIdx.x == 0) {
          st.merges[st.ctl[3]++] = DevMerge{keep, retire, sqrt(all.dist)};
          st.size[keep] = st.size[a] + st.size[b];
          st.size[retire] = 0;
          st.ctl[0] -= 2;
          if (st.ctl[3] == m - 1) {
            st.ctl[1] = -1;
          } else if (st.ctl[0] > 0) {
            st.ctl[1] = st.chain[st.ctl[0] - 1];
          } else {
            int start = st.ctl[2];
            while (start < m && st.size[start] == 0)
              ++start;
            st.ctl[2] = start;
            st.chain[0] = start;
            st.ctl[0] = 1;
            st.ctl[1] = start < m ? start : -1;
          }
        }
      }
    }
    grid.sync();
  }
}

} // namespace

int nnChainMerges(const double *centroids, const int *sizes, int m, int d,
                  std::vector<Edge> &merges) {
  merges.clear();
  if (m < 2)
    return 0;
  int coop = 0;
  (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, 0);
  if (!coop)
    return (int)hipErrorNotSupported;
  const size_t smem = (size_t)d * sizeof(double);
  int perCU = 0;
  Status s;
  s((int)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, nnChainKernel, kChainBlock, smem));
  if (!s.ok())
    return s.code;
  const int blocks = std::max(1, std::min(perCU * numCUs(), (m + kChainBlock - 1) / kChainBlock));
  ChainState st{};
  DevArray<double> c, blockDist;
  DevArray<int> size, chain, ctl, blockIdx;
  DevArray<DevMerge> dm;
  s(c.alloc((size_t)m * d));
  s(size.alloc(m));
  s(chain.alloc(m + 1));
  s(ctl.alloc(4));
  s(blockDist.alloc(blocks));
  s(blockIdx.alloc(blocks));
  s(dm.alloc(m - 1));
  if (!s.ok())
    return s.code;
  s((int)hipMemcpyAsync(c, centroids, (size_t)m * d * sizeof(double), hipMemcpyDeviceToDevice,
                        stream()));
  if (sizes) {
    s((int)hipMemcpyAsync(size, sizes, m * sizeof(int), hipMemcpyDeviceToDevice, stream()));
  } else {
    std::vector<int> ones(m, 1);
    s(upload(size, ones.data(), m * sizeof(int)));
  }
  const int ctl0[4] = {1, 0, 0, 0};
  const int zero = 0;
  s(upload(ctl, ctl0, sizeof(ctl0)));
  s(upload(chain, &zero, sizeof(int)));
  st.c = c;
  st.size = size;
  st.chain = chain;
  st.ctl = ctl;
  st.blockDist = blockDist;
  st.blockIdx_ = blockIdx;
  st.merges = dm;
  st.m = m;
  st.d = d;
  st.maxSteps = 3LL * m + 16; // each step pushes or merges; a chain never exceeds m
  if (s.ok()) {
    void *args[] = {&st};
    s((int)hipLaunchCooperativeKernel((const void *)nnChainKernel, dim3(blocks),
                                      dim3(kChainBlock), args, smem, stream()));
  }
  s(sync());
  int ctlOut[4] = {0, 0, 0, 0};
  s(download(ctlOut, ctl, sizeof(ctlOut)));
  std::vector<DevMerge> out(m - 1);
  s(download(out.data(), dm, (m - 1) * sizeof(DevMerge)));
  if (!s.ok())
    return s.code;
  const int count = std::min(ctlOut[3], m - 1);
  for (int i = 0; i < count; ++i)
    merges.push_back(Edge{out[i].u, out[i].v, out[i].h});
  return 0;
}

int nnChainWard(const float *X, int n, int d, double *tree) {
  if (n < 2)
    return 0;
  DevArray<double> c;
  Status s;
  s(c.alloc((size_t)n * d));
  if (!s.ok())
    return s.code;
  DevArray<float> c32;
  DevArray<int> size, rep, comp;
  s(c32.alloc((size_t)n * d));
  s(size.alloc(n));
  s(rep.alloc(n));
  s(comp.alloc(n));
  Clusters cl{c, c32, size, rep};
  hipLaunchKernelGGL(initClustersKernel, dim3(1024), dim3(256), 0, stream(), X, n, d, cl,
                     comp.get(), nullptr);
  std::vector<Edge> merges;
  if (s.ok())
    s(nnChainMerges(c, nullptr, n, d, merges));
  if (s.ok() && !mergesToTreeOrder(merges, n))
    s(kErrNonFinite);
  if (s.ok())
    treeFromOrderedMerges(merges, n, tree);
  return s.code;
}

int rnnWard(const float *X, int n, int d, const RnnOptions &opt, double *tree) {
  if (n < 2)
    return 0;
  DevArray<double> c64a, c64b;
  DevArray<float> c32a, c32b;
  DevArray<int> sizeA, sizeB, repA, repB, comp, survives, newIndex, blockTotals, numMerges;
  DevArray<unsigned long long> nearest;
  DevArray<DevMerge> merges;
  Status s;
  const size_t nd = (size_t)n * d;
  const int scanBlocks = (n + kScanBlock - 1) / kScanBlock;
  s(c64a.alloc(nd));
  s(c64b.alloc(nd));
  s(c32a.alloc(nd));
  s(c32b.alloc(nd));
  s(sizeA.alloc(n));
  s(sizeB.alloc(n));
  s(repA.alloc(n));
  s(repB.alloc(n));
  s(comp.alloc(n));
  s(survives.alloc(n));
  s(newIndex.alloc(n));
  s(blockTotals.alloc(scanBlocks));
  s(numMerges.alloc(1));
  s(nearest.alloc(n));
  s(merges.alloc(n - 1));
  if (!s.ok())
    return s.code;
  Clusters cur{c64a, c32a, sizeA, repA}, next{c64b, c32b, sizeB, repB};
  hipLaunchKernelGGL(initClustersKernel, dim3(1024), dim3(256), 0, stream(), X, n, d, cur,
                     comp.get(), opt.orig);
  s(fill(numMerges, 0, sizeof(int)));

  std::vector<int> roundEnd;
  int m = n, merged = 0;
  enum { kNone, kDense, kChain } tail = kNone;
  // Pruned passes (points in k-d order): merged clusters take the lower
  // slot's place in the compacted order, so slots stay spatially coherent and
  // the cluster tiles keep tight boxes. Every cluster needs its exact
  // nearest, so no component bound is used; pruning stops for good once a
  // pass computes more than half of the tile pairs.
  bool pruneThisPass = opt.prune;
  while (s.ok() && m > 1) {
    if (opt.allowDense && m <= opt.denseBelow && wardMatrixFits(m)) {
      tail = kDense;
      break;
    }
    s(fill(nearest, 0xff, m * sizeof(unsigned long long)));
    WeightSpec w;
    w.kind = kWeightWard;
    w.size = cur.size;
    if (m == n && opt.mixedFirstPass) {
      // Every cluster is one point (size 1, component = slot): the Ward
      // weight 2 * 1 * 1 / (1 + 1) * sq is sq exactly, and the mixed pass
      // returns every point's exact (FP32 squared distance, slot) nearest.
      int created = 0;
      MixedPass *mp = mixedPassCreate(cur.c32, m, d, &created);
      s(created);
      if (mp) {
        s(mixedPassRun(mp, comp, m, true, true, nearest));
        mixedPassDestroy(mp);
      }
    } else if (pruneThisPass) {
      double share = 1.0;
      s(nearestOtherWalk(cur.c32, m, d, MetricSpec{kSqEuclidean, 2.0f}, w, comp, nullptr,
                         nullptr, nearest, &share, opt.stats));
      pruneThisPass = share <= kPruneGiveUpShare;
    } else {
      s(nearestOtherSymmetric(cur.c32, m, d, MetricSpec{kSqEuclidean, 2.0f}, w, comp, nullptr,
                              nearest));
    }
    {
      PhaseTimer timer(kPhaseRnn);
      const int blocks = (m + kScanBlock - 1) / kScanBlock;
      hipLaunchKernelGGL(survivorsKernel, dim3((m + 255) / 256), dim3(256), 0, stream(),
                         nearest.get(), m, survives.get());
      hipLaunchKernelGGL(scanBlocksKernel, dim3(blocks), dim3(kScanBlock), 0, stream(),
                         survives.get(), m, newIndex.get(), blockTotals.get());
      hipLaunchKernelGGL(scanTotalsKernel, dim3(1), dim3(kScanBlock), 0, stream(),
                         blockTotals.get(), blocks);
      hipLaunchKernelGGL(addOffsetsKernel, dim3(blocks), dim3(kScanBlock), 0, stream(),
                         newIndex.get(), m, blockTotals.get());
      hipLaunchKernelGGL(rnnMergeKernel, dim3(2048), dim3(256), 0, stream(), cur, next,
                         nearest.get(), newIndex.get(), m, d, merges.get(), numMerges.get());
      s((int)hipGetLastError());
    }
    int total = 0;
    s(download(&total, numMerges, sizeof(int)));
    if (!s.ok())
      break;
    const int roundMerges = total - merged;
    merged = total;
    roundEnd.push_back(total);
    std::swap(cur, next);
    const int before = m;
    m -= roundMerges;
    if (m > 1 && (roundMerges == 0 || roundMerges < opt.minProgress * before)) {
      tail = opt.allowDense && wardMatrixFits(m) ? kDense : kChain;
      break;
    }
  }
  std::vector<DevMerge> dm(merged);
  s(download(dm.data(), merges, merged * sizeof(DevMerge)));
  if (!s.ok())
    return s.code;
  std::vector<Edge> all(merged);
  for (int i = 0; i < merged; ++i)
    all[i] = Edge{dm[i].u, dm[i].v, dm[i].h};
  // Within a round merges arrive in atomic order; sort each round so the
  // result does not depend on scheduling. Rounds stay in order (children first).
  for (size_t r = 0; r < roundEnd.size(); ++r)
    std::sort(all.begin() + (r == 0 ? 0 : roundEnd[r - 1]), all.begin() + roundEnd[r], edgeLess);
  if (tail == kDense)
    s(denseTail(cur, m, d, n, all));
  else if (tail == kChain)
    s(chainTail(cur, m, d, all));
  if (s.ok() && !mergesToTreeOrder(all, n))
    s(kErrNonFinite);
  if (s.ok())
    treeFromOrderedMerges(all, n, tree);
  return s.code;
}

} // namespace cuhc
