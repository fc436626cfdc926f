// Dense agglomerative clustering with Lance-Williams updates over a condensed
// distance matrix in HBM (reference src/linkage.cuh: the same linkages, FP64
// updates stored as FP32, lowest-slot ties; a different layout and loop).
//
// Layout: D(i, j), i < j, at i*n - i*(i+1)/2 + (j - i - 1): scipy's pdist order,
// n (n - 1) / 2 floats (20 GB at n = 100k). Slot s holds one active cluster;
// merging slots a < b puts the merged cluster in a and retires b.
//
// Forward nearest cache: fn(i) = argmin over active j > i of (D(i, j), j). The
// closest pair (lowest distance, then lowest i, then lowest j) is
// (i*, fn(i*)) with i* the argmin over slots of (D(i, fn(i)), i). A merge
// changes D(a, k) for every k and retires b, so only rows k < b can need a new
// fn: rows with fn in {a, b} are rescanned unless the merged distance is still
// the row's minimum; row a is always rescanned. Every rescan reads a
// contiguous segment of row k; only the Lance-Williams update itself touches
// the column a (strided).
//
// The whole merge loop is one cooperative persistent kernel (one workgroup per
// CU) with three grid barriers per merge:
//   A  each workgroup takes the minimum over its slots
//   B  all workgroups reduce those minima (redundantly) and know the pair;
//   C  Lance-Williams update of row/column a over all k, fn maintenance, and
//      rescan requests (row chunks) into a work list
//   D  rescans: each chunk's minimum is folded into the row's scratch key
// The tree row is written on the device; there is no host round trip.
#include "linkage.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>

namespace cuhc {

namespace {

namespace cg = cooperative_groups;

constexpr int kChunk = 8192;   // columns per rescan work item
constexpr int kThreads = 256;

// Order-preserving map of a float to 32 bits (all non-NaN values), so that
// (distance, index) fits one u64 key without clamping negative distances.
HD unsigned int orderedBits(float f) {
  const unsigned int b = bitsOf(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HD float fromOrderedBits(unsigned int o) {
  return floatOf((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}
HD unsigned long long denseKey(float f, unsigned int idx) {
  return ((unsigned long long)orderedBits(f) << 32) | idx;
}

HD size_t condensedIndex(long long i, long long j, long long n) {
  // i < j
  return (size_t)(i * n - i * (i + 1) / 2 + (j - i - 1));
}

DEV double lanceWilliams(int linkage, double dak, double dbk, double dab, double na, double nb,
                         double nk) {
  switch (linkage) {
  case kSingle: return fmin(dak, dbk);
  case kComplete: return fmax(dak, dbk);
  case kAverage: return (na * dak + nb * dbk) / (na + nb);
  default: {
    const double total = na + nb + nk;
    const double sq = ((na + nk) * dak * dak + (nb + nk) * dbk * dbk - nk * dab * dab) / total;
    return sqrt(fmax(sq, 0.0));
  }
  }
}

struct DenseState {
  float *D;
  int n;
  int linkage;
  int firstNewId;
  int *size;
  int *id;
  unsigned char *active;
  unsigned long long *fnKey;      // per slot: denseKey(D(i, fn(i)), fn(i)) or empty
  unsigned long long *scratchKey; // per slot: rescan result being accumulated
  unsigned long long *blockBest;  // per workgroup
  int2 *items[2];                 // (row, first column) rescan work, by step parity
  int *itemCount;                 // [2]
  int itemCapacity;
  double *tree; // (n - 1) x 4, device
  unsigned long long *clocks; // null, or [3]: device-clock ticks in selection, update, rescan
};

DEV unsigned long long blockMinKey(unsigned long long v, unsigned long long *red) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    v = minKey(v, shflXorKey(v, m));
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
    red[wave] = v;
  __syncthreads();
  unsigned long long r = red[0];
  for (int w = 1; w < kThreads / 64; ++w)
    r = minKey(r, red[w]);
  return r;
}

// Minimum over active j in [j0, j1) of denseKey(D(row, j), j), by the whole workgroup.
DEV unsigned long long scanRowSegment(const DenseState &s, int row, int j0, int j1,
                                      unsigned long long *red) {
  unsigned long long best = kEmptyKey;
  const size_t base = condensedIndex(row, row + 1, s.n) - (size_t)(row + 1);
  for (int j = j0 + (int)threadIdx.x; j < j1; j += kThreads)
    if (s.active[j])
      best = minKey(best, denseKey(s.D[base + j], (unsigned)j));
  return blockMinKey(best, red);
}

// Enqueues the chunks of row r's forward segment into list p; resets its scratch key.
DEV void enqueueRow(const DenseState &s, int p, int r) {
  const int len = s.n - 1 - r;
  if (len <= 0) {
    s.scratchKey[r] = kEmptyKey;
    return;
  }
  const int chunks = (len + kChunk - 1) / kChunk;
  s.scratchKey[r] = kEmptyKey;
  const int at = atomicAdd(&s.itemCount[p], chunks);
  for (int c = 0; c < chunks && at + c < s.itemCapacity; ++c)
    s.items[p][at + c] = make_int2(r, r + 1 + c * kChunk);
}

__global__ void __launch_bounds__(kThreads) denseInitKernel(DenseState s, int numItems) {
  __shared__ unsigned long long red[kThreads / 64];
  // Work item t: row r, chunk c, enumerated row by row.
  for (int t = blockIdx.x; t < numItems; t += gridDim.x) {
    const int2 it = s.items[0][t];
    const int j1 = min(s.n, it.y + kChunk);
    const unsigned long long k = scanRowSegment(s, it.x, it.y, j1, red);
    if (threadIdx.x == 0 && k != kEmptyKey)
      atomicMin(&s.fnKey[it.x], k);
  }
}

__global__ void __launch_bounds__(kThreads) denseMergeKernel(DenseState s) {
  cg::grid_group grid = cg::this_grid();
  __shared__ unsigned long long red[kThreads / 64];
  __shared__ int shA, shB, shNa, shNb, shIdA, shIdB;
  __shared__ float shDist;
  const int n = s.n;
  const int G = gridDim.x;
  const int per = (n + G - 1) / G;
  const int mine0 = blockIdx.x * per, mine1 = min(n, mine0 + per);
  const int gtid = blockIdx.x * kThreads + threadIdx.x, gthreads = G * kThreads;
  const bool clocked = s.clocks && gtid == 0;
  unsigned long long tSelect = 0, tUpdate = 0, tRescan = 0, t0 = 0, t1 = 0, t2 = 0;

  for (int step = 0; step < n - 1; ++step) {
    const int p = step & 1, q = p ^ 1;
    if (clocked)
      t0 = wall_clock64();
    // A: fold the previous step's rescans into fn, then this workgroup's minimum.
    if (step > 0) {
      const int cnt = min(s.itemCount[q], s.itemCapacity);
      for (int t = threadIdx.x; t < cnt; t += kThreads) {
        const int r = s.items[q][t].x;
        if (r >= mine0 && r < mine1)
          s.fnKey[r] = s.scratchKey[r];
      }
      __syncthreads();
    }
    unsigned long long local = kEmptyKey;
    for (int i = mine0 + (int)threadIdx.x; i < mine1; i += kThreads) {
      const unsigned long long f = s.fnKey[i];
      if (s.active[i] && f != kEmptyKey)
        local = minKey(local, (f & 0xffffffff00000000ull) | (unsigned)i);
    }
    local = blockMinKey(local, red);
    if (threadIdx.x == 0)
      s.blockBest[blockIdx.x] = local;
    grid.sync();

    // B: the pair.
    unsigned long long best = kEmptyKey;
    for (int w = threadIdx.x; w < G; w += kThreads)
      best = minKey(best, s.blockBest[w]);
    best = blockMinKey(best, red);
    if (threadIdx.x == 0) {
      // With two or more active slots the lowest one always has a forward
      // nearest; an empty key is reported (row of -1) and the step is skipped.
      const int a = best == kEmptyKey ? -1 : keyIndex(best);
      const unsigned long long fa = a >= 0 ? s.fnKey[a] : kEmptyKey;
      const int b = fa == kEmptyKey ? -1 : keyIndex(fa);
      shA = b >= 0 ? a : -1;
      shB = b;
      shDist = b >= 0 ? fromOrderedBits((unsigned)(fa >> 32)) : 0.0f;
      shNa = b >= 0 ? s.size[a] : 0;
      shNb = b >= 0 ? s.size[b] : 0;
      shIdA = b >= 0 ? s.id[a] : -1;
      shIdB = b >= 0 ? s.id[b] : -1;
    }
    __syncthreads();
    if (clocked) {
      t1 = wall_clock64();
      tSelect += t1 - t0;
    }
    const int a = shA, b = shB, na = shNa, nb = shNb;
    const float dab = shDist;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      double *row = s.tree + (size_t)step * 4;
      row[0] = (double)min(shIdA, shIdB);
      row[1] = (double)max(shIdA, shIdB);
      row[2] = (double)dab;
      row[3] = (double)(na + nb);
      s.itemCount[q] = 0; // list q was last read in phase A above
    }
    if (a < 0) { // uniform: every workgroup reduced the same minima
      grid.sync();
      grid.sync();
      continue;
    }

    // C: Lance-Williams update of D(a, k) for all active k, fn maintenance.
    for (int k = gtid; k < n; k += gthreads) {
      if (k == a || k == b || !s.active[k])
        continue;
      const size_t iak = k < a ? condensedIndex(k, a, n) : condensedIndex(a, k, n);
      const size_t ibk = k < b ? condensedIndex(k, b, n) : condensedIndex(b, k, n);
      const float dak = s.D[iak], dbk = s.D[ibk];
      const float nd = (float)lanceWilliams(s.linkage, dak, dbk, dab, na, nb, s.size[k]);
      s.D[iak] = nd;
      if (k < a) {
        const unsigned long long f = s.fnKey[k];
        const int fk = keyIndex(f);
        const unsigned long long cand = denseKey(nd, (unsigned)a);
        if (f != kEmptyKey && (fk == a || fk == b)) {
          // The merged cluster replaces the row's nearest if it is no farther;
          // every other entry of the row is unchanged and was >= the old minimum.
          if (orderedBits(nd) <= (unsigned)(f >> 32))
            s.fnKey[k] = cand;
          else
            enqueueRow(s, p, k);
        } else if (cand < f) {
          s.fnKey[k] = cand;
        }
      } else if (k < b) {
        if (keyIndex(s.fnKey[k]) == b && s.fnKey[k] != kEmptyKey)
          enqueueRow(s, p, k); // its nearest retired
      }
    }
    if (gtid == 0) {
      s.active[b] = 0;
      s.fnKey[b] = kEmptyKey;
      enqueueRow(s, p, a);
    }
    grid.sync();
    if (clocked) {
      t2 = wall_clock64();
      tUpdate += t2 - t1;
    }

    // D: rescans; bookkeeping of the merged slot.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      s.size[a] = na + nb;
      s.id[a] = s.firstNewId + step;
    }
    const int cnt = min(s.itemCount[p], s.itemCapacity);
    for (int t = blockIdx.x; t < cnt; t += G) {
      const int2 it = s.items[p][t];
      const unsigned long long k = scanRowSegment(s, it.x, it.y, min(n, it.y + kChunk), red);
      if (threadIdx.x == 0 && k != kEmptyKey)
        atomicMin(&s.scratchKey[it.x], k);
    }
    grid.sync();
    if (clocked)
      tRescan += wall_clock64() - t2;
  }
  if (clocked) {
    s.clocks[0] = tSelect;
    s.clocks[1] = tUpdate;
    s.clocks[2] = tRescan;
  }
}

__global__ void denseSetupKernel(DenseState s, const int *initialSize, const int *initialId) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    s.size[i] = initialSize ? initialSize[i] : 1;
    s.id[i] = initialId ? initialId[i] : i;
    s.active[i] = 1;
    s.fnKey[i] = kEmptyKey;
    s.scratchKey[i] = kEmptyKey;
  }
}

__global__ void denseInitItemsKernel(int2 *items, int n) {
  // Row r's chunks start at sum over rows < r of ceil((n - 1 - r') / kChunk);
  // rows are enumerated by a single thread per row with a closed form below.
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n - 1)
    return;
  // Each row gets ceil((n-1-r)/kChunk) items; prefix over rows computed on
  // the host is avoided by giving every row a fixed stride of maxChunks.
  const int maxChunks = (n - 1 + kChunk - 1) / kChunk;
  const int len = n - 1 - r;
  const int chunks = (len + kChunk - 1) / kChunk;
  for (int c = 0; c < maxChunks; ++c)
    items[(size_t)r * maxChunks + c] = c < chunks ? make_int2(r, r + 1 + c * kChunk)
                                                  : make_int2(r, n); // empty segment
}

} // namespace

int denseLinkageOnCondensed(float *D, int n, int linkage, const int *initialSize,
                            const int *initialId, int firstNewId, double *hostTree) {
  if (n < 2)
    return 0;
  DenseState st{};
  DevArray<int> size, id, itemCount;
  DevArray<unsigned char> active;
  DevArray<unsigned long long> fnKey, scratchKey, blockBest;
  DevArray<int2> items0, items1, initItems;
  DevArray<double> tree;
  DevArray<unsigned long long> clocks;
  Status s;
  const int maxChunks = (n - 1 + kChunk - 1) / kChunk;
  const long long initCount = (long long)(n - 1) * maxChunks;
  // Per step at most one enqueue per row; a row needs at most maxChunks items.
  const long long cap = std::min<long long>((long long)n * maxChunks, 1LL << 30);
  const int G = numCUs();
  s(size.alloc(n));
  s(id.alloc(n));
  s(itemCount.alloc(2));
  s(active.alloc(n));
  s(fnKey.alloc(n));
  s(scratchKey.alloc(n));
  s(blockBest.alloc(G));
  s(items0.alloc(cap));
  s(items1.alloc(cap));
  s(initItems.alloc(initCount));
  s(tree.alloc((size_t)(n - 1) * 4));
  if (!s.ok())
    return s.code;
  st.D = D;
  st.n = n;
  st.linkage = linkage;
  st.firstNewId = firstNewId;
  st.size = size;
  st.id = id;
  st.active = active;
  st.fnKey = fnKey;
  st.scratchKey = scratchKey;
  st.blockBest = blockBest;
  st.items[0] = items0;
  st.items[1] = items1;
  st.itemCount = itemCount;
  st.itemCapacity = (int)cap;
  st.tree = tree;
  st.clocks = nullptr;
  if (timingEnabled()) {
    s(clocks.alloc(3));
    s(fill(clocks, 0, 3 * sizeof(unsigned long long)));
    st.clocks = clocks;
  }
  s(fill(itemCount, 0, 2 * sizeof(int)));
  PhaseTimer timer(kPhaseDense);
  hipLaunchKernelGGL(denseSetupKernel, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0,
                     stream(), st, initialSize, initialId);
  hipLaunchKernelGGL(denseInitItemsKernel, dim3((n + 255) / 256), dim3(256), 0, stream(),
                     initItems.get(), n);
  DenseState initState = st;
  initState.items[0] = initItems;
  hipLaunchKernelGGL(denseInitKernel, dim3((unsigned)std::max<long long>(1, std::min<long long>(initCount, 8LL * G))),
                     dim3(kThreads), 0, stream(), initState, (int)initCount);
  s((int)hipGetLastError());
  if (s.ok()) {
    void *args[] = {&st};
    s((int)hipLaunchCooperativeKernel((const void *)denseMergeKernel, dim3(G), dim3(kThreads),
                                      args, 0, stream()));
  }
  s(sync());
  if (s.ok() && st.clocks) {
    unsigned long long ticks[3];
    s(download(ticks, clocks, sizeof(ticks)));
    const double khz = wallClockKHz();
    for (int i = 0; i < 3; ++i)
      addMs(kPhaseDenseSelect + i, (double)ticks[i] / khz, 1);
  }
  if (s.ok())
    s(download(hostTree, tree, (size_t)(n - 1) * 4 * sizeof(double)));
  if (s.ok())
    for (int step = 0; step < n - 1; ++step)
      if (hostTree[(size_t)step * 4] < 0)
        return kErrNonFinite;
  return s.code;
}

int denseLinkage(const float *X, int n, int d, MetricSpec m, int linkage, double *hostTree) {
  if (n < 2)
    return 0;
  const size_t count = (size_t)n * (n - 1) / 2;
  DevArray<float> D;
  Status s;
  s(D.alloc(count));
  if (!s.ok())
    return s.code;
  s(condensedMatrix(X, n, d, m, D));
  if (s.ok())
    s(denseLinkageOnCondensed(D, n, linkage, nullptr, nullptr, n, hostTree));
  return s.code;
}

} // namespace cuhc
