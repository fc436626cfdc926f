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
// the row's minimum; row a's new fn is the minimum of its new distances.
//
// The merge loop is one cooperative persistent kernel (one workgroup per CU)
// with two grid barriers per merge. Workgroup g owns a contiguous range of
// slots (rows) and is the only writer of their fn, size, id and active flag,
// and of D(k, a) / D(a, k) for its rows k. Merge s:
//   update:  every workgroup reduces the per-workgroup selection records to
//            the pair (a, b); applies the Lance-Williams update to D(a, k) for
//            its own k, maintains its rows' fn, lists its rows that lost their
//            nearest, and publishes its part of row a's new fn;
//   barrier;
//   rescan:  the owner of a reduces row a's parts; all workgroups share the
//            listed rows' rescans in column chunks (atomicMin into fn), and
//            each publishes the minimum (selection key, fn) over its own rows
//            and the chunks it scanned;
//   barrier.
// Per-workgroup records replace same-address atomics, which serialize across
// workgroups. The tree row is written on the device; there is no host round
// trip between merges.

#include "linkage.hpp"

#include <algorithm>

namespace cuhc {

namespace {

constexpr int kThreads = 512;
constexpr int kChunk = 8192;                     // columns per initial-scan work item
constexpr int kRescanChunk = 4096;               // columns per rescan work item
constexpr int kRescanPer = kRescanChunk / kThreads;
constexpr int kMaxOwned = 2048;                  // rows per workgroup (LDS rescan list)
constexpr int kInlineRescans = 5;                // rescan rows carried in an update record
constexpr int kRecordInts = 8;                   // update record: count, rows, padding
constexpr int kGroup = 16;                       // workgroups per first-level barrier counter
constexpr int kLineWords = 32;                   // one 128-byte line in 32-bit words
constexpr unsigned int kSpinLimit = 1u << 22;    // barrier polls (seconds) before aborting

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

// Grid barrier for cooperatively launched (co-resident) workgroups: a
// two-level counter tree, kGroup workgroups per first-level counter and one
// root counter, each on its own line, and a generation word that the last
// arrival advances. No counter sees more than max(kGroup, groups) contending
// atomics. Counters only grow; the last arrival at a level is the one whose
// increment completes a multiple of that level's membership.
// Each wave waits for its own stores; one thread per workgroup then releases
// at agent scope (L2 write-back), arrives, polls the generation with relaxed
// loads (an acquire per poll would invalidate caches on every poll) and
// acquires once. A wait longer than kSpinLimit polls, or an abort word set by
// another workgroup, sets the abort word and returns false, so every
// workgroup leaves the loop.
struct Barrier {
  unsigned int *words; // [kLineWords * (3 + groups)]: generation, abort, groups, root
  int groups;
};

DEV unsigned int fetchAdd(unsigned int *p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

DEV bool gridBarrier(const Barrier &bar, unsigned int &generation, int *ok) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    ++generation;
    unsigned int *gen = bar.words;
    unsigned int *abortWord = bar.words + kLineWords;
    const int group = (int)blockIdx.x / kGroup;
    const unsigned int members = (unsigned int)min(kGroup, (int)gridDim.x - group * kGroup);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int old = fetchAdd(bar.words + kLineWords * (2 + group));
    if ((old + 1) % members == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      const unsigned int r = fetchAdd(bar.words + kLineWords * (2 + bar.groups));
      if ((r + 1) % (unsigned int)bar.groups == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        __hip_atomic_store(gen, generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    int result = 1;
    for (unsigned int polls = 0;; ++polls) {
      if (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == generation)
        break;
      if ((polls & 63) == 63 &&
          (polls >= kSpinLimit ||
           __hip_atomic_load(abortWord, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
        __hip_atomic_store(abortWord, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        result = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok = result;
  }
  __syncthreads();
  return *ok != 0;
}

struct DenseState {
  float *D;
  int n;
  int linkage;
  int firstNewId;
  int *size;
  int *id;
  unsigned char *active;
  unsigned long long *fnKey;       // per slot: denseKey(D(i, fn(i)), fn(i)) or empty
  unsigned long long *selectSlots; // [2 G]: per workgroup (selection key, fn) of its best row
  // Update records, one per workgroup: the minimum over its k > a of
  // (D(a, k), k), and its rows to rescan (count, then up to kInlineRescans
  // rows; all of them in rescanRows[g kMaxOwned ...] when there are more).
  unsigned long long *rowASlots;   // [G]
  int *records;                    // [G kRecordInts]
  int *rescanRows;                 // [G kMaxOwned]
  int2 *items;                     // initial fn: (row, first column) work items
  double *tree;                    // (n - 1) x 4, device
  unsigned long long *clocks;      // null, or [4]: device-clock ticks per part of a merge
  Barrier bar;
};

// Exclusive prefix sum over the workgroup; total receives the sum.
DEV int blockExclusiveScan(int v, int *waveSums, int &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d)
      x += y;
  }
  __syncthreads();
  if (lane == 63)
    waveSums[wave] = x;
  __syncthreads();
  int before = 0, sum = 0;
  for (int w = 0; w < kThreads / 64; ++w) {
    const int sw = waveSums[w];
    before += w < wave ? sw : 0;
    sum += sw;
  }
  total = sum;
  return before + x - v;
}

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

// Lexicographic minimum of (key, payload) pairs.
DEV void pairMin(unsigned long long &k, unsigned long long &p, unsigned long long k2,
                 unsigned long long p2) {
  if (k2 < k || (k2 == k && p2 < p)) {
    k = k2;
    p = p2;
  }
}

DEV void blockMinPair(unsigned long long &k, unsigned long long &p, unsigned long long *redK,
                      unsigned long long *redP) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    pairMin(k, p, shflXorKey(k, m), shflXorKey(p, m));
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    redK[wave] = k;
    redP[wave] = p;
  }
  __syncthreads();
  k = redK[0];
  p = redP[0];
  for (int w = 1; w < kThreads / 64; ++w)
    pairMin(k, p, redK[w], redP[w]);
}

// Minimum over active j in [j0, j1) of denseKey(D(row, j), j), by the whole
// workgroup.
DEV unsigned long long scanRowSegment(const DenseState &s, int row, int j0, int j1,
                                      unsigned long long *red) {
  unsigned long long best = kEmptyKey;
  const size_t base = condensedIndex(row, row + 1, s.n) - (size_t)(row + 1);
  for (int j = j0 + (int)threadIdx.x; j < j1; j += kThreads)
    if (s.active[j])
      best = minKey(best, denseKey(s.D[base + j], (unsigned)j));
  return blockMinKey(best, red);
}

// The same over one rescan chunk (at most kRescanChunk columns), with every
// load issued before the first compare.
DEV unsigned long long scanRescanChunk(const DenseState &s, int row, int j0, int j1,
                                       unsigned long long *red) {
  const size_t base = condensedIndex(row, row + 1, s.n) - (size_t)(row + 1);
  float v[kRescanPer];
  unsigned char act[kRescanPer];
#pragma unroll
  for (int u = 0; u < kRescanPer; ++u) {
    const int j = j0 + (int)threadIdx.x + u * kThreads;
    act[u] = j < j1 ? s.active[j] : 0;
    v[u] = j < j1 ? s.D[base + j] : 0.0f;
  }
  unsigned long long best = kEmptyKey;
#pragma unroll
  for (int u = 0; u < kRescanPer; ++u)
    if (act[u])
      best = minKey(best, denseKey(v[u], (unsigned)(j0 + (int)threadIdx.x + u * kThreads)));
  return blockMinKey(best, red);
}

__global__ void __launch_bounds__(kThreads) denseInitKernel(DenseState s, int numItems) {
  __shared__ unsigned long long red[kThreads / 64];
  // Work item t: row r, chunk c, enumerated row by row.
  for (int t = blockIdx.x; t < numItems; t += gridDim.x) {
    const int2 it = s.items[t];
    const int j1 = min(s.n, it.y + kChunk);
    const unsigned long long k = scanRowSegment(s, it.x, it.y, j1, red);
    if (threadIdx.x == 0 && k != kEmptyKey)
      atomicMin(&s.fnKey[it.x], k);
  }
}

// Selection key of a row: (distance of its forward nearest, row).
DEV unsigned long long selectionKey(unsigned long long fn, int row) {
  return fn == kEmptyKey ? kEmptyKey : (fn & 0xffffffff00000000ull) | (unsigned int)row;
}

__global__ void __launch_bounds__(kThreads) denseMergeKernel(DenseState s) {
  __shared__ unsigned long long redK[kThreads / 64], redP[kThreads / 64];
  __shared__ int ownRescan[kMaxOwned];
  __shared__ int listStart[kThreads + 1];
  __shared__ int listRows[kInlineRescans * kThreads];
  __shared__ int waveSums[kThreads / 64];
  __shared__ int ownRescanCount, barrierOk;
  const int n = s.n;
  const int G = gridDim.x, g = blockIdx.x;
  const int tid = threadIdx.x;
  const int per = (n + G - 1) / G;
  const int mine0 = min(n, g * per), mine1 = min(n, mine0 + per);
  const int maxChunks = (n - 1 + kRescanChunk - 1) / kRescanChunk;
  unsigned int generation = 0;
  const bool clocked = s.clocks && g == 0 && tid == 0;
  unsigned long long ticks[4] = {0, 0, 0, 0}, t0 = 0;
  auto lap = [&](int part) {
    if (clocked) {
      const unsigned long long t = wall_clock64();
      ticks[part] += t - t0;
      t0 = t;
    }
  };

  // Selection records of merge 0 from the initial fn.
  {
    unsigned long long k = kEmptyKey, p = kEmptyKey;
    for (int i = mine0 + tid; i < mine1; i += kThreads) {
      const unsigned long long f = s.fnKey[i];
      pairMin(k, p, selectionKey(f, i), f);
    }
    blockMinPair(k, p, redK, redP);
    if (tid == 0) {
      s.selectSlots[2 * g] = k;
      s.selectSlots[2 * g + 1] = p;
    }
  }
  if (!gridBarrier(s.bar, generation, &barrierOk))
    return;
  if (clocked)
    t0 = wall_clock64();

  for (int step = 0; step < n - 1; ++step) {
    // The pair: the lowest (selection key, fn) record.
    if (tid == 0)
      ownRescanCount = 0;
    unsigned long long best = kEmptyKey, fa = kEmptyKey;
    for (int t = tid; t < G; t += kThreads)
      pairMin(best, fa, s.selectSlots[2 * t], s.selectSlots[2 * t + 1]);
    blockMinPair(best, fa, redK, redP);
    const int a = best == kEmptyKey ? -1 : keyIndex(best);
    const int b = fa == kEmptyKey ? -1 : keyIndex(fa);
    if (a < 0 || b <= a || b >= n) {
      // No pair (non-finite distances): the remaining rows report -1. Every
      // workgroup read the same records, so all of them stop here.
      if (g == 0)
        for (int r = step + tid; r < n - 1; r += kThreads)
          for (int c = 0; c < 4; ++c)
            s.tree[(size_t)r * 4 + c] = -1.0;
      break;
    }
    const float dab = fromOrderedBits((unsigned)(fa >> 32));
    const int na = s.size[a], nb = s.size[b];
    if (g == 0 && tid == 0) {
      const int ia = s.id[a], ib = s.id[b];
      double *row = s.tree + (size_t)step * 4;
      row[0] = (double)min(ia, ib);
      row[1] = (double)max(ia, ib);
      row[2] = (double)dab;
      row[3] = (double)(na + nb);
    }
    lap(0);

    // Lance-Williams update of D(a, k) for this workgroup's k, its rows' fn,
    // and its part of row a's new fn.
    unsigned long long rowAMin = kEmptyKey, selK = kEmptyKey, selP = kEmptyKey;
    for (int k = mine0 + tid; k < mine1; k += kThreads) {
      if (k == a || k == b)
        continue;
      // Every load is issued before the activity test: one round trip.
      const size_t iak = k < a ? condensedIndex(k, a, n) : condensedIndex(a, k, n);
      const size_t ibk = k < b ? condensedIndex(k, b, n) : condensedIndex(b, k, n);
      const unsigned char act = s.active[k];
      const float dak = s.D[iak], dbk = s.D[ibk];
      const unsigned long long f = s.fnKey[k];
      const int nk = s.size[k];
      if (!act)
        continue;
      const float nd = (float)lanceWilliams(s.linkage, dak, dbk, dab, na, nb, nk);
      s.D[iak] = nd;
      unsigned long long nf = f;
      bool rescan = false;
      if (k < a) {
        // Every entry of row k but D(k, a) and D(k, b) is unchanged.
        const int fk = f == kEmptyKey ? -1 : keyIndex(f);
        const unsigned long long cand = denseKey(nd, (unsigned)a);
        if (fk == a || fk == b) {
          // The other entries were >= the old minimum, and any equal one has
          // a higher column than b: (nd, a) is the minimum if nd is no larger.
          if (orderedBits(nd) <= (unsigned)(f >> 32))
            nf = cand;
          else
            rescan = true;
        } else if (cand < f) {
          nf = cand;
        }
      } else {
        rowAMin = minKey(rowAMin, denseKey(nd, (unsigned)k));
        if (k < b && f != kEmptyKey && keyIndex(f) == b)
          rescan = true; // its nearest retired
      }
      if (rescan) {
        nf = kEmptyKey; // rebuilt by atomicMin over the chunks
        ownRescan[atomicAdd(&ownRescanCount, 1)] = k;
      } else {
        pairMin(selK, selP, selectionKey(nf, k), nf);
      }
      if (nf != f)
        s.fnKey[k] = nf;
    }
    rowAMin = blockMinKey(rowAMin, redK); // also orders the ownRescan appends
    const int own = ownRescanCount;
    int *record = s.records + (size_t)g * kRecordInts;
    if (tid == 0) {
      s.rowASlots[g] = rowAMin;
      record[0] = own;
    }
    if (tid < min(own, kInlineRescans))
      record[1 + tid] = ownRescan[tid];
    if (own > kInlineRescans)
      for (int i = tid; i < own; i += kThreads)
        s.rescanRows[(size_t)g * kMaxOwned + i] = ownRescan[i];
    if (tid == 0 && b >= mine0 && b < mine1) {
      s.active[b] = 0;
      s.fnKey[b] = kEmptyKey;
    }
    lap(1);
    if (!gridBarrier(s.bar, generation, &barrierOk))
      return;
    lap(2);

    // Every workgroup's update record (G <= kThreads: one per thread).
    unsigned long long rowAPart = kEmptyKey;
    int cnt = 0;
    int inlineRows[kInlineRescans];
    if (tid < G) {
      rowAPart = s.rowASlots[tid];
      const int4 r0 = *reinterpret_cast<const int4 *>(s.records + (size_t)tid * kRecordInts);
      const int4 r1 = *reinterpret_cast<const int4 *>(s.records + (size_t)tid * kRecordInts + 4);
      cnt = r0.x;
      inlineRows[0] = r0.y;
      inlineRows[1] = r0.z;
      inlineRows[2] = r0.w;
      inlineRows[3] = r1.x;
      inlineRows[4] = r1.y;
    }
    // Row a's new fn, from every workgroup's part.
    if (a >= mine0 && a < mine1) {
      const unsigned long long m = blockMinKey(rowAPart, redK);
      if (tid == 0) {
        s.fnKey[a] = m;
        s.size[a] = na + nb;
        s.id[a] = s.firstNewId + step;
        pairMin(selK, selP, selectionKey(m, a), m);
      }
    }
    // The rescan list: workgroup w's rows at [listStart[w], listStart[w + 1]).
    int count = 0;
    const int start = blockExclusiveScan(cnt, waveSums, count);
    const bool spilled = __syncthreads_or(cnt > kInlineRescans);
    if (tid < G) {
      listStart[tid] = start;
      if (!spilled)
        for (int i = 0; i < cnt; ++i)
          listRows[start + i] = inlineRows[i];
    }
    if (tid == 0)
      listStart[G] = count;
    __syncthreads();
    // Rescans: item t is chunk t % maxChunks of list row t / maxChunks.
    const long long numItems = (long long)count * maxChunks;
    for (long long t = g; t < numItems; t += G) {
      const int q = (int)(t / maxChunks);
      int r;
      if (!spilled) {
        r = listRows[q];
      } else {
        int lo = 0, hi = G; // the w with listStart[w] <= q < listStart[w + 1]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (listStart[mid] <= q)
            lo = mid;
          else
            hi = mid;
        }
        const int i = q - listStart[lo];
        r = listStart[lo + 1] - listStart[lo] <= kInlineRescans
                ? s.records[(size_t)lo * kRecordInts + 1 + i]
                : s.rescanRows[(size_t)lo * kMaxOwned + i];
      }
      const int j0 = r + 1 + (int)(t % maxChunks) * kRescanChunk;
      if (j0 >= n)
        continue;
      const unsigned long long m = scanRescanChunk(s, r, j0, min(n, j0 + kRescanChunk), redK);
      if (tid == 0 && m != kEmptyKey) {
        atomicMin(&s.fnKey[r], m);
        pairMin(selK, selP, selectionKey(m, r), m);
      }
    }
    blockMinPair(selK, selP, redK, redP);
    if (tid == 0) {
      s.selectSlots[2 * g] = selK;
      s.selectSlots[2 * g + 1] = selP;
    }
    lap(3);
    if (!gridBarrier(s.bar, generation, &barrierOk))
      return;
    lap(2);
  }
  if (clocked)
    for (int i = 0; i < 4; ++i)
      s.clocks[i] = ticks[i];
}

__global__ void denseSetupKernel(DenseState s, const int *initialSize, const int *initialId) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    s.size[i] = initialSize ? initialSize[i] : 1;
    s.id[i] = initialId ? initialId[i] : i;
    s.active[i] = 1;
    s.fnKey[i] = kEmptyKey;
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
  const int G = numCUs();
  if ((n + G - 1) / G > kMaxOwned || G > kThreads)
    return kErrInvalidShape; // beyond any condensed matrix that fits in HBM
  DenseState st{};
  DevArray<int> size, id, records, rescanRows;
  DevArray<unsigned char> active;
  DevArray<unsigned long long> fnKey, selectSlots, rowASlots;
  DevArray<int2> initItems;
  DevArray<double> tree;
  DevArray<unsigned long long> clocks;
  DevArray<unsigned int> barrierWords;
  Status s;
  const int maxChunks = (n - 1 + kChunk - 1) / kChunk;
  const long long initCount = (long long)(n - 1) * maxChunks;
  const int groups = (G + kGroup - 1) / kGroup;
  const size_t numBarrierWords = (size_t)kLineWords * (3 + groups);
  s(size.alloc(n));
  s(id.alloc(n));
  s(active.alloc(n));
  s(fnKey.alloc(n));
  s(selectSlots.alloc(2 * (size_t)G));
  s(rowASlots.alloc(G));
  s(records.alloc((size_t)G * kRecordInts));
  s(rescanRows.alloc((size_t)G * kMaxOwned));
  s(initItems.alloc(initCount));
  s(tree.alloc((size_t)(n - 1) * 4));
  s(barrierWords.alloc(numBarrierWords));
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
  st.selectSlots = selectSlots;
  st.rowASlots = rowASlots;
  st.records = records;
  st.rescanRows = rescanRows;
  st.items = initItems;
  st.tree = tree;
  st.clocks = nullptr;
  st.bar.words = barrierWords;
  st.bar.groups = groups;
  if (timingEnabled()) {
    s(clocks.alloc(4));
    s(fill(clocks, 0, 4 * sizeof(unsigned long long)));
    st.clocks = clocks;
  }
  s(fill(barrierWords, 0, numBarrierWords * sizeof(unsigned int)));
  PhaseTimer timer(kPhaseDense);
  hipLaunchKernelGGL(denseSetupKernel, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0,
                     stream(), st, initialSize, initialId);
  hipLaunchKernelGGL(denseInitItemsKernel, dim3((n + 255) / 256), dim3(256), 0, stream(),
                     initItems.get(), n);
  hipLaunchKernelGGL(denseInitKernel,
                     dim3((unsigned)std::max<long long>(1, std::min<long long>(initCount, 8LL * G))),
                     dim3(kThreads), 0, stream(), st, (int)initCount);
  s((int)hipGetLastError());
  if (s.ok()) {
    void *args[] = {&st};
    s((int)hipLaunchCooperativeKernel((const void *)denseMergeKernel, dim3(G), dim3(kThreads),
                                      args, 0, stream()));
  }
  s(sync());
  unsigned int aborted = 0;
  if (s.ok())
    s(download(&aborted, barrierWords.get() + kLineWords, sizeof(aborted)));
  if (s.ok() && aborted)
    return (int)hipErrorLaunchTimeOut; // a barrier wait exceeded its limit
  if (s.ok() && st.clocks) {
    unsigned long long ticks[4];
    s(download(ticks, clocks, sizeof(ticks)));
    const double khz = wallClockKHz();
    addMs(kPhaseDenseSelect, (double)ticks[0] / khz, 1);
    addMs(kPhaseDenseUpdate, (double)ticks[1] / khz, 1);
    addMs(kPhaseDenseBarrier, (double)ticks[2] / khz, 1);
    addMs(kPhaseDenseRescan, (double)ticks[3] / khz, 1);
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
