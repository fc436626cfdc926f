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
// with one grid barrier per merge. Workgroup g owns a contiguous range of
// slots (rows): their fn and size live in its LDS, it alone writes D(k, a) /
// D(a, k) for its rows k, and every workgroup keeps the active flags of all
// slots as a bitmask in LDS. Phase s (between two barriers):
//   select:  every workgroup reads all records of phase s - 1 and reduces them
//            the same way: the previous merged row's new fn (the minimum of
//            the workgroups' parts) and the pair (a, b) = the lowest
//            (selection key, fn) over every row, that one included; b's bit
//            is cleared; the owner of the previous merged slot installs its
//            fn, size and id (sizes and ids change one phase late, so no read
//            in the merge's own phase races the write);
//   update:  Lance-Williams update of D(a, k) for its own k, its rows' fn,
//            and its part of row a's new fn (minimum over its k > a);
//   rescan:  its rows whose nearest was a or b and is not the merged distance
//            are scanned again by the whole workgroup (no other workgroup
//            changes an entry of such a row in this phase, and D(k, a) is the
//            value it just stored);
//   record:  the minimum (selection key, fn) over its rows but a, and its part
//            of row a's fn;
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
constexpr int kMaxOwned = 2048;                  // rows per workgroup (LDS row state)
constexpr int kUpdatePer = 2;                    // owned rows per thread per load batch
constexpr int kRecordWords = 8;                  // u64 words per workgroup record
constexpr int kInlineRows = 6;                   // pending rows carried in a record
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
  // Cluster size and scipy id by slot. The owner of a merged slot writes its
  // new size and id one phase after the merge (every workgroup still reads
  // the old size in the merge's own phase, and knows the new one itself).
  int *size;
  int *id;
  unsigned long long *fnKey;    // initial forward nearest per slot (denseInitKernel)
  unsigned long long *fnRescan; // per slot: the rescan's minimum (atomicMin) while pending
  // One record per workgroup and publish (kRecordWords words): [0] and [1]
  // the minimum (selection key, fn) over its settled rows, [2] its part of
  // the merged row's new fn, [3] the lowest lower-bound key of its pending
  // rows, [4] the number of rows it made pending in the phase, [5..7] the
  // first kInlineRows of them (two per word); all of them in pendingRows[g
  // kMaxOwned ...] when there are more. Two buffers, alternating by publish:
  // consecutive publishes are separated by a grid barrier, so a buffer is
  // rewritten only after every workgroup has passed the barrier that follows
  // its last reader's next publish.
  unsigned long long *records; // [2][G kRecordWords]
  int *pendingRows;            // [2][G kMaxOwned]
  int2 *items;                 // initial fn: (row, first column) work items
  double *tree;                // (n - 1) x 4, device
  unsigned long long *clocks;  // null, or [4]: device-clock ticks per part of a merge
  Barrier bar;
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

// Workgroup reduction of (k, p) by lexicographic minimum and of x, y and z
// by minimum, in one LDS exchange (red: 5 kThreads / 64 words).
DEV void blockReduce5(unsigned long long &k, unsigned long long &p, unsigned long long &x,
                      unsigned long long &y, unsigned long long &z, unsigned long long *red) {
  constexpr int W = kThreads / 64;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    pairMin(k, p, shflXorKey(k, m), shflXorKey(p, m));
    x = minKey(x, shflXorKey(x, m));
    y = minKey(y, shflXorKey(y, m));
    z = minKey(z, shflXorKey(z, m));
  }
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[wave] = k;
    red[W + wave] = p;
    red[2 * W + wave] = x;
    red[3 * W + wave] = y;
    red[4 * W + wave] = z;
  }
  __syncthreads();
  k = red[0];
  p = red[W];
  x = red[2 * W];
  y = red[3 * W];
  z = red[4 * W];
  for (int w = 1; w < W; ++w) {
    pairMin(k, p, red[w], red[W + w]);
    x = minKey(x, red[2 * W + w]);
    y = minKey(y, red[3 * W + w]);
    z = minKey(z, red[4 * W + w]);
  }
}

// A blockReduce5 without z plus, in the same exchange, the exclusive prefix sum of cnt
// (start; total receives the sum) and whether any cnt exceeds kInlineRows.
DEV void blockReduce4Scan(unsigned long long &k, unsigned long long &p, unsigned long long &x,
                          unsigned long long &y, int cnt, int &start, int &total,
                          bool &spilled, unsigned long long *red, int *redInt) {
  constexpr int W = kThreads / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    pairMin(k, p, shflXorKey(k, m), shflXorKey(p, m));
    x = minKey(x, shflXorKey(x, m));
    y = minKey(y, shflXorKey(y, m));
  }
  int incl = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if (lane >= d)
      incl += v;
  }
  const bool spill = __ballot(cnt > kInlineRows) != 0;
  __syncthreads();
  if (lane == 63)
    redInt[wave] = incl;
  if (lane == 0) {
    red[wave] = k;
    red[W + wave] = p;
    red[2 * W + wave] = x;
    red[3 * W + wave] = y;
    redInt[W + wave] = spill;
  }
  __syncthreads();
  k = red[0];
  p = red[W];
  x = red[2 * W];
  y = red[3 * W];
  int before = 0, sum = 0;
  bool any = false;
  for (int w = 0; w < W; ++w) {
    if (w > 0) {
      pairMin(k, p, red[w], red[W + w]);
      x = minKey(x, red[2 * W + w]);
      y = minKey(y, red[3 * W + w]);
    }
    before += w < wave ? redInt[w] : 0;
    sum += redInt[w];
    any |= redInt[W + w] != 0;
  }
  start = before + incl - cnt;
  total = sum;
  spilled = any;
}

// Minimum over j in [j0, j1) of denseKey(D(row, j), j), by the whole
// workgroup (every slot is active before the first merge).
DEV unsigned long long scanRowSegment(const DenseState &s, int row, int j0, int j1,
                                      unsigned long long *red) {
  unsigned long long best = kEmptyKey;
  const size_t base = condensedIndex(row, row + 1, s.n) - (size_t)(row + 1);
  for (int j = j0 + (int)threadIdx.x; j < j1; j += kThreads)
    best = minKey(best, denseKey(s.D[base + j], (unsigned)j));
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

DEV bool isActive(const unsigned int *activeBits, int j) {
  return (activeBits[j >> 5] >> (j & 31)) & 1u;
}

// Minimum over active j in [j0, j1) (at most kRescanChunk columns), j != skip,
// of denseKey(D(row, j), j), by the whole workgroup, every load issued
// before the first compare.
DEV unsigned long long scanRescanChunk(const DenseState &s, const unsigned int *activeBits,
                                       int row, int j0, int j1, int skip,
                                       unsigned long long *red) {
  const size_t base = condensedIndex(row, row + 1, s.n) - (size_t)(row + 1);
  float v[kRescanPer];
#pragma unroll
  for (int u = 0; u < kRescanPer; ++u) {
    const int j = j0 + (int)threadIdx.x + u * kThreads;
    v[u] = j < j1 ? s.D[base + j] : 0.0f;
  }
  unsigned long long best = kEmptyKey;
#pragma unroll
  for (int u = 0; u < kRescanPer; ++u) {
    const int j = j0 + (int)threadIdx.x + u * kThreads;
    if (j < j1 && j != skip && isActive(activeBits, j))
      best = minKey(best, denseKey(v[u], (unsigned)j));
  }
  return blockMinKey(best, red);
}

__global__ void __launch_bounds__(kThreads) denseMergeKernel(DenseState s) {
  extern __shared__ unsigned int activeBits[]; // (n + 31) / 32 words: bit j = slot j active
  __shared__ unsigned long long ownFn[kMaxOwned];
  __shared__ int ownSize[kMaxOwned];
  __shared__ unsigned char ownPending[kMaxOwned];
  // This workgroup's pending rows of the two latest phases and their
  // lower-bound keys: list[older] is being (or was) rescanned, list[older ^ 1]
  // is filled by the current update.
  __shared__ int pendRow[2][kMaxOwned];
  __shared__ unsigned long long pendKey[2][kMaxOwned];
  __shared__ int pendCount[2];
  // The rows every workgroup made pending in the previous phase:
  // workgroup w's at [listStart[w], listStart[w + 1]).
  __shared__ int listStart[kThreads + 1];
  __shared__ int listRows[kInlineRows * kThreads];
  __shared__ unsigned long long redK[kThreads / 64], redP[kThreads / 64];
  __shared__ unsigned long long red4[5 * (kThreads / 64)];
  __shared__ int redInt[2 * (kThreads / 64)];
  __shared__ int barrierOk;
  const int n = s.n;
  const int G = gridDim.x, g = blockIdx.x;
  const int tid = threadIdx.x;
  const int per = (n + G - 1) / G;
  const int mine0 = min(n, g * per), mine1 = min(n, mine0 + per);
  const int words = (n + 31) >> 5;
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
  int published = 0; // publishes so far: the latest is in buffer (published - 1) & 1
  auto recordsOf = [&](int buffer) { return s.records + (size_t)buffer * G * kRecordWords; };
  auto pendingOf = [&](int buffer) { return s.pendingRows + (size_t)buffer * G * kMaxOwned; };
  // Publishes this workgroup's record: settled rows, row-a part, pending
  // lower bound, and the rows made pending in this phase (list[older ^ 1]);
  // folds this thread's part of a rescan chunk of row chunkRow (-1: none) into
  // that row's result in the same reduction.
  auto publish = [&](unsigned long long selK, unsigned long long selP,
                     unsigned long long rowA, int older, unsigned long long chunkMin,
                     int chunkRow) {
    const int buffer = published & 1;
    unsigned long long *record = recordsOf(buffer) + (size_t)g * kRecordWords;
    unsigned long long pmin = kEmptyKey;
    for (int l = 0; l < 2; ++l)
      for (int i = tid; i < pendCount[l]; i += kThreads)
        if (isActive(activeBits, pendRow[l][i]))
          pmin = minKey(pmin, pendKey[l][i]);
    blockReduce5(selK, selP, rowA, pmin, chunkMin, red4);
    if (tid == 0 && chunkRow >= 0 && chunkMin != kEmptyKey)
      atomicMin(&s.fnRescan[chunkRow], chunkMin);
    const int fresh = older ^ 1, count = pendCount[fresh];
    if (count > kInlineRows)
      for (int i = tid; i < count; i += kThreads)
        pendingOf(buffer)[(size_t)g * kMaxOwned + i] = pendRow[fresh][i];
    if (tid == 0) {
      record[0] = selK;
      record[1] = selP;
      record[2] = rowA;
      record[3] = pmin;
      record[4] = (unsigned long long)count;
      for (int w = 0; w < kInlineRows / 2; ++w) {
        const int r0 = 2 * w < count ? pendRow[fresh][2 * w] : 0;
        const int r1 = 2 * w + 1 < count ? pendRow[fresh][2 * w + 1] : 0;
        record[5 + w] = (unsigned long long)(unsigned)r0 | ((unsigned long long)(unsigned)r1 << 32);
      }
    }
    ++published;
  };
  // The q-th row of the previous phase's pending list (listStart ready).
  auto listRow = [&](int q, bool spilled) {
    if (!spilled)
      return listRows[q];
    int lo = 0, hi = G; // the w with listStart[w] <= q < listStart[w + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (listStart[mid] <= q)
        lo = mid;
      else
        hi = mid;
    }
    const int i = q - listStart[lo];
    const int buffer = (published - 1) & 1;
    if (listStart[lo + 1] - listStart[lo] > kInlineRows)
      return pendingOf(buffer)[(size_t)lo * kMaxOwned + i];
    const unsigned long long w = recordsOf(buffer)[(size_t)lo * kRecordWords + 5 + i / 2];
    return (int)(unsigned)(i & 1 ? w >> 32 : w);
  };
  // Rescans of the listed rows, shared by all workgroups in column chunks:
  // item t is chunk t % maxChunks of list row t / maxChunks. Column skip
  // (the merged slot, whose entries their owners are rewriting) is left out.
  auto rescans = [&](int count, bool spilled, int skip, long long first) {
    const long long numItems = (long long)count * maxChunks;
    for (long long t = first; t < numItems; t += G) {
      const int r = listRow((int)(t / maxChunks), spilled);
      const int j0 = r + 1 + (int)(t % maxChunks) * kRescanChunk;
      if (j0 >= n || !isActive(activeBits, r))
        continue;
      const unsigned long long m =
          scanRescanChunk(s, activeBits, r, j0, min(n, j0 + kRescanChunk), skip, redK);
      if (tid == 0 && m != kEmptyKey)
        atomicMin(&s.fnRescan[r], m);
    }
  };
  // Settles this workgroup's rows of list[l]: their fn is the rescan's result.
  auto install = [&](int l) {
    for (int i = tid; i < pendCount[l]; i += kThreads) {
      const int r = pendRow[l][i];
      ownPending[r - mine0] = 0;
      ownFn[r - mine0] = isActive(activeBits, r) ? s.fnRescan[r] : kEmptyKey;
    }
    __syncthreads();
    if (tid == 0)
      pendCount[l] = 0;
  };
  // The settled rows' minimum (selection key, fn), all rows but skipRow.
  auto settledMin = [&](unsigned long long &k, unsigned long long &p, int skipRow) {
    for (int i = mine0 + tid; i < mine1; i += kThreads)
      if (i != skipRow && !ownPending[i - mine0] && isActive(activeBits, i))
        pairMin(k, p, selectionKey(ownFn[i - mine0], i), ownFn[i - mine0]);
  };

  // Every slot starts active and settled; this workgroup's rows' state moves
  // to LDS, and the records of merge 0 come from the initial fn.
  for (int w = tid; w < words; w += kThreads) {
    const int rest = n - 32 * w;
    activeBits[w] = rest >= 32 ? 0xffffffffu : (1u << rest) - 1u;
  }
  for (int i = mine0 + tid; i < mine1; i += kThreads) {
    ownFn[i - mine0] = s.fnKey[i];
    ownSize[i - mine0] = s.size[i];
    ownPending[i - mine0] = 0;
  }
  if (tid < 2)
    pendCount[tid] = 0;
  __syncthreads();
  int older = 0;
  {
    unsigned long long k = kEmptyKey, p = kEmptyKey;
    settledMin(k, p, -1);
    publish(k, p, kEmptyKey, older, kEmptyKey, -1);
  }
  if (!gridBarrier(s.bar, generation, &barrierOk))
    return;
  if (clocked)
    t0 = wall_clock64();

  int prevA = -1, prevSize = 0; // the previous merge's slot a and merged size
  for (int step = 0; step < n - 1; ++step) {
    // This phase's inputs are loaded together: the latest records, and the
    // rescan results of this workgroup's rows rescanned in the previous phase.
    unsigned long long k = kEmptyKey, p = kEmptyKey, part = kEmptyKey, pmin = kEmptyKey;
    int cnt = 0;
    int inl[kInlineRows] = {};
    if (tid < G) {
      const unsigned long long *r = recordsOf((published - 1) & 1) + (size_t)tid * kRecordWords;
      unsigned long long w[kRecordWords];
#pragma unroll
      for (int i = 0; i < kRecordWords; ++i)
        w[i] = r[i];
      k = w[0];
      p = w[1];
      part = w[2];
      pmin = w[3];
      cnt = (int)w[4];
#pragma unroll
      for (int i = 0; i < kInlineRows / 2; ++i) {
        inl[2 * i] = (int)(unsigned)w[5 + i];
        inl[2 * i + 1] = (int)(unsigned)(w[5 + i] >> 32);
      }
    }
    const int numInstall = pendCount[older];
    int installRow = -1;
    unsigned long long installFn = kEmptyKey;
    if (tid < numInstall) {
      installRow = pendRow[older][tid];
      installFn = s.fnRescan[installRow];
    }
    // Every workgroup reduces the records the same way. The previous merged
    // row's new fn is the minimum of the parts.
    int listCount = 0, start = 0;
    bool spilled = false;
    blockReduce4Scan(k, p, part, pmin, cnt, start, listCount, spilled, red4, redInt);
    if (prevA >= 0)
      pairMin(k, p, selectionKey(part, prevA), part);
    if (tid < G) {
      listStart[tid] = start;
      if (!spilled)
        for (int i = 0; i < cnt; ++i)
          listRows[start + i] = inl[i];
    }
    if (tid == 0) {
      listStart[G] = listCount;
      pendCount[older] = 0;
      if (prevA >= mine0 && prevA < mine1) {
        ownFn[prevA - mine0] = part;
        ownSize[prevA - mine0] = prevSize;
        s.size[prevA] = prevSize;
        s.id[prevA] = s.firstNewId + step - 1;
      }
    }
    // Rows rescanned in the previous phase settle; this phase's list is new.
    if (installRow >= 0) {
      ownPending[installRow - mine0] = 0;
      ownFn[installRow - mine0] = isActive(activeBits, installRow) ? installFn : kEmptyKey;
    }
    for (int i = kThreads + tid; i < numInstall; i += kThreads) {
      const int r = pendRow[older][i];
      ownPending[r - mine0] = 0;
      ownFn[r - mine0] = isActive(activeBits, r) ? s.fnRescan[r] : kEmptyKey;
    }
    __syncthreads();
    older ^= 1;
    bool rescanned = false;
    if (!(k < pmin)) {
      // A pending row could hold the lowest key: rescan the previous phase's
      // rows now, settle every pending row, and select again from records
      // that hold no pending row (two extra barriers; rare).
      rescans(listCount, spilled, -1, g);
      if (!gridBarrier(s.bar, generation, &barrierOk))
        return;
      install(older);
      unsigned long long k2 = kEmptyKey, p2 = kEmptyKey;
      settledMin(k2, p2, -1);
      publish(k2, p2, kEmptyKey, older, kEmptyKey, -1);
      if (!gridBarrier(s.bar, generation, &barrierOk))
        return;
      k = kEmptyKey;
      p = kEmptyKey;
      if (tid < G) {
        const unsigned long long *r = recordsOf((published - 1) & 1) + (size_t)tid * kRecordWords;
        k = r[0];
        p = r[1];
      }
      blockMinPair(k, p, redK, redP);
      rescanned = true;
    }
    const int a = k == kEmptyKey ? -1 : keyIndex(k);
    const int b = p == kEmptyKey ? -1 : keyIndex(p);
    if (a < 0 || b <= a || b >= n) {
      // No pair (non-finite distances): the remaining rows report -1. Every
      // workgroup reduced the same records, so all of them stop here.
      if (g == 0)
        for (int r = step + tid; r < n - 1; r += kThreads)
          for (int c = 0; c < 4; ++c)
            s.tree[(size_t)r * 4 + c] = -1.0;
      break;
    }
    const float dab = fromOrderedBits((unsigned)(p >> 32));
    // Sizes and ids in memory are at least one phase old; the previous
    // merge's cluster is known here. Their loads join the update's loads.
    const int na = a == prevA ? prevSize : s.size[a];
    const int nb = b == prevA ? prevSize : s.size[b];
    int ia = 0, ib = 0;
    if (g == 0 && tid == 0) {
      const int prevId = s.firstNewId + step - 1;
      ia = a == prevA ? prevId : s.id[a];
      ib = b == prevA ? prevId : s.id[b];
    }
    // The update skips b itself, so its bit is cleared without a barrier
    // here (the next one, after the update, publishes it to the rescans).
    if (tid == 0)
      activeBits[b >> 5] &= ~(1u << (b & 31));
    // The first rescan item of this workgroup: its loads are issued before
    // the update (no entry it reads is written in this phase but D(row, a),
    // which the scan leaves out).
    const long long numItems = rescanned ? 0 : (long long)listCount * maxChunks;
    int firstRow = -1, firstJ0 = n;
    float firstV[kRescanPer];
    if (g < numItems) {
      firstRow = listRow(g / maxChunks, spilled);
      firstJ0 = firstRow + 1 + (g % maxChunks) * kRescanChunk;
      if (firstJ0 < n) {
        const size_t base = condensedIndex(firstRow, firstRow + 1, n) - (size_t)(firstRow + 1);
#pragma unroll
        for (int u = 0; u < kRescanPer; ++u) {
          const int j = firstJ0 + tid + u * kThreads;
          firstV[u] = j < n && j < firstJ0 + kRescanChunk ? s.D[base + j] : 0.0f;
        }
      }
    }
    lap(0);

    // Lance-Williams update of D(a, k) for this workgroup's k, its settled
    // rows' fn, and its part of row a's new fn.
    unsigned long long rowAMin = kEmptyKey, selK = kEmptyKey, selP = kEmptyKey;
    const int fresh = older ^ 1;
    // Rows in batches of kUpdatePer per thread, every load of a batch issued
    // before the first is used: one memory round trip per batch.
    for (int batch = mine0; batch < mine1; batch += kUpdatePer * kThreads) {
      bool live[kUpdatePer];
      size_t iakOf[kUpdatePer];
      float dakOf[kUpdatePer], dbkOf[kUpdatePer];
#pragma unroll
      for (int u = 0; u < kUpdatePer; ++u) {
        const int kk = batch + tid + u * kThreads;
        live[u] = kk < mine1 && kk != a && kk != b && isActive(activeBits, kk);
        if (live[u]) {
          iakOf[u] = kk < a ? condensedIndex(kk, a, n) : condensedIndex(a, kk, n);
          const size_t ibk = kk < b ? condensedIndex(kk, b, n) : condensedIndex(b, kk, n);
          dakOf[u] = s.D[iakOf[u]];
          dbkOf[u] = s.D[ibk];
        }
      }
#pragma unroll
      for (int u = 0; u < kUpdatePer; ++u) {
        if (!live[u])
          continue;
        const int kk = batch + tid + u * kThreads;
        const size_t iak = iakOf[u];
        const float dak = dakOf[u], dbk = dbkOf[u];
        const unsigned long long f = ownFn[kk - mine0];
        const float nd =
            (float)lanceWilliams(s.linkage, dak, dbk, dab, na, nb, ownSize[kk - mine0]);
        s.D[iak] = nd;
        if (kk > a)
          rowAMin = minKey(rowAMin, denseKey(nd, (unsigned)kk));
        if (ownPending[kk - mine0]) {
          // Being rescanned without column a: its new D(kk, a) joins here.
          if (kk < a)
            atomicMin(&s.fnRescan[kk], denseKey(nd, (unsigned)a));
          continue;
        }
        unsigned long long nf = f;
        bool rescan = false;
        if (kk < a) {
          // Every entry of row kk but D(kk, a) and D(kk, b) is unchanged.
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
        } else if (kk < b && f != kEmptyKey && keyIndex(f) == b) {
          rescan = true; // its nearest retired
        }
        if (rescan) {
          // Pending until the next phase's rescan; its new key is at least
          // the old one (every linkage here is reducible).
          const int r = atomicAdd(&pendCount[fresh], 1);
          pendRow[fresh][r] = kk;
          pendKey[fresh][r] = selectionKey(f, kk);
          ownPending[kk - mine0] = 1;
          s.fnRescan[kk] = kEmptyKey;
          nf = kEmptyKey;
        } else {
          pairMin(selK, selP, selectionKey(nf, kk), nf);
        }
        ownFn[kk - mine0] = nf;
      }
    }
    if (tid == 0 && b >= mine0 && b < mine1)
      ownFn[b - mine0] = kEmptyKey;
    if (g == 0 && tid == 0) {
      double *row = s.tree + (size_t)step * 4;
      row[0] = (double)min(ia, ib);
      row[1] = (double)max(ia, ib);
      row[2] = (double)dab;
      row[3] = (double)(na + nb);
    }
    __syncthreads();
    lap(1);

    // The first rescan item, loaded before the update (the same item for
    // every thread of the workgroup); its minimum joins the record's
    // reduction.
    unsigned long long firstMin = kEmptyKey;
    const bool firstLive = g < numItems && firstJ0 < n && isActive(activeBits, firstRow);
    if (firstLive) {
      const int j1 = min(n, firstJ0 + kRescanChunk);
#pragma unroll
      for (int u = 0; u < kRescanPer; ++u) {
        const int j = firstJ0 + tid + u * kThreads;
        if (j < j1 && j != a && isActive(activeBits, j))
          firstMin = minKey(firstMin, denseKey(firstV[u], (unsigned)j));
      }
    }
    // The rest of the previous phase's pending rows, rescanned by all
    // workgroups.
    if (!rescanned)
      rescans(listCount, spilled, a, (long long)g + G);
    publish(selK, selP, rowAMin, older, firstMin, firstLive ? firstRow : -1);
    prevA = a;
    prevSize = na + nb;
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
  DevArray<int> size, id;
  DevArray<unsigned long long> fnKey, records, fnRescan;
  DevArray<int> pendingRows;
  DevArray<int2> initItems;
  DevArray<double> tree;
  DevArray<unsigned long long> clocks;
  DevArray<unsigned int> barrierWords;
  Status s;
  const int maxChunks = (n - 1 + kChunk - 1) / kChunk;
  const long long initCount = (long long)(n - 1) * maxChunks;
  const int groups = (G + kGroup - 1) / kGroup;
  const size_t numBarrierWords = (size_t)kLineWords * (3 + groups);
  // The merge kernel keeps one bit per slot in LDS beside its static arrays.
  const size_t activeBytes = (size_t)(n + 31) / 32 * sizeof(unsigned int);
  int device = 0, maxLds = 0;
  s((int)hipGetDevice(&device));
  s((int)hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if (!s.ok())
    return s.code;
  hipFuncAttributes attr{};
  s((int)hipFuncGetAttributes(&attr, (const void *)denseMergeKernel));
  if (!s.ok())
    return s.code;
  if (activeBytes + attr.sharedSizeBytes > (size_t)maxLds)
    return kErrInvalidShape; // beyond any condensed matrix that fits in HBM
  s(size.alloc(n));
  s(id.alloc(n));
  s(fnKey.alloc(n));
  s(records.alloc((size_t)2 * G * kRecordWords));
  s(fnRescan.alloc(n));
  s(pendingRows.alloc((size_t)2 * G * kMaxOwned));
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
  st.fnKey = fnKey;
  st.records = records;
  st.fnRescan = fnRescan;
  st.pendingRows = pendingRows;
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
                                      args, (unsigned int)activeBytes, stream()));
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
