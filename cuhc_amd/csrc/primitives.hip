// Exclusive scans and stable LSD radix sorts for wave64 (see primitives.hpp).
//
// Both work on tiles of 4,096 items, one 256-thread workgroup (four waves)
// per tile. The scan reduces each tile, scans the tile sums (recursively,
// 4,096 sums per level), then scans every tile from its offset. A sort pass
// counts each tile's 8-bit digits, scans the counts in digit-major order (so
// every (digit, tile) gets its global offset), and scatters: the tile's
// items go in index order, 256 at a time, each wave ranking its lanes by
// digit with eight ballots (the lanes holding the same digit, and how many of
// them lie below), the waves ordered through LDS counts. Items keep their
// relative order within a digit, so each pass, and the sort, is stable.
#include "primitives.hpp"

#include "common.hpp"
#include "runtime.hpp"

#include <algorithm>

namespace cuhc {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int kTile = kThreads * kPerThread;
constexpr int kRadixBits = 8;
constexpr int kDigits = 1 << kRadixBits;

size_t align256(size_t bytes) { return (bytes + 255) / 256 * 256; }
int tilesOf(int count) { return (count + kTile - 1) / kTile; }

DEV int waveSum(int v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    v += __shfl_xor(v, m, 64);
  return v;
}

__global__ void __launch_bounds__(kThreads)
    tileSumsKernel(const int *__restrict__ in, int count, int *__restrict__ sums) {
  __shared__ int waveTotal[kThreads / 64];
  const size_t base = (size_t)blockIdx.x * kTile;
  int acc = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const size_t i = base + (size_t)j * kThreads + threadIdx.x;
    if (i < (size_t)count)
      acc += in[i];
  }
  acc = waveSum(acc);
  if ((threadIdx.x & 63) == 0)
    waveTotal[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    sums[blockIdx.x] = waveTotal[0] + waveTotal[1] + waveTotal[2] + waveTotal[3];
}

// Thread t scans items [t * 16, t * 16 + 16) of the tile in registers; the
// thread totals are scanned across the wave with shuffles and across the
// four waves through LDS. Every item is read before the barrier and written
// after it, so in == out is safe.
__global__ void __launch_bounds__(kThreads)
    scanTilesKernel(const int *in, int *out, int count, const int *__restrict__ tileOffsets) {
  __shared__ int waveTotal[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t first = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kPerThread;
  int v[kPerThread];
  int total = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const size_t i = first + j;
    v[j] = i < (size_t)count ? in[i] : 0;
    const int x = v[j];
    v[j] = total; // exclusive within the thread
    total += x;
  }
  int inclusive = total;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const int up = __shfl_up(inclusive, m, 64);
    if (lane >= m)
      inclusive += up;
  }
  if (lane == 63)
    waveTotal[wave] = inclusive;
  __syncthreads();
  int offset = (tileOffsets ? tileOffsets[blockIdx.x] : 0) + inclusive - total;
  for (int w = 0; w < wave; ++w)
    offset += waveTotal[w];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const size_t i = first + j;
    if (i < (size_t)count)
      out[i] = offset + v[j];
  }
}

template <class Key>
__global__ void __launch_bounds__(kThreads)
    digitCountsKernel(const Key *__restrict__ keys, int count, int shift, unsigned int digitMask,
                      int tiles, int *__restrict__ counts) {
  __shared__ int hist[kDigits];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int j = 0; j < kPerThread; ++j) {
    const size_t i = base + (size_t)j * kThreads + threadIdx.x;
    if (i < (size_t)count)
      atomicAdd(&hist[(unsigned int)(keys[i] >> shift) & digitMask], 1);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

template <class Key, class Value>
__global__ void __launch_bounds__(kThreads)
    scatterKernel(const Key *__restrict__ keysIn, const Value *__restrict__ valuesIn,
                  Key *__restrict__ keysOut, Value *__restrict__ valuesOut, int count, int shift,
                  unsigned int digitMask, int tiles, const int *__restrict__ offsets) {
  __shared__ int placed[kDigits];              // next output slot of each digit
  __shared__ int waveCount[kThreads / 64][kDigits];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  placed[threadIdx.x] = offsets[(size_t)threadIdx.x * tiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w)
    waveCount[w][threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  const size_t base = (size_t)blockIdx.x * kTile;
  for (int j = 0; j < kPerThread; ++j) {
    const size_t i = base + (size_t)j * kThreads + threadIdx.x;
    const bool valid = i < (size_t)count;
    Key key = 0;
    Value value = 0;
    unsigned int digit = 0;
    if (valid) {
      key = keysIn[i];
      value = valuesIn[i];
      digit = (unsigned int)(key >> shift) & digitMask;
    }
    // The lanes of this wave holding the same digit.
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool set = (digit >> b) & 1u;
      const unsigned long long bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    const int rank = __popcll(peers & below);
    if (valid && rank == 0)
      waveCount[wave][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = placed[digit] + rank;
      for (int w = 0; w < wave; ++w)
        pos += waveCount[w][digit];
      keysOut[pos] = key;
      valuesOut[pos] = value;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      add += waveCount[w][threadIdx.x];
      waveCount[w][threadIdx.x] = 0;
    }
    placed[threadIdx.x] += add;
    __syncthreads();
  }
}

template <class Key, class Value> size_t sortTempBytes(int count) {
  const int tiles = std::max(1, tilesOf(count));
  const int countsLen = kDigits * tiles;
  return align256((size_t)count * sizeof(Key)) + align256((size_t)count * sizeof(Value)) +
         2 * align256((size_t)countsLen * sizeof(int)) + scanTempBytes(countsLen);
}

template <class Key, class Value>
int sortPairs(const Key *keysIn, Key *keysOut, const Value *valuesIn, Value *valuesOut, int count,
              int bits, void *temp) {
  if (count <= 0)
    return 0;
  Status s;
  const int passes = (bits + kRadixBits - 1) / kRadixBits;
  if (passes == 0) {
    s((int)hipMemcpyAsync(keysOut, keysIn, (size_t)count * sizeof(Key),
                          hipMemcpyDeviceToDevice, stream()));
    s((int)hipMemcpyAsync(valuesOut, valuesIn, (size_t)count * sizeof(Value),
                          hipMemcpyDeviceToDevice, stream()));
    return s.code;
  }
  const int tiles = tilesOf(count);
  const int countsLen = kDigits * tiles;
  char *t = static_cast<char *>(temp);
  Key *keysAlt = reinterpret_cast<Key *>(t);
  t += align256((size_t)count * sizeof(Key));
  Value *valuesAlt = reinterpret_cast<Value *>(t);
  t += align256((size_t)count * sizeof(Value));
  int *counts = reinterpret_cast<int *>(t);
  t += align256((size_t)countsLen * sizeof(int));
  int *offsets = reinterpret_cast<int *>(t);
  t += align256((size_t)countsLen * sizeof(int));
  void *scanTemp = t;
  const Key *srcK = keysIn;
  const Value *srcV = valuesIn;
  for (int p = 0; p < passes && s.ok(); ++p) {
    // The last pass lands in the output buffers; the passes alternate.
    const bool toOut = ((passes - 1 - p) & 1) == 0;
    Key *dstK = toOut ? keysOut : keysAlt;
    Value *dstV = toOut ? valuesOut : valuesAlt;
    const int shift = p * kRadixBits;
    const int width = std::min(kRadixBits, bits - shift);
    const unsigned int mask = (1u << width) - 1u;
    hipLaunchKernelGGL((digitCountsKernel<Key>), dim3(tiles), dim3(kThreads), 0, stream(), srcK,
                       count, shift, mask, tiles, counts);
    s((int)hipGetLastError());
    s(exclusiveScan(counts, offsets, countsLen, scanTemp));
    hipLaunchKernelGGL((scatterKernel<Key, Value>), dim3(tiles), dim3(kThreads), 0, stream(),
                       srcK, srcV, dstK, dstV, count, shift, mask, tiles, offsets);
    s((int)hipGetLastError());
    srcK = dstK;
    srcV = dstV;
  }
  return s.code;
}

} // namespace

size_t scanTempBytes(int count) {
  const int tiles = tilesOf(count);
  if (tiles <= 1)
    return 256;
  return align256((size_t)tiles * sizeof(int)) + scanTempBytes(tiles);
}

int exclusiveScan(const int *in, int *out, int count, void *temp) {
  if (count <= 0)
    return 0;
  const int tiles = tilesOf(count);
  if (tiles == 1) {
    hipLaunchKernelGGL(scanTilesKernel, dim3(1), dim3(kThreads), 0, stream(), in, out, count,
                       nullptr);
    return (int)hipGetLastError();
  }
  int *sums = static_cast<int *>(temp);
  void *rest = static_cast<char *>(temp) + align256((size_t)tiles * sizeof(int));
  Status s;
  hipLaunchKernelGGL(tileSumsKernel, dim3(tiles), dim3(kThreads), 0, stream(), in, count, sums);
  s((int)hipGetLastError());
  s(exclusiveScan(sums, sums, tiles, rest));
  hipLaunchKernelGGL(scanTilesKernel, dim3(tiles), dim3(kThreads), 0, stream(), in, out, count,
                     sums);
  s((int)hipGetLastError());
  return s.code;
}

size_t sortTempBytes32(int count) { return sortTempBytes<unsigned int, int>(count); }

int radixSortPairs(const unsigned int *keysIn, unsigned int *keysOut, const int *valuesIn,
                   int *valuesOut, int count, int bits, void *temp) {
  return sortPairs(keysIn, keysOut, valuesIn, valuesOut, count, bits, temp);
}

size_t sortTempBytes64(int count) { return sortTempBytes<unsigned long long, unsigned int>(count); }

int radixSortPairs(const unsigned long long *keysIn, unsigned long long *keysOut,
                   const unsigned int *valuesIn, unsigned int *valuesOut, int count, int bits,
                   void *temp) {
  return sortPairs(keysIn, keysOut, valuesIn, valuesOut, count, bits, temp);
}

} // namespace cuhc

// Test hooks (device arrays; the call returns when the result is in place).
extern "C" int cuhc_exclusive_scan(const int *in, int *out, int count) {
  using namespace cuhc;
  if (count < 0)
    return kErrInvalidShape;
  DevArray<unsigned char> temp;
  Status s;
  s(temp.alloc(scanTempBytes(count)));
  if (s.ok())
    s(exclusiveScan(in, out, count, temp));
  s(sync());
  return s.code;
}

extern "C" int cuhc_radix_sort_pairs_u32(const unsigned int *keysIn, unsigned int *keysOut,
                                         const int *valuesIn, int *valuesOut, int count,
                                         int bits) {
  using namespace cuhc;
  if (count < 0 || bits < 0 || bits > 32)
    return kErrInvalidShape;
  DevArray<unsigned char> temp;
  Status s;
  s(temp.alloc(sortTempBytes32(count)));
  if (s.ok())
    s(radixSortPairs(keysIn, keysOut, valuesIn, valuesOut, count, bits, temp));
  s(sync());
  return s.code;
}

extern "C" int cuhc_radix_sort_pairs_u64(const unsigned long long *keysIn,
                                         unsigned long long *keysOut,
                                         const unsigned int *valuesIn, unsigned int *valuesOut,
                                         int count, int bits) {
  using namespace cuhc;
  if (count < 0 || bits < 0 || bits > 64)
    return kErrInvalidShape;
  DevArray<unsigned char> temp;
  Status s;
  s(temp.alloc(sortTempBytes64(count)));
  if (s.ok())
    s(radixSortPairs(keysIn, keysOut, valuesIn, valuesOut, count, bits, temp));
  s(sync());
  return s.code;
}
