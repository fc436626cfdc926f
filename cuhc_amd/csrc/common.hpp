// Shared device/host definitions for the gfx950 clustering kernels.
//
// Distance semantics (the exact FP32 value every kernel and the oracle agree on)
// follow the reference's metric folds (reference src/distances.cuh:21-83):
//   sqeuclidean  acc = fma(a_k - b_k, a_k - b_k, acc) over k = 0..d-1 in order
//   euclidean    sqrt(sqeuclidean)
//   manhattan    acc + |a_k - b_k|
//   chebyshev    max(acc, |a_k - b_k|)
//   minkowski    sum |a_k - b_k|^p, then ^(1/p)
//   cosine       (sum ab, sum aa, sum bb) by fma; 1 - ab / sqrt(aa * bb), or 1
//                when the denominator is 0
// Every fold is a no-op for a zero feature pair, so tiles pad features with 0.
// The library is built with -ffp-contract=off: the only fused operations are
// the explicit fmaf calls, and division and sqrt are correctly rounded.
//
// Argmin filters: for metrics with kAccFilter, accBound(x) is an upper bound
// on every accumulator whose finish() is <= x, so an argmin pass can compare
// the raw accumulator with it and finish only the pairs that pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <float.h>

#define HD __host__ __device__ __forceinline__
#define DEV __device__ __forceinline__

namespace cuhc {

// ---------------------------------------------------------------------------
// Error codes (reference src/capi.cu:26-30, src/mst.cuh:80-81)

constexpr int kErrInvalidMetric = -1;
constexpr int kErrInvalidShape = -2;
constexpr int kErrInvalidLinkage = -3;
constexpr int kErrInvalidAlgorithm = -4;
constexpr int kErrDisconnected = -5;
constexpr int kErrNonFinite = -6;
constexpr int kErrWorkspace = -7; // no counterpart: the fixed workspace is too small

enum MetricId : int {
  kEuclidean = 0,
  kSqEuclidean = 1,
  kManhattan = 2,
  kChebyshev = 3,
  kMinkowski = 4,
  kCosine = 5,
};

// ---------------------------------------------------------------------------
// Metric folds

struct SqEuclidean {
  using Acc = float;
  static constexpr bool kBoxBound = true;
  static constexpr bool kWide = false;
  static constexpr bool kAccFilter = true;
  HD float accBound(float x) const { return x; }
  HD Acc init() const { return 0.0f; }
  HD Acc fold(Acc acc, float a, float b) const {
    const float diff = a - b;
    return fmaf(diff, diff, acc);
  }
  HD float finish(Acc acc) const { return acc; }
};

struct Euclidean : SqEuclidean {
  HD float finish(Acc acc) const { return sqrtf(acc); }
  // sqrtf(a) <= x implies a <= x^2 (1 + 2^-23); the float product and the
  // factor 1 + 2^-20 cover that and their own rounding, FLT_MIN underflow.
  HD float accBound(float x) const { return x * x * (1.0f + 0x1p-20f) + FLT_MIN; }
};

struct Manhattan {
  using Acc = float;
  static constexpr bool kBoxBound = true;
  static constexpr bool kWide = false;
  static constexpr bool kAccFilter = true;
  HD float accBound(float x) const { return x; }
  HD Acc init() const { return 0.0f; }
  HD Acc fold(Acc acc, float a, float b) const { return acc + fabsf(a - b); }
  HD float finish(Acc acc) const { return acc; }
};

struct Chebyshev {
  using Acc = float;
  static constexpr bool kBoxBound = true;
  static constexpr bool kWide = false;
  static constexpr bool kAccFilter = true;
  HD float accBound(float x) const { return x; }
  HD Acc init() const { return 0.0f; }
  HD Acc fold(Acc acc, float a, float b) const { return fmaxf(acc, fabsf(a - b)); }
  HD float finish(Acc acc) const { return acc; }
};

// |x|^p through powf. Kept out of line on the device: inlined into the
// unrolled tile loop, hundreds of powf copies would flood the instruction
// cache. Not bit-identical to other libraries' powf; Minkowski is compared
// with a tolerance.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __noinline__ inline float minkowskiTerm(float x, float p) { return powf(x, p); }
#else
inline float minkowskiTerm(float x, float p) { return powf(x, p); }
#endif

struct Minkowski {
  float p;
  using Acc = float;
  static constexpr bool kBoxBound = true;
  static constexpr bool kWide = false;
  static constexpr bool kAccFilter = false;
  HD float accBound(float) const { return INFINITY; }
  HD Acc init() const { return 0.0f; }
  HD Acc fold(Acc acc, float a, float b) const {
    const float x = fabsf(a - b);
    return x > 0.0f ? acc + minkowskiTerm(x, p) : acc;
  }
  HD float finish(Acc acc) const { return powf(acc, 1.0f / p); }
};

struct CosAcc {
  float ab, aa, bb;
};

struct Cosine {
  using Acc = CosAcc;
  static constexpr bool kBoxBound = false;
  static constexpr bool kWide = true;
  static constexpr bool kAccFilter = false;
  HD float accBound(float) const { return INFINITY; }
  HD Acc init() const { return CosAcc{0.0f, 0.0f, 0.0f}; }
  HD Acc fold(Acc acc, float a, float b) const {
    acc.ab = fmaf(a, b, acc.ab);
    acc.aa = fmaf(a, a, acc.aa);
    acc.bb = fmaf(b, b, acc.bb);
    return acc;
  }
  HD float finish(Acc acc) const {
    const float denom = sqrtf(acc.aa * acc.bb);
    return denom > 0.0f ? 1.0f - acc.ab / denom : 1.0f;
  }
};

// ---------------------------------------------------------------------------
// (weight, index) keys: the float bits in the high word, the index in the low
// word. Unsigned order equals (weight, index) order for weights >= 0; negative
// weights and NaN clamp to 0 (reference src/reductions.cuh:64-71).

constexpr unsigned long long kEmptyKey = ~0ull;

HD unsigned int bitsOf(float f) { return __builtin_bit_cast(unsigned int, f); }
HD float floatOf(unsigned int u) { return __builtin_bit_cast(float, u); }

HD unsigned long long packKey(float w, unsigned int index) {
  const float c = w > 0.0f ? w : 0.0f;
  return ((unsigned long long)bitsOf(c) << 32) | index;
}
HD float keyWeight(unsigned long long key) { return floatOf((unsigned int)(key >> 32)); }
HD int keyIndex(unsigned long long key) { return (int)(key & 0xffffffffull); }

HD unsigned long long minKey(unsigned long long a, unsigned long long b) { return a < b ? a : b; }

// ---------------------------------------------------------------------------
// Edge weights: per-point value loaded once per row/column, then a symmetric
// combination with the pair distance (reference src/distance_matrix.cuh:202-219,
// src/hdbscan.cuh:342-351). distBound(w) is an upper bound on every distance
// whose weight is <= w (for any per-point values).

struct PlainWeight {
  using Value = float;
  DEV Value value(int) const { return 0.0f; }
  DEV float operator()(float dist, Value, Value) const { return dist; }
  DEV float distBound(float w) const { return w; }
  static constexpr bool kHasValue = false;
  static constexpr bool kDistBound = true;
};

// kScaled: alpha != 1 (the division is a separate instantiation: a select
// between dist and dist / alpha would evaluate the division for every pair).
template <bool kScaled> struct MReachWeightT {
  const float *core;
  float alpha;
  using Value = float;
  DEV Value value(int p) const { return core[p]; }
  DEV float operator()(float dist, Value ci, Value cj) const {
    return fmaxf(fmaxf(ci, cj), kScaled ? dist / alpha : dist);
  }
  // fl(dist / alpha) <= w implies dist <= w alpha (1 + 2^-24).
  DEV float distBound(float w) const {
    return kScaled ? w * alpha * (1.0f + 0x1p-21f) + FLT_MIN : w;
  }
  static constexpr bool kHasValue = true;
  static constexpr bool kDistBound = true;
};

// Ward's weight for the nearest-neighbour choice between cluster centroids:
// 2 n_i n_j / (n_i + n_j) * |c_i - c_j|^2 in FP32 with a correctly rounded
// division. Symmetric bit for bit. Merge heights are recomputed in FP64.
struct WardWeight {
  const int *size;
  using Value = float;
  DEV Value value(int p) const { return (float)size[p]; }
  DEV float operator()(float sq, Value si, Value sj) const {
    return (2.0f * (si * sj)) / (si + sj) * sq;
  }
  DEV float distBound(float) const { return INFINITY; }
  static constexpr bool kHasValue = true;
  static constexpr bool kDistBound = false; // the weight scales the distance by cluster sizes
};

// ---------------------------------------------------------------------------
// wave64 helpers

DEV unsigned long long shflXorKey(unsigned long long v, int mask) {
  const unsigned int lo = __shfl_xor((unsigned int)(v & 0xffffffffu), mask, 64);
  const unsigned int hi = __shfl_xor((unsigned int)(v >> 32), mask, 64);
  return ((unsigned long long)hi << 32) | lo;
}

DEV float atomicMinNonNegFloat(float *addr, float v) {
  // Non-negative floats order like their bit patterns.
  return __uint_as_float(atomicMin((unsigned int *)addr, __float_as_uint(v)));
}

} // namespace cuhc
