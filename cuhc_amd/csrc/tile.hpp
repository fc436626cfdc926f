// FP32 direct-difference distance tiles for gfx950 (wave64).
//
// A workgroup of 256 threads (4 waves) owns a DIM x DIM tile of the distance
// matrix, DIM = 16 * TR. Thread t = (tr, tc) = (t / 16, t % 16) owns TR x TR
// outputs: tile rows (r / 4) * 64 + 4 tr + r % 4 and the same pattern for
// columns. The 16 threads of one tile row are 16 consecutive lanes of one wave,
// so row reductions are xor-shuffles inside a 16-lane group, and the 4 threads
// of a wave that share a column tc are lanes tc, tc+16, tc+32, tc+48.
//
// Features are staged in LDS transposed (feature-major, rows padded to DIM + 4
// floats so the transposed stores spread over the banks) FILL features at a
// time; each thread reads its 4-row groups as float4. Each output folds the
// features in order 0..d-1 with the metric's fold, which fixes the FP32 value
// bit for bit (the semantics of reference src/distance_matrix.cuh:108-175).
// Features beyond d are zero-padded: a no-op for every built-in fold.
#pragma once

#include "common.hpp"

namespace cuhc {

template <int TR_, int FILL_> struct Tile {
  static constexpr int TR = TR_;          // outputs per thread side (4 or 8)
  static constexpr int DIM = 16 * TR;     // tile side in points
  static constexpr int FILL = FILL_;      // features staged per LDS fill (multiple of 8)
  static constexpr int LD = DIM + 4;      // padded LDS row (floats)
  static constexpr int THREADS = 256;
  static_assert(TR == 4 || TR == 8, "TR");
  static_assert(FILL % 8 == 0, "FILL");

  DEV static int tr() { return threadIdx.x >> 4; }
  DEV static int tc() { return threadIdx.x & 15; }
  DEV static int rowInTile(int r) { return (r >> 2) * 64 + 4 * tr() + (r & 3); }
  DEV static int colInTile(int c) { return (c >> 2) * 64 + 4 * tc() + (c & 3); }
};

template <class T> struct TileSmem {
  float a[T::FILL][T::LD];
  float b[T::FILL][T::LD];
};

// Copies features [f0, f0 + nf8) of points [p0, p0 + DIM) into s[k][p]
// (zero outside the matrix or past d). A thread issues its loads in batches
// of up to four before storing them to LDS, so a fill costs one to four
// memory round trips instead of one per element. idx / nf8 is
// (idx * magic) >> 20 with magic = ceil(2^20 / nf8): exact for idx < 2^12
// and nf8 <= 32 (the rounding adds less than 2^-8, below the 1 / nf8 gap to
// the next integer).
template <class T>
DEV void stageSlice(float (*s)[T::LD], const float *__restrict__ pts, int n, int d, int p0, int f0,
                    int nf8) {
  constexpr int kPer = T::DIM * T::FILL / T::THREADS;
  constexpr int kBatch = kPer < 4 ? kPer : 4;
  static_assert(T::DIM * T::FILL <= 4096 && T::FILL <= 32 && kPer % kBatch == 0,
                "division by multiplication, batches");
  const int total = T::DIM * nf8;
  const unsigned magic = ((1u << 20) + (unsigned)nf8 - 1u) / (unsigned)nf8;
#pragma unroll
  for (int b = 0; b < kPer; b += kBatch) {
    float v[kBatch];
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int idx = (int)threadIdx.x + (b + i) * T::THREADS;
      const int p = (int)(((unsigned)idx * magic) >> 20), k = idx - p * nf8;
      const int point = p0 + p, f = f0 + k;
      v[i] = (idx < total && point < n && f < d) ? pts[(size_t)point * d + f] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int idx = (int)threadIdx.x + (b + i) * T::THREADS;
      const int p = (int)(((unsigned)idx * magic) >> 20), k = idx - p * nf8;
      if (idx < total)
        s[k][p] = v[i];
    }
  }
}

// No extra staging (see computeTile).
struct NoExtraStage {
  struct Value {};
  DEV Value load() const { return {}; }
  DEV void store(const Value &) const {}
};

// out[r][c] = metric(A[rowA0 + rowInTile(r)], B[colB0 + colInTile(c)]), or
// the raw accumulator when kFinish is false (scalar accumulators only).
// Entries outside the matrices are garbage. Contains __syncthreads(); every
// thread of the workgroup must call it. extra.load() is issued with the first
// fill's loads and extra.store() runs after that fill's LDS stores, before
// its barrier: per-tile data the caller stages in LDS shares the fill's
// memory round trip. With kRowsResident (only for d <= T::FILL: one fill)
// sm.a already holds the row slice (stageRows) and only the columns are
// staged.
template <class T, class Metric, bool kFinish = true, bool kRowsResident = false,
          class Extra = NoExtraStage>
DEV void computeTile(TileSmem<T> &sm, const float *__restrict__ A, int nA,
                     const float *__restrict__ B, int nB, int d, const Metric &metric, int rowA0,
                     int colB0, float (&out)[T::TR][T::TR], const Extra &extra = Extra{}) {
  constexpr int TR = T::TR;
  using Acc = typename Metric::Acc;
  Acc acc[TR][TR];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int c = 0; c < TR; ++c)
      acc[r][c] = metric.init();

  const int tr4 = 4 * T::tr(), tc4 = 4 * T::tc();
  for (int f0 = 0; f0 < d; f0 += T::FILL) {
    const int nf = min(T::FILL, d - f0);
    const int nf8 = (nf + 7) & ~7;
    __syncthreads(); // previous fill fully consumed
    typename Extra::Value ev{};
    if (f0 == 0)
      ev = extra.load();
    if (!kRowsResident)
      stageSlice<T>(sm.a, A, nA, d, rowA0, f0, nf8);
    stageSlice<T>(sm.b, B, nB, d, colB0, f0, nf8);
    if (f0 == 0)
      extra.store(ev);
    __syncthreads();
    for (int k0 = 0; k0 < nf8; k0 += 8) {
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const int k = k0 + kk;
        float av[TR], bv[TR];
#pragma unroll
        for (int h = 0; h < TR / 4; ++h) {
          const float4 a4 = *reinterpret_cast<const float4 *>(&sm.a[k][h * 64 + tr4]);
          const float4 b4 = *reinterpret_cast<const float4 *>(&sm.b[k][h * 64 + tc4]);
          av[4 * h + 0] = a4.x; av[4 * h + 1] = a4.y; av[4 * h + 2] = a4.z; av[4 * h + 3] = a4.w;
          bv[4 * h + 0] = b4.x; bv[4 * h + 1] = b4.y; bv[4 * h + 2] = b4.z; bv[4 * h + 3] = b4.w;
        }
#pragma unroll
        for (int r = 0; r < TR; ++r)
#pragma unroll
          for (int c = 0; c < TR; ++c)
            acc[r][c] = metric.fold(acc[r][c], av[r], bv[c]);
      }
    }
  }
  __syncthreads(); // LDS free for the caller / next tile
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int c = 0; c < TR; ++c) {
      if constexpr (kFinish)
        out[r][c] = metric.finish(acc[r][c]);
      else
        out[r][c] = acc[r][c];
    }
}

// Stages rows [rowA0, rowA0 + DIM) of A into sm.a once, for the
// kRowsResident tiles of a workgroup whose rows stay fixed (d <= T::FILL).
// Contains __syncthreads().
template <class T>
DEV void stageRows(TileSmem<T> &sm, const float *__restrict__ A, int nA, int d, int rowA0) {
  __syncthreads();
  stageSlice<T>(sm.a, A, nA, d, rowA0, 0, (d + 7) & ~7);
  __syncthreads();
}

// Tile geometry per metric: wide (cosine) accumulators get 4x4 outputs per
// thread so they stay in registers.
template <class Metric, int FILL> struct TileFor {
  using type = Tile<Metric::kWide ? 4 : 8, FILL>;
};

// Linear index over the upper triangle of tile pairs, t = J (J + 1) / 2 + I
// with I <= J, decoded with a double sqrt and an exact integer fix-up.
DEV void decodeTrianglePair(long long t, int &I, int &J) {
  long long j = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (j * (j + 1) / 2 > t)
    --j;
  while ((j + 1) * (j + 2) / 2 <= t)
    ++j;
  J = (int)j;
  I = (int)(t - j * (j + 1) / 2);
}

} // namespace cuhc
