// MFMA issue-rate probe: the matrix-core throughput that the candidate pass's
// instruction mix can reach on this GPU, measured with synthetic register and
// LDS operands (no memory traffic). It prices the candidate kernel against an
// attainable ceiling as well as against the data-sheet peak. No reference
// counterpart; used by the benchmark only.
//
// Shapes (flops per wave per step: 524,288 for 0-2 and 6, 262,144 for 3):
//   0  v_mfma_f32_32x32x16_f16, 8 accumulators updated in place (GEMM form)
//   1  v_mfma_f32_16x16x32_f16, 16 accumulators updated in place (GEMM form)
//   2  the candidate loop: 4 row blocks x 2 K steps x 2 column blocks, C seeded
//      with column norms read from LDS, A operands read from LDS, each block
//      reduced to this lane's minimum of 16 values and compared with a
//      per-row threshold under one wave-uniform ballot
//   3  the same loop on v_mfma_f32_16x16x32_f16 (d = 32 in one K step): 8 row
//      blocks of 16 x 2 column blocks of 16, seeded C, LDS operands, each
//      lane's 8 values per row block reduced and compared under one ballot
//   4  shape 2 with 2 row blocks per wave at four waves per SIMD (at most 128
//      VGPRs), 262,144 flops per wave per step
//   5  shape 3 with 4 row blocks per wave at four waves per SIMD, 131,072
//      flops per wave per step
//   6  shape 2 with a symmetric epilogue: besides each row's minimum, each
//      column's minimum over the wave's 128 rows (in-lane over the row
//      blocks) is compared with a per-column threshold read from LDS, so one
//      value serves the pair in both directions; counted as twice its flops
#include "common.hpp"
#include "runtime.hpp"

#include "../../include/cuhc.h"

#include <cmath>

namespace cuhc {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kProbeThreads = 256;

// The empty asm statements make an operand opaque to the compiler each step,
// so loop-invariant MFMAs are not hoisted; they emit no instruction.
template <int SHAPE>
__global__ void __launch_bounds__(kProbeThreads, 2) mfmaProbeKernel(int iters, float *out) {
  const int lane = threadIdx.x & 63;
  float t = 0.0f;
  if (SHAPE == 0 || SHAPE == 1) {
    half8 a[2], b[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a[0][q] = (_Float16)(q * 0.01f);
      a[1][q] = (_Float16)(q * 0.02f);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int q = 0; q < 8; ++q)
        b[i][q] = (_Float16)((i + q + lane) * 0.001f);
    if (SHAPE == 0) {
      f32x16 acc[8] = {};
#pragma unroll 1
      for (int it = 0; it < iters; ++it) {
        asm volatile("" : "+v"(a[0]), "+v"(a[1]));
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int i = 0; i < 8; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[k], b[i], acc[i], 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i)
        t += acc[i][0] + acc[i][15];
    } else {
      f32x4 acc[16] = {};
#pragma unroll 1
      for (int it = 0; it < iters; ++it) {
        asm volatile("" : "+v"(a[0]), "+v"(a[1]));
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int i = 0; i < 16; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[k], b[i & 7], acc[i], 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        t += acc[i][0] + acc[i][3];
    }
  } else if (SHAPE == 3) {
    __shared__ __align__(16) _Float16 sx[256 * 32];
    __shared__ __align__(16) float sn[256];
    const int l16 = lane & 15, q = lane >> 4;
    for (int i = threadIdx.x; i < 256 * 32; i += kProbeThreads)
      sx[i] = (_Float16)((i % 7) * 0.01f);
    sn[threadIdx.x] = (float)(threadIdx.x % 13);
    __syncthreads();
    half8 bop[8];
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        bop[rb][k] = (_Float16)((rb + k + lane) * -0.001f);
    float th[8];
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
      th[rb] = -1.0f - rb; // below every value: the record branch stays cold
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      const int cb = (it & 7) * 2; // column blocks 0..15 of the tile
      half8 a[2];
      f32x4 seed[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        a[c] = *reinterpret_cast<const half8 *>(
            &sx[((cb + c) * 16 + l16) * 32 + 8 * (q ^ (l16 & 3))]);
        seed[c] = *reinterpret_cast<const f32x4 *>(&sn[(cb + c) * 16 + 4 * q]);
      }
      f32x4 acc[2][8];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          acc[c][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[c], bop[rb], seed[c], 0, 0, 0);
      bool record = false;
      float m[8];
#pragma unroll
      for (int rb = 0; rb < 8; ++rb) {
        m[rb] = fminf(fminf(fminf(acc[0][rb][0], acc[0][rb][1]), fminf(acc[0][rb][2], acc[0][rb][3])),
                      fminf(fminf(acc[1][rb][0], acc[1][rb][1]), fminf(acc[1][rb][2], acc[1][rb][3])));
        record |= m[rb] <= th[rb];
      }
      if (__ballot(record) != 0) {
#pragma unroll
        for (int rb = 0; rb < 8; ++rb)
          th[rb] = fminf(th[rb], m[rb] - 1.0f);
      }
      t += m[0];
    }
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
      t += th[rb];
  } else {
    // 256 columns of 32 halves, their norms and (shape 6) their thresholds,
    // as one staged column tile.
    __shared__ __align__(16) _Float16 sx[256 * 32];
    __shared__ __align__(16) float sn[256];
    __shared__ __align__(16) float sth[256];
    const int l32 = lane & 31, hi = lane >> 5;
    for (int i = threadIdx.x; i < 256 * 32; i += kProbeThreads)
      sx[i] = (_Float16)((i % 7) * 0.01f);
    sn[threadIdx.x] = (float)(threadIdx.x % 13);
    sth[threadIdx.x] = -1.0f - (float)(threadIdx.x % 5); // below every value
    __syncthreads();
    half8 bop[4][2];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          bop[rb][kb][q] = (_Float16)((rb + kb + q + lane) * -0.001f);
    float th[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
      th[rb] = -1.0f - rb; // below every value: the record branch stays cold
    auto block = [&](int cb, f32x16 (&acc)[4]) {
      f32x16 seed;
      half8 a[2];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 v = *reinterpret_cast<const float4 *>(&sn[cb * 32 + g * 8 + hi * 4]);
        seed[4 * g] = v.x;
        seed[4 * g + 1] = v.y;
        seed[4 * g + 2] = v.z;
        seed[4 * g + 3] = v.w;
      }
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
        a[kb] = *reinterpret_cast<const half8 *>(
            &sx[(cb * 32 + l32) * 32 + 8 * ((kb * 2 + hi) ^ (l32 & 3))]);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], bop[rb][0], seed, 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], bop[rb][1], acc[rb], 0, 0, 0);
    };
    auto minOf = [](const f32x16 &v) {
      float m = v[0];
#pragma unroll
      for (int r = 1; r < 16; ++r)
        m = fminf(m, v[r]);
      return m;
    };
    // Symmetric epilogue: each of the lane's 16 column slots, minimum over
    // the row blocks, against the column's threshold (slot r is column
    // 8 * (r / 4) + 4 * hi + r % 4 of the block).
    auto columnTest = [&](int cb, const f32x16 (&acc)[4]) {
      bool hit = false;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 t = *reinterpret_cast<const float4 *>(&sth[cb * 32 + g * 8 + hi * 4]);
        const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * g + q;
          const float c = fminf(fminf(acc[0][r], acc[1][r]), fminf(acc[2][r], acc[3][r]));
          hit |= c <= tv[q];
        }
      }
      return hit;
    };
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      const int cb = (it & 3) * 2; // column blocks 0..7 of the tile
      f32x16 acc[4];
      float mA[4];
      bool record = false;
      block(cb, acc);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
        mA[rb] = minOf(acc[rb]);
      if (SHAPE == 6)
        record |= columnTest(cb, acc);
      block(cb + 1, acc);
      if (SHAPE == 6)
        record |= columnTest(cb + 1, acc);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        mA[rb] = fminf(mA[rb], minOf(acc[rb]));
        record |= mA[rb] <= th[rb];
      }
      if (__ballot(record) != 0) {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
          th[rb] = fminf(th[rb], mA[rb] - 1.0f);
      }
      t += mA[0];
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
      t += th[rb];
  }
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = t;
}

// Shapes 4 and 5: the candidate loops with fewer row blocks per wave, so that
// four waves fit on each SIMD.
template <int SHAPE>
__global__ void __launch_bounds__(kProbeThreads, 4) mfmaProbeSlimKernel(int iters, float *out) {
  const int lane = threadIdx.x & 63;
  float t = 0.0f;
  __shared__ __align__(16) _Float16 sx[256 * 32];
  __shared__ __align__(16) float sn[256];
  for (int i = threadIdx.x; i < 256 * 32; i += kProbeThreads)
    sx[i] = (_Float16)((i % 7) * 0.01f);
  sn[threadIdx.x] = (float)(threadIdx.x % 13);
  __syncthreads();
  if (SHAPE == 4) {
    const int l32 = lane & 31, hi = lane >> 5;
    half8 bop[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          bop[rb][kb][q] = (_Float16)((rb + kb + q + lane) * -0.001f);
    float th[2] = {-1.0f, -2.0f}; // below every value: the record branch stays cold
    auto block = [&](int cb, f32x16 (&acc)[2]) {
      f32x16 seed;
      half8 a[2];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 v = *reinterpret_cast<const float4 *>(&sn[cb * 32 + g * 8 + hi * 4]);
        seed[4 * g] = v.x;
        seed[4 * g + 1] = v.y;
        seed[4 * g + 2] = v.z;
        seed[4 * g + 3] = v.w;
      }
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
        a[kb] = *reinterpret_cast<const half8 *>(
            &sx[(cb * 32 + l32) * 32 + 8 * ((kb * 2 + hi) ^ (l32 & 3))]);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], bop[rb][0], seed, 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], bop[rb][1], acc[rb], 0, 0, 0);
    };
    auto minOf = [](const f32x16 &v) {
      float m = v[0];
#pragma unroll
      for (int r = 1; r < 16; ++r)
        m = fminf(m, v[r]);
      return m;
    };
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      const int cb = (it & 3) * 2;
      f32x16 acc[2];
      float mA[2];
      block(cb, acc);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        mA[rb] = minOf(acc[rb]);
      block(cb + 1, acc);
      bool record = false;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        mA[rb] = fminf(mA[rb], minOf(acc[rb]));
        record |= mA[rb] <= th[rb];
      }
      if (__ballot(record) != 0) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
          th[rb] = fminf(th[rb], mA[rb] - 1.0f);
      }
      t += mA[0];
    }
    t += th[0] + th[1];
  } else {
    const int l16 = lane & 15, q = lane >> 4;
    half8 bop[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        bop[rb][k] = (_Float16)((rb + k + lane) * -0.001f);
    float th[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
      th[rb] = -1.0f - rb; // below every value: the record branch stays cold
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      const int cb = (it & 7) * 2;
      half8 a[2];
      f32x4 seed[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        a[c] = *reinterpret_cast<const half8 *>(
            &sx[((cb + c) * 16 + l16) * 32 + 8 * (q ^ (l16 & 3))]);
        seed[c] = *reinterpret_cast<const f32x4 *>(&sn[(cb + c) * 16 + 4 * q]);
      }
      f32x4 acc[2][4];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
          acc[c][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[c], bop[rb], seed[c], 0, 0, 0);
      bool record = false;
      float m[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        m[rb] = fminf(fminf(fminf(acc[0][rb][0], acc[0][rb][1]), fminf(acc[0][rb][2], acc[0][rb][3])),
                      fminf(fminf(acc[1][rb][0], acc[1][rb][1]), fminf(acc[1][rb][2], acc[1][rb][3])));
        record |= m[rb] <= th[rb];
      }
      if (__ballot(record) != 0) {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
          th[rb] = fminf(th[rb], m[rb] - 1.0f);
      }
      t += m[0];
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
      t += th[rb];
  }
  out[(size_t)blockIdx.x * kProbeThreads + threadIdx.x] = t;
}

} // namespace
} // namespace cuhc

extern "C" int cuhc_mfma_probe(int shape, int workgroups, int iters, float *msPerLaunch,
                               double *flopsPerLaunch) {
  using namespace cuhc;
  if (shape < 0 || shape > 6 || workgroups < 1 || iters < 1 || !msPerLaunch || !flopsPerLaunch)
    return kErrInvalidShape;
  Status s;
  DevArray<float> out;
  s(out.alloc((size_t)workgroups * kProbeThreads));
  hipEvent_t start = nullptr, stop = nullptr;
  s((int)hipEventCreate(&start));
  s((int)hipEventCreate(&stop));
  constexpr int kReps = 3;
  auto launch = [&] {
    switch (shape) {
    case 0:
      hipLaunchKernelGGL(mfmaProbeKernel<0>, dim3(workgroups), dim3(kProbeThreads), 0, stream(),
                         iters, out.get());
      break;
    case 1:
      hipLaunchKernelGGL(mfmaProbeKernel<1>, dim3(workgroups), dim3(kProbeThreads), 0, stream(),
                         iters, out.get());
      break;
    case 2:
      hipLaunchKernelGGL(mfmaProbeKernel<2>, dim3(workgroups), dim3(kProbeThreads), 0, stream(),
                         iters, out.get());
      break;
    case 3:
      hipLaunchKernelGGL(mfmaProbeKernel<3>, dim3(workgroups), dim3(kProbeThreads), 0, stream(),
                         iters, out.get());
      break;
    case 4:
      hipLaunchKernelGGL(mfmaProbeSlimKernel<4>, dim3(workgroups), dim3(kProbeThreads), 0,
                         stream(), iters, out.get());
      break;
    case 5:
      hipLaunchKernelGGL(mfmaProbeSlimKernel<5>, dim3(workgroups), dim3(kProbeThreads), 0,
                         stream(), iters, out.get());
      break;
    default:
      hipLaunchKernelGGL(mfmaProbeKernel<6>, dim3(workgroups), dim3(kProbeThreads), 0, stream(),
                         iters, out.get());
      break;
    }
    s((int)hipGetLastError());
  };
  float ms = 0.0f;
  if (s.ok()) {
    launch(); // warm-up
    s((int)hipEventRecord(start, stream()));
    for (int r = 0; r < kReps && s.ok(); ++r)
      launch();
    s((int)hipEventRecord(stop, stream()));
    s((int)hipEventSynchronize(stop));
    s((int)hipEventElapsedTime(&ms, start, stop));
  }
  if (start)
    (void)hipEventDestroy(start);
  if (stop)
    (void)hipEventDestroy(stop);
  if (!s.ok())
    return s.code;
  *msPerLaunch = ms / kReps;
  const double flopsPerMfma = (shape == 3 || shape == 5) ? 16384.0 : 32768.0;
  const double mfmasPerStep = (shape == 4 || shape == 5) ? 8.0 : 16.0;
  const double directions = shape == 6 ? 2.0 : 1.0; // a symmetric value serves both rows
  *flopsPerLaunch =
      (double)workgroups * (kProbeThreads / 64) * iters * mfmasPerStep * flopsPerMfma * directions;
  return 0;
}
