// Mixed-precision nearest-other-component pass for Boruvka (euclidean and
// squared euclidean): FP16 Gram candidates on MFMA with a certified per-row
// error bound, then an exact FP32 recheck of every candidate. No reference
// counterpart; the result equals the FP32 direct-form pass bit for bit.
//
// Scaled points x' = (x - mu) * s (mu the mean, s a power of two with
// max |x'| <= 2^14) are rounded to FP16: xh. Per point: N = |xh|^2 (FP32),
// R >= |xh|, rho >= |xh - x'| (exact FP16 rounding error, FP64).
//
// MFMA (v_mfma_f32_32x32x16_f16): for a block of 32 columns j and 32 rows i,
//   v_ij = N_j + sum_k xh_jk * (-2 xh_ik)      (C seeded with N_j)
// so S_ij = v_ij + N_i estimates |xh_i - xh_j|^2 with
//   |S_ij - |xh_i - xh_j|^2| <= E_i = (K + 8) 2^-23 (R_i + R_max)^2
// (every FP32 addition may round; K products are exact), and the exact scaled
// distance r = |x'_i - x'_j| = s |x_i - x_j| satisfies | r - |xh_i - xh_j| |
// <= delta_i = rho_i + rho_max. The FP32 direct-form value sigma of the pair
// (the fma fold) is within (d + 4) 2^-24 of the exact squared distance.
// Per row, from m = the kTopK-th smallest half-block minimum of v seen so far
// (kTopK distinct columns have v <= m):
//   A     = sqrt(m + N_i + E_i) + delta_i            >= r of those columns
//   theta = (A sqrt(kappa) + delta_i)^2 + E_i - N_i
// with kappa covering both FP32 roundings and a 2^-21 margin for ties of
// sqrtf. A column with v_ij > theta has a larger FP32 value than the row's
// kTopK-th nearest (or than its sqrt ties), so it is not among the row's
// kTopK nearest; every column with v_ij <= theta is a candidate.
//
// Rows are the points still to be searched this round (in component order);
// columns are all points permuted so that every component is a contiguous
// range, which makes the same-component mask a per-row interval: column tiles
// inside every row's own component are skipped, tiles crossing one are
// masked per value, all others need no mask. Each wave owns 128 rows (4 row
// blocks of 32, the MFMA N dimension); the workgroup's 4 waves share column
// tiles of 128 points staged in LDS (double-buffered, the next tile's global
// loads in flight while the current one is computed). A lane holds 16 of the
// 32 columns of each MFMA block for its row; it records (half-block, minimum)
// in its own list when the minimum is within its theta (the two halves of a
// row share their minimum once per tile, so theta stays tight). The recheck
// kernel recomputes the FP32 direct-form weights of the recorded half-blocks
// still within the row's final theta and keeps the exact kTopK smallest keys
// (weight, original index). The next round takes a row's nearest point in
// another component from that list when one of its entries has left the
// row's component (the entries were the nearest outside a subset of the new
// component, so the first such entry is exact); only rows whose whole list
// joined their component are searched again. Rows whose list overflows are
// rechecked in full, eight rows sharing each read of a 4,096-column chunk. This
// translation unit is compiled with -fno-honor-nans (inputs are validated
// finite), which lets fminf map to v_min/v_min3 without canonicalisation.
#include "mst.hpp"

#include "primitives.hpp"

#include <algorithm>
#include <cfloat>
#include <atomic>
#include <cmath>
#include <numeric>
#include <type_traits>

namespace cuhc {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Columns per staged tile: 256 for DP <= 32 (two workgroups of 256 threads
// still fit a CU's LDS with three stages), 128 for DP = 64.
HD constexpr int tileColsFor(int DP) { return DP <= 32 ? 256 : 128; }
constexpr int kWaveRows = 128;     // 4 row blocks of 32
constexpr int kBlockRows = 4 * kWaveRows;
constexpr int kCandCap = 64; // list entries per (row, half)
constexpr int kTopK = 4; // nearest points kept per row for the next round
constexpr int kDryTilesFirstRound = 16;
constexpr int kDryTilesLaterRounds = 4; // per column split
constexpr int kMaxSplits = 256;         // column splits of a later round

std::atomic<int> gLayoutChecked{0}; // 0 unknown, 1 ok, -1 mismatch
unsigned long long gStats[4];

// D register r of lane l holds row m = (r / 4) * 8 + (l / 32) * 4 + r % 4 and
// column n = l % 32; lane l supplies A[l % 32][8 (l / 32) .. +7] and
// B[8 (l / 32) .. +7][l % 32]. Checked once per process by layoutProbe.
DEV int mOfReg(int r, int lane) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

__global__ void layoutProbeKernel(int *bad) {
  const int lane = threadIdx.x;
  half8 a, b;
  for (int i = 0; i < 8; ++i) {
    const int k = 8 * (lane >> 5) + i;
    a[i] = (_Float16)(k == 0 ? (float)(lane & 31) : (k == 1 ? 1.0f : 0.0f));
    b[i] = (_Float16)(k == 0 ? 1.0f : (k == 1 ? 32.0f * (lane & 31) : 0.0f));
  }
  f32x16 c;
  for (int r = 0; r < 16; ++r)
    c[r] = 0.0f;
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  int wrong = 0;
  for (int r = 0; r < 16; ++r)
    wrong += c[r] != (float)(mOfReg(r, lane) + 32 * (lane & 31));
  if (wrong)
    atomicAdd(bad, wrong);
}

// ---------------------------------------------------------------------------
// Preparation: mean, scale, FP16 points and per-point bounds.

__global__ void columnSumsKernel(const float *X, int n, int d, double *sums) {
  // Each workgroup sums a contiguous chunk of rows; thread t owns feature t % d.
  const int rowsPer = (n + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * rowsPer, r1 = min(n, r0 + rowsPer);
  const int lanes = (blockDim.x / d) * d;
  if ((int)threadIdx.x >= lanes)
    return;
  const int f = threadIdx.x % d, group = threadIdx.x / d, groups = lanes / d;
  double acc = 0.0;
  for (int r = r0 + group; r < r1; r += groups)
    acc += (double)X[(size_t)r * d + f];
  atomicAdd(&sums[f], acc);
}

__global__ void maxDevKernel(const float *X, int n, int d, const float *mu, unsigned int *maxBits) {
  float local = 0.0f;
  const size_t total = (size_t)n * d;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x)
    local = fmaxf(local, fabsf((float)((double)X[e] - (double)mu[e % d])));
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    local = fmaxf(local, __shfl_xor(local, m, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax(maxBits, __float_as_uint(local));
}

__global__ void toHalfKernel(const float *X, int n, int d, int DP, const float *mu, double scale,
                             _Float16 *xh, float *norm, float *R, float *rho,
                             unsigned int *maxR, unsigned int *maxRho) {
  float lr = 0.0f, lrho = 0.0f;
  // A thread converts one point, eight features at a time: the row is read in
  // 16-byte loads when X is 16-byte aligned and d is a multiple of 4 (every
  // row then is), and each eight halves leave in one 16-byte store.
  const bool vec4 = (d & 3) == 0 && ((size_t)X & 15) == 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double nn = 0.0, err = 0.0;
    const float *row = X + (size_t)i * d;
    for (int k0 = 0; k0 < DP; k0 += 8) {
      float v[8];
      if (vec4 && k0 + 8 <= d) {
        const float4 a = *reinterpret_cast<const float4 *>(row + k0);
        const float4 b = *reinterpret_cast<const float4 *>(row + k0 + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q)
          v[q] = k0 + q < d ? row[k0 + q] : 0.0f;
      }
      half8 h8;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + q;
        _Float16 h = (_Float16)0.0f;
        if (k < d) {
          const double xs = ((double)v[q] - (double)mu[k]) * scale;
          h = (_Float16)(float)xs;
          const double hd = (double)(float)h;
          nn += hd * hd;
          err += (hd - xs) * (hd - xs);
        }
        h8[q] = h;
      }
      *reinterpret_cast<half8 *>(xh + (size_t)i * DP + k0) = h8;
    }
    const float nf = (float)nn;
    norm[i] = nf;
    const float r = (float)(sqrt(nn) * (1.0 + 1e-6));
    const float e = (float)(sqrt(err) * (1.0 + 1e-6) + 1e-30);
    R[i] = r;
    rho[i] = e;
    lr = fmaxf(lr, r);
    lrho = fmaxf(lrho, e);
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    lr = fmaxf(lr, __shfl_xor(lr, m, 64));
    lrho = fmaxf(lrho, __shfl_xor(lrho, m, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(maxR, __float_as_uint(lr));
    atomicMax(maxRho, __float_as_uint(lrho));
  }
}

// Position of 16-byte chunk k (of DP / 8) of column c within its column in
// the permuted copy and in LDS: an XOR swizzle that puts the chunk the MFMA
// operand loads read in the same step for 8 or 16 consecutive columns in
// different LDS banks (checked for 32 banks with 8-lane phases and 64 banks
// with 16-lane phases).
HD int swizzleChunk(int c, int k, int chunks) {
  if (chunks == 2)
    return k ^ (((c >> 2) ^ (c >> 3)) & 1);
  if (chunks == 4)
    return k ^ (((c >> 1) ^ (c >> 3)) & 3);
  return k ^ ((c ^ (c >> 3)) & 7);
}

// Norm slots per column tile in normp (the tile's norms, then padding to 256,
// so that each of the four waves stages one 256-byte slice).
constexpr int kNormSlots = 256;

// Permuted copies of the FP16 points (swizzled chunks) and norms; padding
// columns get zero coordinates and +inf norms (v = +inf, never a candidate).
__global__ void permuteHalfKernel(const _Float16 *xh, const float *norm, const int *perm, int n,
                                  int nPad, int DP, int tileCols, _Float16 *xhp, float *normp) {
  const int vecs = DP / 8;
  const size_t total = (size_t)nPad * vecs;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(e / vecs), v = (int)(e - (size_t)p * vecs);
    half8 h;
    if (p < n) {
      h = *reinterpret_cast<const half8 *>(xh + (size_t)perm[p] * DP + 8 * v);
    } else {
      for (int i = 0; i < 8; ++i)
        h[i] = (_Float16)0.0f;
    }
    const int c = p % tileCols;
    *reinterpret_cast<half8 *>(xhp + (size_t)p * DP + 8 * swizzleChunk(c, v, vecs)) = h;
    if (v == 0) {
      const size_t slot = (size_t)(p / tileCols) * kNormSlots + c;
      normp[slot] = p < n ? norm[perm[p]] : INFINITY;
      if (c + tileCols < kNormSlots)
        normp[slot + tileCols] = INFINITY;
    }
  }
}

// ---------------------------------------------------------------------------
// Candidate pass

struct CandArgs {
  const _Float16 *xhp; // nPad x DP, permuted
  const float *normp;  // nPad, permuted
  const _Float16 *xh;  // n x DP, original order (row operands)
  const float *norm, *R, *rho; // per point, original order
  const int *perm;     // position -> original index
  const int *rows;     // numRows x 3: position (-1: padding slot), component start, end
  int numRows, n, nPad;
  float Rmax, rhoMax;
  float gamma;  // (K + 8) 2^-23
  double kappa; // ((1 + eps)(1 + 2^-21) / (1 - eps)) with eps = (d + 4) 2^-24
  uint2 *cand;  // (numRows x 2 halves) x kCandCap: (half-block, block minimum bits); half-block
                // b is lane half b & 1 of the 32-column block b >> 1
  int *candCount; // numRows x 2 (zeroed before a launch with column splits)
  unsigned int *theta;  // numRows, order-preserving bits (thetaBits), minimum over splits
  unsigned int *rowMin; // numRows, thetaBits of the row minimum shared by column splits
  unsigned long long *recorded; // [0] entries recorded, [4] row-column pairs computed
  int rank; // theta from the rank-th smallest half-block minimum (1 or kTopK)
  int dryTiles; // tiles of the dry run before recording starts (at least 1)
};

// Order-preserving 32-bit image of a float (any sign), for atomicMin across
// the workgroups that share a row; 0xffffffff is "none".
HD unsigned int thetaBits(float f) {
  const unsigned int b = bitsOf(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HD float thetaFromBits(unsigned int o) {
  return floatOf((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}
constexpr unsigned int kNoTheta = 0xffffffffu;

__device__ __noinline__ float thetaOf(float m, float Ni, float Ei, float deltai, double sqrtKappa) {
  if (!(m < INFINITY))
    return INFINITY;
  const double A = sqrt(fmax((double)m + (double)Ni + (double)Ei, 0.0)) + (double)deltai;
  const double t = A * sqrtKappa + (double)deltai;
  const double th = t * t + (double)Ei - (double)Ni;
  // Round up to the next float so the FP32 comparison stays conservative.
  const float f = (float)th;
  return (double)f >= th ? f : __uint_as_float(__float_as_uint(f) + (f >= 0.0f ? 1u : -1u));
}

// Operands of one 32-column block: A (column points) and the C seed (their norms).
template <int KB> struct BlockOps {
  half8 a[KB];
  f32x16 seed;
};

template <int KB>
DEV void loadBlock(BlockOps<KB> &o, const _Float16 *sx, const float *sn, int cb, int l32, int hi) {
  constexpr int DP = 16 * KB;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 s4 = *reinterpret_cast<const float4 *>(&sn[cb * 32 + g * 8 + hi * 4]);
    o.seed[4 * g + 0] = s4.x;
    o.seed[4 * g + 1] = s4.y;
    o.seed[4 * g + 2] = s4.z;
    o.seed[4 * g + 3] = s4.w;
  }
  const int c = cb * 32 + l32;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb)
    o.a[kb] = *reinterpret_cast<const half8 *>(
        &sx[c * DP + 8 * swizzleChunk(c, kb * 2 + hi, DP / 8)]);
}

// Per-lane candidate state of one row block kept in registers (this lane's
// half of the row's columns). The rest of the row's state (component range,
// norm and bounds, row minimum) is read rarely and lives in LDS (RowShared).
struct RowLane {
  float top[kTopK]; // ascending: this lane's smallest half-block minima (see tile end)
  float th;     // theta of the row minimum at the last tile end
  int count;    // entries recorded in this lane's list
};

// Struct-of-arrays per (wave, row block, lane): consecutive lanes, consecutive
// banks.
struct RowShared {
  int cs[4][4][64], ce[4][4][64]; // the row's component: permuted positions [cs, ce)
  float N[4][4][64], E[4][4][64], D[4][4][64]; // norm, accumulation bound, rounding radius
  float rowMin[4][4][64];                      // the row minimum over both halves
};

// v for one row block: C seeded with the column norms, A the column points,
// B this row block's -2 xh.
template <int KB>
DEV f32x16 mfmaBlock(const BlockOps<KB> &o, const half8 (&bop)[KB]) {
  f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.a[0], bop[0], o.seed, 0, 0, 0);
#pragma unroll
  for (int kb = 1; kb < KB; ++kb)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.a[kb], bop[kb], acc, 0, 0, 0);
  return acc;
}

// This lane's minimum over its 16 values of one block (same-component columns
// masked to +inf when the tile crosses the row's component).
DEV float blockMin(f32x16 acc, bool anyMask, int j0, int lane, int cs, int ce) {
  if (anyMask) {
    // Column of register r: j0 + mOfReg(r, lane) = x + (r / 4) * 8 + r % 4.
    const unsigned span = (unsigned)(ce - cs);
    const int x = j0 + (lane >> 5) * 4 - cs;
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if ((unsigned)(x + (r >> 2) * 8 + (r & 3)) < span)
        acc[r] = INFINITY;
  }
  // A tree of three-way minima.
  const float m0 = fminf(fminf(acc[0], acc[1]), acc[2]);
  const float m1 = fminf(fminf(acc[3], acc[4]), acc[5]);
  const float m2 = fminf(fminf(acc[6], acc[7]), acc[8]);
  const float m3 = fminf(fminf(acc[9], acc[10]), acc[11]);
  const float m4 = fminf(fminf(acc[12], acc[13]), acc[14]);
  return fminf(fminf(fminf(m0, m1), m2), fminf(fminf(m3, m4), acc[15]));
}

// Sorted insertion into the lane's kTopK smallest half-block minima.
DEV void insertTop(RowLane &rl, float m) {
#pragma unroll
  for (int q = kTopK - 1; q > 0; --q)
    rl.top[q] = fminf(rl.top[q], fmaxf(rl.top[q - 1], m));
  rl.top[0] = fminf(rl.top[0], m);
}

// One list per (row, half). With column splits several workgroups append to
// it, so the position comes from an atomic counter; otherwise from the lane's
// own count (stored at the end).
DEV void recordBlock(RowLane &rl, int slot, float m, int block, const CandArgs &a, int hi,
                     bool shared) {
  const size_t list = (size_t)slot * 2 + hi;
  const int pos = shared ? atomicAdd(&a.candCount[list], 1) : rl.count;
  if (pos < kCandCap)
    a.cand[list * kCandCap + pos] = make_uint2((unsigned)block, __float_as_uint(m));
  ++rl.count;
}

// Records this lane's 16 columns of each of two blocks when they can hold a
// candidate (the lane's theta is never below the row's final one, so nothing
// needed is missed), and inserts the two minima into the lane's smallest
// ones. th is -inf for rows outside the range, so one comparison also drops
// wholly masked blocks. A minimum above th is never among the row's kTopK
// smallest, so skipping its insertion only keeps theta conservative.
DEV void recordPair(RowLane &rl, int slot, float mA, float mB, int blockA, const CandArgs &a,
                    int hi, bool shared, bool dry) {
  if (fminf(mA, mB) <= rl.th) {
    insertTop(rl, mA);
    insertTop(rl, mB);
    if (!dry && mA <= rl.th)
      recordBlock(rl, slot, mA, blockA, a, hi, shared);
    if (!dry && mB <= rl.th)
      recordBlock(rl, slot, mB, blockA + 2, a, hi, shared);
  }
}

// s_waitcnt with vmcnt = N and the other counters left unconstrained (gfx9
// encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N> DEV void waitVmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// LDS-direct loads (global_load_lds): each lane's 16 (or 4) bytes at g land
// in LDS at lds + lane * 16 (or * 4); lds is wave-uniform. Written as inline
// assembly so that the compiler, which cannot tell which LDS bytes these loads
// write, does not add a full vmcnt(0) wait before every LDS read; the kernel
// waits for them itself (waitVmcnt, then a barrier). M0 is set in the same
// statement and holds nothing else in this kernel.
DEV unsigned ldsOffset(const void *lds) {
  return __builtin_amdgcn_readfirstlane(
      (unsigned)(size_t)(__attribute__((address_space(3))) const char *)lds);
}
DEV void ldsDirectLoad16(const void *g, const void *lds) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
               :
               : "s"(ldsOffset(lds)), "v"(g)
               : "memory");
}
DEV void ldsDirectLoad4(const void *g, const void *lds) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off"
               :
               : "s"(ldsOffset(lds)), "v"(g)
               : "memory");
}

// SPLIT: several workgroups (gridDim.y column splits) share each row block.
// A separate instantiation keeps the atomic list position, whose return the
// wave must wait for, out of the unsplit kernel.
template <int KB, bool SPLIT>
__global__ void __launch_bounds__(256, 2) mfmaCandidateKernel(CandArgs a) {
  constexpr int DP = 16 * KB;
  constexpr int kTileCols = tileColsFor(DP);
  constexpr int TILE_BYTES = kTileCols * DP * 2;
  constexpr int LOADS = TILE_BYTES / (256 * 16); // 16-byte LDS-direct loads per thread per tile
  static_assert(TILE_BYTES % (256 * 16) == 0, "tile staging");
  constexpr int STAGES = 3;
  __shared__ __align__(16) _Float16 sx[STAGES][kTileCols * DP];
  __shared__ __align__(16) float sn[STAGES][kNormSlots];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hi = lane >> 5, l32 = lane & 31;
  const int rowBase = blockIdx.x * kBlockRows + wave * kWaveRows;
  const double sqrtKappa = sqrt(a.kappa);

  // Row operands (-2 xh_i) and per-row state. Row block rb of this wave is
  // rows rowBase + rb * 32 + l32 (slot); slots at or past numRows and slots
  // with position -1 are padding (zero operands, th = -inf: nothing recorded,
  // component start -1 in LDS).
  __shared__ RowShared rs;
  int myValid = 0; // valid rows among this lane's (hi == 0 lanes count each row once)
  half8 bop[4][KB];
  RowLane rl[4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
    const int r = rowBase + rb * 32 + l32;
    const int pos = r < a.numRows ? a.rows[3 * r] : -1;
    const bool valid = pos >= 0;
    int orig = 0, cs = -1, ce = 0;
    if (valid) {
      cs = a.rows[3 * r + 1];
      ce = a.rows[3 * r + 2];
      orig = a.perm[pos];
    }
    myValid += valid && hi == 0;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      half8 h;
      if (valid) {
        h = *reinterpret_cast<const half8 *>(a.xh + (size_t)orig * DP + kb * 16 + 8 * hi);
        h = h * (_Float16)(-2.0f);
      } else {
        for (int i = 0; i < 8; ++i)
          h[i] = (_Float16)0.0f;
      }
      bop[rb][kb] = h;
    }
    rs.cs[wave][rb][lane] = cs;
    rs.ce[wave][rb][lane] = ce;
    rs.N[wave][rb][lane] = valid ? a.norm[orig] : 0.0f;
    const float Rs = (valid ? a.R[orig] : 0.0f) + a.Rmax;
    rs.E[wave][rb][lane] = a.gamma * Rs * Rs * (1.0f + 1e-6f);
    rs.D[wave][rb][lane] = (valid ? a.rho[orig] : 0.0f) + a.rhoMax;
    rs.rowMin[wave][rb][lane] = INFINITY;
#pragma unroll
    for (int q = 0; q < kTopK; ++q)
      rl[rb].top[q] = INFINITY;
    rl[rb].th = valid ? FLT_MAX : -INFINITY;
    rl[rb].count = 0;
  }

  // Column tiles stream into LDS by direct (LDS-DMA) loads, three buffers
  // deep: tile t + 2 is requested while tile t is computed. Each wave copies
  // LOADS consecutive 1 KB slices of the tile and one 256-byte slice of its
  // norm slots; the slices land in LDS in global order, and the global copy is
  // already swizzled for the operand loads.
  const char *gsrc = reinterpret_cast<const char *>(a.xhp);
  const char *gnorm = reinterpret_cast<const char *>(a.normp);
  auto request = [&](int t, int stage) {
    char *dstTile = reinterpret_cast<char *>(sx[stage]);
#pragma unroll
    for (int j = 0; j < LOADS; ++j) {
      const int slice = j * 4 + wave;
      ldsDirectLoad16(gsrc + (size_t)t * TILE_BYTES + slice * 1024 + lane * 16,
                      dstTile + slice * 1024);
    }
    ldsDirectLoad4(gnorm + ((size_t)t * kNormSlots + wave * 64 + lane) * 4,
                   reinterpret_cast<char *>(sn[stage]) + wave * 256);
  };
  const int numTiles = a.nPad / kTileCols;
  int validRows = myValid;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    validRows += __shfl_xor(validRows, m, 64);
  unsigned long long computedTiles = 0;
  // Column tiles [tA, tB) lie inside the component of every row of the
  // workgroup (rows are listed in component order, so late rounds with large
  // components skip most tiles): they are neither loaded nor visited.
  __shared__ int blockMaxCs, blockMinCe;
  if (threadIdx.x == 0) {
    blockMaxCs = 0;
    blockMinCe = 0x7fffffff;
  }
  __syncthreads();
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
    if (rs.cs[wave][rb][lane] >= 0) {
      atomicMax(&blockMaxCs, rs.cs[wave][rb][lane]);
      atomicMin(&blockMinCe, rs.ce[wave][rb][lane]);
    }
  __syncthreads();
  const int tA = (blockMaxCs + kTileCols - 1) / kTileCols, tB = blockMinCe / kTileCols;
  const int skipped = tB > tA ? tB - tA : 0;
  const int allIters = numTiles - skipped;
  // Column split blockIdx.y of gridDim.y takes an equal share of the tiles
  // (several workgroups per row block when rows are few).
  const int splits = gridDim.y;
  constexpr bool shared = SPLIT;
  const int it0 = (int)((long long)allIters * blockIdx.y / splits);
  const int iters = (int)((long long)allIters * (blockIdx.y + 1) / splits) - it0;
  auto tileOf = [&](int i) {
    const int j = it0 + i;
    return skipped && j >= tA ? j + skipped : j;
  };
  // Steps [0, dryTiles) are a dry run over the first dryTiles tiles: they set
  // every row's theta from those tiles' smallest half-block minima without
  // recording anything, so recording starts against a theta near the row's
  // final one (a streaming top-kTopK records about kTopK ln(columns seen /
  // columns of the dry run) times while theta tightens). The remaining steps
  // visit every tile, the dry run's again.
  const int dryTiles = iters > 0 ? min(a.dryTiles, iters) : 0;
  const int totalSteps = dryTiles + iters;
  auto tileOfStep = [&](int j) { return tileOf(j < dryTiles ? j : j - dryTiles); };
  if (totalSteps > 0) {
    request(tileOfStep(0), 0);
    if (totalSteps > 1)
      request(tileOfStep(1), 1);
  }
  for (int j = 0; j < totalSteps; ++j) {
    const bool dry = j < dryTiles;
    const bool lastDry = j + 1 == dryTiles;
    const int ii = dry ? j : j - dryTiles;
    const int t = tileOfStep(j), buf = j % STAGES;
    // Tile t's loads are this wave's oldest outstanding ones; only the next
    // tile's (LOADS + 1) may remain (memory operations complete in order).
    if (j + 1 < totalSteps)
      waitVmcnt<LOADS + 1>();
    else
      waitVmcnt<0>();
    // Every wave's slices of tile t are in LDS, and every wave is done with
    // the previous tile, whose buffer the tile after next now reuses.
    __syncthreads();
    if (j + 2 < totalSteps)
      request(tileOfStep(j + 2), (j + 2) % STAGES);
    const int c0 = t * kTileCols, c1 = c0 + kTileCols;
    // Wave-uniform decisions about the same-component mask.
    bool inter = false, contained = true;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int cs = rs.cs[wave][rb][lane], ce = rs.ce[wave][rb][lane];
      if (cs < 0)
        continue; // padding slot
      inter |= ce > c0 && cs < c1;
      contained &= cs <= c0 && ce >= c1;
    }
    const bool skipAll = __ballot(!contained) == 0;
    const bool anyMask = __ballot(inter) != 0;
    if (!skipAll) {
      computedTiles += !dry;
      // The masked and the mask-free tile are separate code paths, so the
      // inner loop carries no per-block mask branch.
      auto runTile = [&](auto maskTag) {
        constexpr bool kMask = decltype(maskTag)::value;
        BlockOps<KB> o0, o1;
        loadBlock<KB>(o0, sx[buf], sn[buf], 0, l32, hi);
#pragma unroll 1
        for (int cb = 0; cb < kTileCols / 32; cb += 2) {
          // All four row blocks' MFMA chains are issued before any epilogue
          // reads a result, so they overlap in the matrix pipeline.
          f32x16 acc[4];
          loadBlock<KB>(o1, sx[buf], sn[buf], cb + 1, l32, hi);
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
            acc[rb] = mfmaBlock<KB>(o0, bop[rb]);
          float mA[4];
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
            mA[rb] = blockMin(acc[rb], kMask, c0 + cb * 32, lane,
                              kMask ? rs.cs[wave][rb][lane] : 0,
                              kMask ? rs.ce[wave][rb][lane] : 0);
          if (cb + 2 < kTileCols / 32)
            loadBlock<KB>(o0, sx[buf], sn[buf], cb + 2, l32, hi);
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
            acc[rb] = mfmaBlock<KB>(o1, bop[rb]);
          float mB[4];
          bool record = false;
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) {
            mB[rb] = blockMin(acc[rb], kMask, c0 + (cb + 1) * 32, lane,
                              kMask ? rs.cs[wave][rb][lane] : 0,
                              kMask ? rs.ce[wave][rb][lane] : 0);
            record |= fminf(mA[rb], mB[rb]) <= rl[rb].th;
          }
          // One wave-uniform branch for the four row blocks: in the common
          // case no lane records anything. The dry run only keeps minima
          // (theta is FLT_MAX until its first merge). (The minima are updated
          // in this one branch only: a second update path would make the
          // compiler copy them between registers every iteration.)
          if (__ballot(record) != 0) {
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
              recordPair(rl[rb], rowBase + rb * 32 + l32, mA[rb], mB[rb],
                         (t * (kTileCols / 32) + cb) * 2 + hi, a, hi, shared, dry);
          }
        }
      };
      if (anyMask)
        runTile(std::true_type{});
      else
        runTile(std::false_type{});
      // Tile end (every fourth tile, the end of the dry run and the last
      // tile): both halves of each row agree on the row's kTopK smallest
      // half-block minima and theta. The merged list stays with the lane of
      // half 0 and half 1 starts empty, so no minimum is counted twice; after
      // the dry run both start empty (its tiles are visited again). Between
      // merges theta stays that of the last merge, only looser: more blocks
      // are recorded, none is missed. At 10^6 points, d = 32, merging every
      // tile, second, fourth and eighth tile takes the first round 72.6,
      // 70.0, 68.2 and 69.9 ms, with 0, 0, about 20 and about 680 rows whose
      // lists overflow (each then scans every column).
      const bool mergeNow = lastDry || (ii & 3) == 3 || (!dry && ii + 1 == iters);
#pragma unroll
      for (int rb = 0; rb < 4 && mergeNow; ++rb) {
        float c[kTopK];
#pragma unroll
        for (int q = 0; q < kTopK; ++q)
          c[q] = fminf(rl[rb].top[q], __shfl_xor(rl[rb].top[kTopK - 1 - q], 32, 64));
        // c is bitonic (ascending against descending); sort it.
        static_assert(kTopK == 4, "bitonic merge of four");
        auto cx = [](float &x, float &y) {
          const float lo = fminf(x, y), hi2 = fmaxf(x, y);
          x = lo;
          y = hi2;
        };
        cx(c[0], c[2]);
        cx(c[1], c[3]);
        cx(c[0], c[1]);
        cx(c[2], c[3]);
        const bool keep = !lastDry && hi == 0;
#pragma unroll
        for (int q = 0; q < kTopK; ++q)
          rl[rb].top[q] = keep ? c[q] : INFINITY;
        float m = a.rank == kTopK ? c[kTopK - 1] : c[0];
        const bool improved = m < rs.rowMin[wave][rb][lane] && rs.cs[wave][rb][lane] >= 0;
        if (shared) {
          // Publish an improvement; take the best any split has published.
          unsigned int old = kNoTheta;
          if (improved && hi == 0)
            old = atomicMin(&a.rowMin[rowBase + rb * 32 + l32], thetaBits(m));
          old = __shfl(old, l32, 64);
          if (old != kNoTheta)
            m = fminf(m, thetaFromBits(old));
        }
        if (improved) {
          rs.rowMin[wave][rb][lane] = m;
          rl[rb].th = thetaOf(m, rs.N[wave][rb][lane], rs.E[wave][rb][lane],
                              rs.D[wave][rb][lane], sqrtKappa);
        }
      }
    }
  }
  unsigned long long rec = 0;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
    const int slot = rowBase + rb * 32 + l32;
    if (rs.cs[wave][rb][lane] >= 0) {
      if (!shared)
        a.candCount[(size_t)slot * 2 + hi] = rl[rb].count;
      if (hi == 0)
        atomicMin(&a.theta[slot], thetaBits(rl[rb].th));
      rec += rl[rb].count;
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
    rec += __shfl_xor(rec, m, 64);
  if (lane == 0 && rec)
    atomicAdd(a.recorded, rec);
  if (lane == 0 && computedTiles && validRows)
    atomicAdd(a.recorded + 4, computedTiles * validRows * kTileCols);
}

// ---------------------------------------------------------------------------
// Exact FP32 recheck: one wave per row. The row's list entries (kCandCap per
// half) are read 32 per half at a time, one per lane; the entries still within
// the row's final theta are taken four at a time, 16 lanes per half-block, one
// column per lane. Column points come from the permuted FP32 copy xp (the same
// values as X), so no load waits on the permutation.

struct RecheckArgs {
  const float *X;  // n x d, original order
  const float *xp; // n x d, permuted like the columns
  int n, d;
  const int *perm;
  const int *rows;
  int numRows;
  const uint2 *cand;
  const int *candCount;
  const unsigned int *theta; // thetaBits
  bool sqeuclidean;
  unsigned long long *nearest; // by original index
  unsigned long long *topk;    // by original index: kTopK ascending keys (kEmptyKey padded)
  int rank;                    // keys certified per row (the candidate pass's rank)
  // kStatStripes stripes of one 128-byte line each: [1] half-blocks rechecked,
  // [2] pairs, [3] overflow rows (striped so that a million waves do not
  // serialise on one address).
  unsigned long long *stats;
  int *overflowRows;  // rows whose list overflowed, appended at overflowCount
  int *overflowCount; // zero before the recheck
};

// Columns one workgroup scans for the rows whose lists overflowed.
constexpr int kOverflowWaveCols = 4096;

constexpr int kStatStripes = 64;
constexpr int kStripeWords = 16;

// FP32 direct-form weight of (row in LDS, column position col): the plain fold.
DEV unsigned long long recheckColumn(const RecheckArgs &a, const float *xi, int col, int cs,
                                     int ce) {
  if (col >= a.n || (col >= cs && col < ce))
    return kEmptyKey;
  const int j = a.perm[col];
  const float *xj = a.xp + (size_t)col * a.d;
  float acc = 0.0f;
  if ((a.d & 3) == 0) {
    // Every load is issued before the fold uses it (d <= 64).
    float4 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (4 * q < a.d)
        v[q] = *reinterpret_cast<const float4 *>(xj + 4 * q);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (4 * q < a.d) {
        const int k = 4 * q;
        float diff = xi[k] - v[q].x;
        acc = fmaf(diff, diff, acc);
        diff = xi[k + 1] - v[q].y;
        acc = fmaf(diff, diff, acc);
        diff = xi[k + 2] - v[q].z;
        acc = fmaf(diff, diff, acc);
        diff = xi[k + 3] - v[q].w;
        acc = fmaf(diff, diff, acc);
      }
    }
  } else {
    for (int k = 0; k < a.d; ++k) {
      const float diff = xi[k] - xj[k];
      acc = fmaf(diff, diff, acc);
    }
  }
  const float w = a.sqeuclidean ? acc : sqrtf(acc);
  return w != INFINITY ? packKey(w, (unsigned)j) : kEmptyKey;
}

// Sorted insertion of key into this lane's kTopK smallest keys (ascending).
DEV void insertKey(unsigned long long (&best)[kTopK], unsigned long long key) {
#pragma unroll
  for (int q = kTopK - 1; q > 0; --q)
    best[q] = key < best[q] ? (key > best[q - 1] ? key : best[q - 1]) : best[q];
  best[0] = minKey(best[0], key);
}

// The wave's kTopK smallest keys (every lane gets them): repeatedly the
// minimum of the lanes' heads, popped from the lane that holds it (keys are
// distinct: distinct indices).
DEV void waveTopK(unsigned long long (&best)[kTopK], unsigned long long (&top)[kTopK]) {
#pragma unroll
  for (int q = 0; q < kTopK; ++q) {
    unsigned long long m = best[0];
#pragma unroll
    for (int s2 = 32; s2 > 0; s2 >>= 1)
      m = minKey(m, shflXorKey(m, s2));
    top[q] = m;
    if (m != kEmptyKey && best[0] == m) {
#pragma unroll
      for (int r2 = 0; r2 + 1 < kTopK; ++r2)
        best[r2] = best[r2 + 1];
      best[kTopK - 1] = kEmptyKey;
    }
  }
}

// A row's result: its first rank keys (every column that could precede them
// was checked) and its nearest.
DEV void writeRowKeys(const RecheckArgs &a, int pos, const unsigned long long (&top)[kTopK],
                      int lane) {
  if (lane < kTopK)
    a.topk[(size_t)a.perm[pos] * kTopK + lane] = lane >= a.rank ? kEmptyKey
                                                 : lane == 0   ? top[0]
                                                 : lane == 1   ? top[1]
                                                 : lane == 2   ? top[2]
                                                               : top[3];
  if (lane == 0)
    a.nearest[a.perm[pos]] = top[0];
}

__global__ void __launch_bounds__(256) recheckKernel(RecheckArgs a) {
  static_assert(kCandCap % 32 == 0, "lists are read 32 entries per half at a time");
  __shared__ float rowX[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = blockIdx.x * 4 + wave;
  if (r >= a.numRows || a.rows[3 * r] < 0)
    return; // wave-uniform (padding slot or past the end); no block-wide barrier below
  // Independent loads: the row, its counts, theta and the first 32 list slots
  // of each half (slots past a count hold stale data and are masked below).
  const int pos = a.rows[3 * r], cs = a.rows[3 * r + 1], ce = a.rows[3 * r + 2];
  const int c0 = a.candCount[(size_t)r * 2], c1 = a.candCount[(size_t)r * 2 + 1];
  if (c0 > kCandCap || c1 > kCandCap) {
    // The list lost entries: the row is scanned in full by the overflow
    // kernels (wave-uniform exit).
    if (lane == 0)
      a.overflowRows[atomicAdd(a.overflowCount, 1)] = r;
    return;
  }
  const float th = thetaFromBits(a.theta[r]);
  const int half = lane >> 5, e32 = lane & 31;
  const size_t listBase = ((size_t)r * 2 + half) * kCandCap;
  uint2 entry = a.cand[listBase + e32];
  float *xi = rowX[wave];
  for (int k = lane; k < a.d; k += 64)
    xi[k] = a.xp[(size_t)pos * a.d + k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // This lane's kTopK smallest keys (ascending), merged over the wave below.
  unsigned long long best[kTopK];
#pragma unroll
  for (int q = 0; q < kTopK; ++q)
    best[q] = kEmptyKey;
  unsigned long long blocks = 0;
  {
    const int sub = lane >> 4, idx = lane & 15;
    const int count = half ? c1 : c0, longest = max(c0, c1);
    // 32 entries of each half at a time.
#pragma unroll 1
    for (int w0 = 0; w0 < longest; w0 += 32) {
      if (w0 > 0)
        entry = a.cand[listBase + w0 + e32];
      const bool listed = w0 + e32 < count;
      unsigned long long live = __ballot(listed && __uint_as_float(entry.y) <= th);
      blocks += (unsigned long long)__popcll(live);
      while (live) {
        // The next four live entries (wave-uniform); this lane's group takes one.
        int pick[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          pick[g] = live ? __ffsll((long long)live) - 1 : -1;
          live &= live ? live - 1 : 0ull;
        }
        const int e = sub == 0 ? pick[0] : sub == 1 ? pick[1] : sub == 2 ? pick[2] : pick[3];
        const int blk = (int)__shfl(entry.x, e < 0 ? 0 : e, 64);
        if (e >= 0) {
          // Half-block b (32-column block b >> 1, lane half b & 1) covers
          // columns (b >> 1) * 32 + g * 8 + (b & 1) * 4 + q, g, q in 0..3.
          const int col = (blk >> 1) * 32 + (idx >> 2) * 8 + (blk & 1) * 4 + (idx & 3);
          insertKey(best, recheckColumn(a, xi, col, cs, ce));
        }
      }
    }
  }
  unsigned long long top[kTopK];
  waveTopK(best, top);
  writeRowKeys(a, pos, top, lane);
  if (lane == 0 && blocks) {
    unsigned long long *st = a.stats + (size_t)(r % kStatStripes) * kStripeWords;
    atomicAdd(&st[1], blocks);
    atomicAdd(&st[2], blocks * 16);
  }
}

// Rows whose list overflowed, kOverflowRowsPerGroup per workgroup (two per
// wave) in the y dimension: workgroup x scans columns [x * kOverflowWaveCols,
// (x + 1) * kOverflowWaveCols), staged 64 at a time in LDS so that every
// column is read once for all of the workgroup's rows, and leaves each row's
// kTopK smallest keys in partial[(row * gridDim.x + x) * kTopK ..] (row: the
// row's place in this launch). The fold is recheckColumn's, feature by
// feature, on the same values.
constexpr int kOverflowRowsPerGroup = 8;
__global__ void __launch_bounds__(256) overflowScanKernel(RecheckArgs a, int first, int rowsNow,
                                                         unsigned long long *partial) {
  __shared__ float rowX[kOverflowRowsPerGroup][64];
  __shared__ float colX[64][65]; // [feature][column], padded: lane c reads column c
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int local0 = blockIdx.y * kOverflowRowsPerGroup; // this launch's rows
  const int rows = min(kOverflowRowsPerGroup, rowsNow - local0);
  const int d = a.d;
  for (int e = threadIdx.x; e < kOverflowRowsPerGroup * 64; e += 256) {
    const int rr = e >> 6, k = e & 63;
    if (rr < rows && k < d)
      rowX[rr][k] = a.xp[(size_t)a.rows[3 * a.overflowRows[first + local0 + rr]] * d + k];
  }
  // This wave's two rows.
  int cs[2], ce[2];
  bool valid[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int rr = 2 * wave + q;
    valid[q] = rr < rows;
    const int r = valid[q] ? a.overflowRows[first + local0 + rr] : 0;
    cs[q] = valid[q] ? a.rows[3 * r + 1] : 0;
    ce[q] = valid[q] ? a.rows[3 * r + 2] : 0;
  }
  unsigned long long best[2][kTopK];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int t = 0; t < kTopK; ++t)
      best[q][t] = kEmptyKey;
  const long long c0 = (long long)blockIdx.x * kOverflowWaveCols;
  const int c1 = (int)min((long long)a.n, c0 + kOverflowWaveCols);
  for (int s0 = (int)c0; s0 < c1; s0 += 64) {
    __syncthreads(); // rowX written, or the previous stage consumed
    for (int e = threadIdx.x; e < 64 * d; e += 256) {
      const int c = e / d, k = e - c * d;
      if (s0 + c < c1)
        colX[k][c] = a.xp[(size_t)(s0 + c) * d + k];
    }
    __syncthreads();
    const int col = s0 + lane;
    if (col < c1) {
      const int j = a.perm[col];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (!valid[q] || (col >= cs[q] && col < ce[q]))
          continue;
        const float *xi = rowX[2 * wave + q];
        float acc = 0.0f;
        for (int k = 0; k < d; ++k) {
          const float diff = xi[k] - colX[k][lane];
          acc = fmaf(diff, diff, acc);
        }
        const float w = a.sqeuclidean ? acc : sqrtf(acc);
        if (w != INFINITY)
          insertKey(best[q], packKey(w, (unsigned)j));
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    unsigned long long top[kTopK];
    waveTopK(best[q], top);
    if (valid[q] && lane < kTopK)
      partial[((size_t)(local0 + 2 * wave + q) * gridDim.x + blockIdx.x) * kTopK + lane] =
          lane == 0 ? top[0] : lane == 1 ? top[1] : lane == 2 ? top[2] : top[3];
  }
}

// One wave per overflowed row: the kTopK smallest of its column chunks' keys.
__global__ void __launch_bounds__(64) overflowMergeKernel(RecheckArgs a, int first,
                                                         const unsigned long long *partial,
                                                         int chunks) {
  const int lane = threadIdx.x, o = blockIdx.x, r = a.overflowRows[first + o];
  const int pos = a.rows[3 * r];
  unsigned long long best[kTopK];
#pragma unroll
  for (int q = 0; q < kTopK; ++q)
    best[q] = kEmptyKey;
  const unsigned long long *keys = partial + (size_t)o * chunks * kTopK;
  for (int i = lane; i < chunks * kTopK; i += 64)
    insertKey(best, keys[i]);
  unsigned long long top[kTopK];
  waveTopK(best, top);
  writeRowKeys(a, pos, top, lane);
  if (lane == 0) {
    unsigned long long *st = a.stats + (size_t)(r % kStatStripes) * kStripeWords;
    atomicAdd(&st[3], 1ull);
    atomicAdd(&st[2], (unsigned long long)a.n);
  }
}

// Device preparation of a round: positions 0..n-1, each component's range in
// the sorted order, and the list of rows (position in the sorted order and the
// row's component range) to search.
__global__ void positionsKernel(int *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = i;
}

__global__ void componentRangesKernel(const int *__restrict__ sortedComp, int n, int *start,
                                      int *end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = sortedComp[i];
  if (i == 0 || sortedComp[i - 1] != c)
    start[c] = i;
  if (i == n - 1 || sortedComp[i + 1] != c)
    end[c] = i + 1;
}

// (size << 32 | component) of the largest component.
__global__ void largestComponentKernel(const int *__restrict__ start, const int *__restrict__ end,
                                       int numComp, unsigned long long *largest) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long key = 0;
  if (c < numComp)
    key = ((unsigned long long)(end[c] - start[c]) << 32) | (unsigned int)c;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    const unsigned long long other = shflXorKey(key, m);
    key = other > key ? other : key;
  }
  if ((threadIdx.x & 63) == 0 && key != 0)
    atomicMax(largest, key);
}

// A row is settled when an entry of its kept list (the kTopK nearest outside
// its component when it was last searched) lies in another component now: the
// first such entry is its nearest there, written to nearest[]. Every other
// row is searched again, except the rows of a component holding more than
// half of the points: that component proposes no edge this round (its
// nearest[] entries are emptied). Boruvka only needs each proposed edge to be
// its component's smallest outgoing one, and every other component still
// proposes its own (the large component joins through theirs); its rows'
// kept lists stay valid, as they hold the nearest points outside a subset of
// the component.
__global__ void staleRowsKernel(const int *__restrict__ perm, const int *__restrict__ comp,
                                const unsigned long long *__restrict__ topk,
                                unsigned long long *__restrict__ nearest, int n, bool firstRound,
                                const unsigned long long *__restrict__ largest, int *flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n)
    return;
  if (i == n) {
    flag[n] = 0;
    return;
  }
  const int p = perm[i];
  bool stale = true;
  if (!firstRound) {
    const int own = comp[p];
    const unsigned long long big = *largest;
    if (2 * (long long)(big >> 32) > (long long)n && own == (int)(big & 0xffffffffu)) {
      nearest[p] = kEmptyKey;
      flag[i] = 0;
      return;
    }
#pragma unroll
    for (int q = 0; q < kTopK && stale; ++q) {
      const unsigned long long key = topk[(size_t)p * kTopK + q];
      if (key == kEmptyKey)
        break;
      if (comp[(int)keyIndex(key)] != own) {
        nearest[p] = key;
        stale = false;
      }
    }
  }
  flag[i] = stale ? 1 : 0;
}

// Row slots: a component with at least a workgroup's worth of stale rows gets
// whole workgroups of its own (its row count rounded up, padding slots marked
// by position -1), so that its workgroups can skip its own column range; the
// other components' rows are packed after them.
__global__ void componentSlotsKernel(const int *__restrict__ start, const int *__restrict__ end,
                                     const int *__restrict__ index, int numComp, int *bigSlots,
                                     int *smallSlots) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > numComp)
    return;
  if (c == numComp) {
    bigSlots[c] = 0;
    smallSlots[c] = 0;
    return;
  }
  const int count = index[end[c]] - index[start[c]];
  const bool big = count >= kBlockRows;
  bigSlots[c] = big ? (count + kBlockRows - 1) / kBlockRows * kBlockRows : 0;
  smallSlots[c] = big ? 0 : count;
}

__global__ void rowListKernel(const int *__restrict__ perm, const int *__restrict__ comp,
                              const int *__restrict__ flag, const int *__restrict__ index,
                              const int *__restrict__ start, const int *__restrict__ end,
                              const int *__restrict__ bigSlots, const int *__restrict__ bigBase,
                              const int *__restrict__ smallBase, int numComp, int n, int *rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i])
    return;
  const int c = comp[perm[i]];
  const int within = index[i] - index[start[c]];
  const int r = bigSlots[c] ? bigBase[c] + within : bigBase[numComp] + smallBase[c] + within;
  rows[3 * r] = i;
  rows[3 * r + 1] = start[c];
  rows[3 * r + 2] = end[c];
}

int checkLayout() {
  int state = gLayoutChecked.load();
  if (state != 0)
    return state > 0 ? 0 : (int)hipErrorInvalidDeviceFunction;
  DevArray<int> bad;
  Status s;
  s(bad.alloc(1));
  s(fill(bad, 0, sizeof(int)));
  if (s.ok()) {
    hipLaunchKernelGGL(layoutProbeKernel, dim3(1), dim3(64), 0, stream(), bad.get());
    s((int)hipGetLastError());
  }
  int wrong = 0;
  s(download(&wrong, bad, sizeof(int)));
  if (!s.ok())
    return s.code;
  gLayoutChecked.store(wrong == 0 ? 1 : -1);
  return wrong == 0 ? 0 : (int)hipErrorInvalidDeviceFunction;
}

} // namespace

bool mixedPassSupports(int d) { return d >= 1 && d <= 64; }

struct MixedPass {
  int n = 0, d = 0, DP = 0, nPad = 0, tileCols = 0;
  const float *X = nullptr;
  DevArray<_Float16> xh, xhp;
  DevArray<float> norm, normp, R, rho, mu, xp;
  DevArray<unsigned int> theta, rowMin;
  DevArray<int> perm, rows, candCount;
  DevArray<int> positions, sortedComp, compStart, compEnd, rowFlag, rowIndex;
  DevArray<int> bigSlots, smallSlots, bigBase, smallBase;
  DevArray<unsigned char> sortTemp, scanTemp;
  size_t sortBytes = 0, scanBytes = 0;
  DevArray<uint2> cand;
  DevArray<unsigned long long> stats, recheckStats, topk, largest;
  DevArray<int> overflowRows, overflowCount;
  float Rmax = 0.0f, rhoMax = 0.0f;
};

MixedPass *mixedPassCreate(const float *X, int n, int d, int *status) {
  *status = checkLayout();
  if (*status)
    return nullptr;
  MixedPass *mp = new MixedPass;
  mp->n = n;
  mp->d = d;
  mp->DP = d <= 16 ? 16 : d <= 32 ? 32 : 64; // a power of two: swizzled 16-byte chunks
  mp->tileCols = tileColsFor(mp->DP);
  mp->nPad = (n + mp->tileCols - 1) / mp->tileCols * mp->tileCols;
  mp->X = X;
  Status s;
  DevArray<double> sums;
  DevArray<unsigned int> maxes;
  s(sums.alloc(d));
  s(maxes.alloc(3));
  s(mp->mu.alloc(d));
  s(mp->xh.alloc((size_t)n * mp->DP));
  s(mp->xhp.alloc((size_t)mp->nPad * mp->DP));
  s(mp->norm.alloc(n));
  s(mp->normp.alloc((size_t)mp->nPad / mp->tileCols * kNormSlots));
  s(mp->R.alloc(n));
  s(mp->rho.alloc(n));
  s(mp->perm.alloc(n));
  // Row slots: at most twice the rows (workgroup-aligned components, padding).
  s(mp->rows.alloc((size_t)3 * 2 * n));
  s(mp->candCount.alloc((size_t)2 * 2 * n));
  s(mp->theta.alloc((size_t)2 * n));
  s(mp->rowMin.alloc((size_t)2 * n));
  s(mp->xp.alloc((size_t)n * d));
  s(mp->cand.alloc((size_t)2 * 2 * n * kCandCap));
  s(mp->topk.alloc((size_t)n * kTopK));
  s(mp->largest.alloc(1));
  s(mp->stats.alloc(5));
  s(mp->recheckStats.alloc((size_t)kStatStripes * kStripeWords));
  s(mp->overflowRows.alloc((size_t)2 * n)); // one entry per row slot at most
  s(mp->overflowCount.alloc(1));
  s(mp->positions.alloc(n));
  s(mp->sortedComp.alloc(n));
  s(mp->compStart.alloc(n));
  s(mp->compEnd.alloc(n));
  s(mp->rowFlag.alloc(n + 1));
  s(mp->rowIndex.alloc(n + 1));
  s(mp->bigSlots.alloc(n + 1));
  s(mp->smallSlots.alloc(n + 1));
  s(mp->bigBase.alloc(n + 1));
  s(mp->smallBase.alloc(n + 1));
  mp->sortBytes = sortTempBytes32(n);
  mp->scanBytes = scanTempBytes(n + 1);
  s(mp->sortTemp.alloc(mp->sortBytes));
  s(mp->scanTemp.alloc(mp->scanBytes));
  if (s.ok())
    hipLaunchKernelGGL(positionsKernel, dim3((n + 255) / 256), dim3(256), 0, stream(),
                       mp->positions.get(), n);
  if (s.ok()) {
    s(fill(sums, 0, d * sizeof(double)));
    s(fill(maxes, 0, 3 * sizeof(unsigned int)));
    s(fill(mp->stats, 0, 5 * sizeof(unsigned long long)));
    s(fill(mp->recheckStats, 0, mp->recheckStats.bytes()));
    hipLaunchKernelGGL(columnSumsKernel, dim3(std::max(1, std::min(1024, n / 64 + 1))),
                       dim3(256), 0, stream(), X, n, d, sums.get());
    std::vector<double> hs(d);
    s(download(hs.data(), sums, d * sizeof(double)));
    std::vector<float> hmu(d);
    for (int k = 0; k < d; ++k)
      hmu[k] = (float)(hs[k] / n);
    s(upload(mp->mu, hmu.data(), d * sizeof(float)));
    hipLaunchKernelGGL(maxDevKernel, dim3(1024), dim3(256), 0, stream(), X, n, d, mp->mu.get(),
                       maxes.get());
    unsigned int mb = 0;
    s(download(&mb, maxes, sizeof(mb)));
    const float maxDev = floatOf(mb);
    // Power-of-two scale: max |x'| in (2^13, 2^14].
    int e = 0;
    (void)std::frexp(maxDev > 0.0f && std::isfinite(maxDev) ? maxDev : 1.0f, &e);
    const double scale = std::ldexp(1.0, 14 - e);
    hipLaunchKernelGGL(toHalfKernel, dim3(1024), dim3(256), 0, stream(), X, n, d, mp->DP,
                       mp->mu.get(), scale, mp->xh.get(), mp->norm.get(), mp->R.get(),
                       mp->rho.get(), maxes.get() + 1, maxes.get() + 2);
    unsigned int rr[2] = {0, 0};
    s(download(rr, maxes.get() + 1, sizeof(rr)));
    mp->Rmax = floatOf(rr[0]);
    mp->rhoMax = floatOf(rr[1]);
  }
  *status = s.code;
  if (!s.ok()) {
    delete mp;
    return nullptr;
  }
  return mp;
}

void mixedPassDestroy(MixedPass *mp) {
  if (!mp)
    return;
  unsigned long long st[5] = {0, 0, 0, 0, 0};
  std::vector<unsigned long long> stripes((size_t)kStatStripes * kStripeWords);
  if (download(stripes.data(), mp->recheckStats, mp->recheckStats.bytes()) == 0 &&
      download(st, mp->stats, sizeof(st)) == 0) {
    for (int k = 0; k < kStatStripes; ++k)
      for (int i = 1; i < 4; ++i)
        st[i] += stripes[(size_t)k * kStripeWords + i];
    for (int i = 0; i < 4; ++i)
      __atomic_fetch_add(&gStats[i], st[i], __ATOMIC_RELAXED);
    addUnits(kPhaseMfmaCandidates, (long long)st[4]); // pairs the MFMA tiles computed
  }
  delete mp;
}

int mixedPassRun(MixedPass *mp, const int *comp, int numComponents, bool sqeuclidean,
                 bool firstRound, unsigned long long *nearest) {
  const int n = mp->n;
  Status s;
  // Component-contiguous order: a stable radix sort of the positions by
  // component (so by component, then original index), and each component's
  // range in it.
  int bits = 1;
  while (bits < 31 && (numComponents - 1) >> bits)
    ++bits;
  // Component labels are non-negative: sorted as unsigned keys.
  s(radixSortPairs(reinterpret_cast<const unsigned int *>(comp),
                   reinterpret_cast<unsigned int *>(mp->sortedComp.get()), mp->positions.get(),
                   mp->perm.get(), n, bits, mp->sortTemp.get()));
  const dim3 grid((n + 255) / 256);
  hipLaunchKernelGGL(componentRangesKernel, grid, dim3(256), 0, stream(), mp->sortedComp.get(), n,
                     mp->compStart.get(), mp->compEnd.get());
  // Rows to search: all in the first round; later only points whose kept
  // list has no entry outside their component (the others are settled here),
  // and none of a component holding more than half of the points. Listed in
  // component order.
  s(fill(mp->largest, 0, sizeof(unsigned long long)));
  hipLaunchKernelGGL(largestComponentKernel, dim3((numComponents + 255) / 256), dim3(256), 0,
                     stream(), mp->compStart.get(), mp->compEnd.get(), numComponents,
                     mp->largest.get());
  hipLaunchKernelGGL(staleRowsKernel, dim3((n + 1 + 255) / 256), dim3(256), 0, stream(),
                     mp->perm.get(), comp, mp->topk.get(), nearest, n, firstRound,
                     mp->largest.get(), mp->rowFlag.get());
  s(exclusiveScan(mp->rowFlag.get(), mp->rowIndex.get(), n + 1, mp->scanTemp.get()));
  hipLaunchKernelGGL(componentSlotsKernel, dim3((numComponents + 1 + 255) / 256), dim3(256), 0,
                     stream(), mp->compStart.get(), mp->compEnd.get(), mp->rowIndex.get(),
                     numComponents, mp->bigSlots.get(), mp->smallSlots.get());
  s(exclusiveScan(mp->bigSlots.get(), mp->bigBase.get(), numComponents + 1,
                  mp->scanTemp.get()));
  s(exclusiveScan(mp->smallSlots.get(), mp->smallBase.get(), numComponents + 1,
                  mp->scanTemp.get()));
  s(fill(mp->rows, 0xff, mp->rows.bytes())); // every slot starts as padding (position -1)
  hipLaunchKernelGGL(rowListKernel, grid, dim3(256), 0, stream(), mp->perm.get(), comp,
                     mp->rowFlag.get(), mp->rowIndex.get(), mp->compStart.get(),
                     mp->compEnd.get(), mp->bigSlots.get(), mp->bigBase.get(),
                     mp->smallBase.get(), numComponents, n, mp->rows.get());
  s((int)hipGetLastError());
  int slotTotals[2] = {0, 0};
  s(download(&slotTotals[0], mp->bigBase.get() + numComponents, sizeof(int)));
  s(download(&slotTotals[1], mp->smallBase.get() + numComponents, sizeof(int)));
  const int numRows = slotTotals[0] + slotTotals[1]; // row slots, padding included
  if (!s.ok() || numRows == 0)
    return s.code;
  hipLaunchKernelGGL(permuteHalfKernel, dim3(2048), dim3(256), 0, stream(), mp->xh.get(),
                     mp->norm.get(), mp->perm.get(), n, mp->nPad, mp->DP, mp->tileCols, mp->xhp.get(),
                     mp->normp.get());
  s(gatherRows(mp->X, mp->perm, n, mp->d, mp->xp));
  CandArgs ca;
  ca.xhp = mp->xhp;
  ca.normp = mp->normp;
  ca.xh = mp->xh;
  ca.norm = mp->norm;
  ca.R = mp->R;
  ca.rho = mp->rho;
  ca.perm = mp->perm;
  ca.rows = mp->rows;
  ca.numRows = numRows;
  ca.n = n;
  ca.nPad = mp->nPad;
  ca.Rmax = mp->Rmax;
  ca.rhoMax = mp->rhoMax;
  const int KB = mp->DP / 16;
  ca.gamma = (float)((16 * KB + 8) * std::ldexp(1.0, -23));
  const double eps = (mp->d + 4) * std::ldexp(1.0, -24);
  ca.kappa = (1.0 + eps) * (1.0 + std::ldexp(1.0, -21)) / (1.0 - eps) * (1.0 + 1e-9);
  ca.cand = mp->cand;
  ca.candCount = mp->candCount;
  ca.theta = mp->theta;
  ca.rowMin = mp->rowMin;
  ca.recorded = mp->stats.get();
  // The first round keeps each point's kTopK nearest, which settle most rows
  // of the second round; later rounds keep the nearest only.
  ca.rank = firstRound ? kTopK : 1;
  // The first round keeps the kTopK nearest of every point: a longer dry run
  // (kDryTilesFirstRound tiles, 0.4% of its columns) starts its records much
  // closer to their final theta. Later rounds search few rows over many
  // column splits; each split's dry run of kDryTilesLaterRounds tiles keeps
  // its records, before the splits share theta, within the lists.
  ca.dryTiles = firstRound ? kDryTilesFirstRound : kDryTilesLaterRounds;
  const int blocks = (numRows + kBlockRows - 1) / kBlockRows;
  // Column splits (at most kMaxSplits, each keeping at least 2,048 columns):
  // the fewest splits whose workgroups fill the last wave of resident slots
  // (two per CU) at least 90%, else the best fill. Late rounds with few rows
  // would otherwise leave CUs idle (rounds 2-5 at 10^6 points: 6.6 ms with
  // at most 16 splits, about 3.2 ms with 128 or 256; with 256 no list of
  // those rounds overflowed, against about 20 with 128), and a grid just
  // above a multiple of the resident slots would run its last wave nearly
  // empty.
  const long long slots = 2LL * numCUs();
  const int maxSplits = std::max(1, std::min(kMaxSplits, mp->nPad / 2048));
  int splits = 1;
  double bestFill = 0.0;
  for (int sp = 1; sp <= maxSplits; ++sp) {
    const long long units = (long long)blocks * sp;
    const double fill = (double)units / (double)(((units + slots - 1) / slots) * slots);
    if (fill > bestFill + 1e-9) {
      bestFill = fill;
      splits = sp;
    }
    if (fill >= 0.9)
      break;
  }
  s(fill(mp->theta, 0xff, (size_t)numRows * sizeof(unsigned int)));
  if (splits > 1) {
    s(fill(mp->candCount, 0, (size_t)numRows * 2 * sizeof(int)));
    s(fill(mp->rowMin, 0xff, (size_t)numRows * sizeof(unsigned int)));
  }
  auto launch = [&](auto unsplit, auto split) {
    if (splits > 1)
      hipLaunchKernelGGL(split, dim3(blocks, splits), dim3(256), 0, stream(), ca);
    else
      hipLaunchKernelGGL(unsplit, dim3(blocks), dim3(256), 0, stream(), ca);
  };
  {
    PhaseTimer timer(kPhaseMfmaCandidates);
    switch (KB) {
    case 1: launch(mfmaCandidateKernel<1, false>, mfmaCandidateKernel<1, true>); break;
    case 2: launch(mfmaCandidateKernel<2, false>, mfmaCandidateKernel<2, true>); break;
    default: launch(mfmaCandidateKernel<4, false>, mfmaCandidateKernel<4, true>); break;
    }
    s((int)hipGetLastError());
  }
  RecheckArgs ra;
  ra.X = mp->X;
  ra.xp = mp->xp;
  ra.n = n;
  ra.d = mp->d;
  ra.perm = mp->perm;
  ra.rows = mp->rows;
  ra.numRows = numRows;
  ra.cand = mp->cand;
  ra.candCount = mp->candCount;
  ra.theta = mp->theta;
  ra.sqeuclidean = sqeuclidean;
  ra.nearest = nearest;
  ra.topk = mp->topk;
  ra.rank = ca.rank;
  ra.stats = mp->recheckStats;
  ra.overflowRows = mp->overflowRows;
  ra.overflowCount = mp->overflowCount;
  s(fill(mp->overflowCount, 0, sizeof(int)));
  int overflowed = 0;
  {
    PhaseTimer timer(kPhaseRecheck);
    hipLaunchKernelGGL(recheckKernel, dim3((numRows + 3) / 4), dim3(256), 0, stream(), ra);
    s((int)hipGetLastError());
  }
  s(download(&overflowed, mp->overflowCount, sizeof(int)));
  if (s.ok() && overflowed > 0) {
    // Rows whose lists overflowed, at most 1,024 per batch and at most 32 MiB
    // of partial keys: every column, kOverflowWaveCols per workgroup.
    PhaseTimer timer(kPhaseRecheck);
    const int groups = (int)(((long long)n + kOverflowWaveCols - 1) / kOverflowWaveCols);
    const int batch = std::max(1, std::min({overflowed, 1024, (32 << 20) / (groups * kTopK * 8)}));
    DevArray<unsigned long long> partial;
    s(partial.alloc((size_t)batch * groups * kTopK));
    for (int first = 0; s.ok() && first < overflowed; first += batch) {
      const int rowsNow = std::min(batch, overflowed - first);
      const int rowGroups = (rowsNow + kOverflowRowsPerGroup - 1) / kOverflowRowsPerGroup;
      hipLaunchKernelGGL(overflowScanKernel, dim3(groups, rowGroups), dim3(256), 0, stream(), ra,
                         first, rowsNow, partial.get());
      hipLaunchKernelGGL(overflowMergeKernel, dim3(rowsNow), dim3(64), 0, stream(), ra, first,
                         partial.get(), groups);
      s((int)hipGetLastError());
    }
  }
  return s.code;
}

} // namespace cuhc

extern "C" void cuhc_mixed_stats(unsigned long long *recorded, unsigned long long *rechecked,
                                 unsigned long long *pairs, unsigned long long *overflows) {
  using cuhc::gStats;
  *recorded = __atomic_exchange_n(&gStats[0], 0ull, __ATOMIC_RELAXED);
  *rechecked = __atomic_exchange_n(&gStats[1], 0ull, __ATOMIC_RELAXED);
  *pairs = __atomic_exchange_n(&gStats[2], 0ull, __ATOMIC_RELAXED);
  *overflows = __atomic_exchange_n(&gStats[3], 0ull, __ATOMIC_RELAXED);
}
