// Training and inference kernels of the radius-proposal network (mlp.hpp).
//
// Forward/backward: one thread per sample keeps both hidden layers in
// registers and writes the activations and back-propagated errors it needs
// (h1, h2, dz1, dz2, dy) to a work buffer. Gradients: FP64 sums over fixed
// sample chunks, then over the chunks in order (deterministic, no atomics).
// Adam: one thread per parameter.
#include "mlp.hpp"
#include "common.hpp"

#include <cmath>
#include <vector>

namespace cuhc {

namespace {

constexpr int F = kMlpIn, H = kMlpHidden, P = kMlpParams;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H;
// Work layout per sample: h1[H], h2[H], dz1[H], dz2[H], dy.
constexpr int WS = 4 * H + 1;

DEV void forwardOne(const float *__restrict__ p, const float *x, float (&h1)[H], float (&h2)[H],
                    float &y) {
#pragma unroll
  for (int h = 0; h < H; ++h) {
    float z = p[OB1 + h];
#pragma unroll
    for (int f = 0; f < F; ++f)
      z = fmaf(p[OW1 + h * F + f], x[f], z);
    h1[h] = tanhf(z);
  }
#pragma unroll 4
  for (int h = 0; h < H; ++h) {
    float z = p[OB2 + h];
#pragma unroll
    for (int g = 0; g < H; ++g)
      z = fmaf(p[OW2 + h * H + g], h1[g], z);
    h2[h] = tanhf(z);
  }
  float out = p[OB3];
#pragma unroll
  for (int h = 0; h < H; ++h)
    out = fmaf(p[OW3 + h], h2[h], out);
  y = out;
}

__global__ void __launch_bounds__(256) forwardKernel(const float *__restrict__ p,
                                                     const float *__restrict__ X, int m,
                                                     float *__restrict__ y) {
  __shared__ float sp[P];
  for (int i = threadIdx.x; i < P; i += blockDim.x)
    sp[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m)
    return;
  float x[F], h1[H], h2[H], out;
#pragma unroll
  for (int f = 0; f < F; ++f)
    x[f] = X[(size_t)s * F + f];
  forwardOne(sp, x, h1, h2, out);
  y[s] = out;
}

__global__ void __launch_bounds__(256) backwardKernel(const float *__restrict__ p,
                                                      const float *__restrict__ X,
                                                      const float *__restrict__ t, int m,
                                                      float *__restrict__ work) {
  __shared__ float sp[P];
  for (int i = threadIdx.x; i < P; i += blockDim.x)
    sp[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m)
    return;
  float x[F], h1[H], h2[H], y;
#pragma unroll
  for (int f = 0; f < F; ++f)
    x[f] = X[(size_t)s * F + f];
  forwardOne(sp, x, h1, h2, y);
  const float dy = y - t[s];
  float dz2[H];
#pragma unroll
  for (int h = 0; h < H; ++h)
    dz2[h] = dy * sp[OW3 + h] * (1.0f - h2[h] * h2[h]);
  float *w = work + (size_t)s * WS;
#pragma unroll 4
  for (int g = 0; g < H; ++g) {
    float back = 0.0f;
#pragma unroll
    for (int h = 0; h < H; ++h)
      back = fmaf(sp[OW2 + h * H + g], dz2[h], back);
    w[2 * H + g] = back * (1.0f - h1[g] * h1[g]); // dz1
  }
#pragma unroll
  for (int h = 0; h < H; ++h) {
    w[h] = h1[h];
    w[H + h] = h2[h];
    w[3 * H + h] = dz2[h];
  }
  w[4 * H] = dy;
}

// Gradient in two deterministic stages. Stage 1: workgroup (g, c) sums, for
// the 64 parameters of group g (one per lane), the products of samples chunk c
// (kChunk samples, a quarter per wave, the waves' sums added in wave order) in
// FP64 into partial[c][q]. Stage 2: one thread per parameter adds the chunks
// in order and divides by m; thread P writes the loss (mean of 0.5 dy^2).
constexpr int kChunk = 256;

DEV double productOf(int q, const float *__restrict__ X, const float *__restrict__ w, int s) {
  if (q < OB1) {
    const int h = (q - OW1) / F, f = (q - OW1) % F;
    return (double)w[2 * H + h] * X[(size_t)s * F + f];
  }
  if (q < OW2)
    return w[2 * H + (q - OB1)];
  if (q < OB2) {
    const int h = (q - OW2) / H, g = (q - OW2) % H;
    return (double)w[3 * H + h] * w[g];
  }
  if (q < OW3)
    return w[3 * H + (q - OB2)];
  if (q < OB3)
    return (double)w[4 * H] * w[H + (q - OW3)];
  if (q == OB3)
    return w[4 * H];
  return 0.5 * (double)w[4 * H] * w[4 * H]; // q == P: the loss term
}

__global__ void __launch_bounds__(256)
    gradientPartialKernel(const float *__restrict__ X, const float *__restrict__ work, int m,
                          double *__restrict__ partial) {
  __shared__ double waveSum[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane; // 0..P (P = the loss)
  const int chunk = blockIdx.y;
  const int first = chunk * kChunk, last = min(m, first + kChunk);
  const int per = kChunk / 4;
  double acc = 0.0;
  if (q <= P)
    for (int s = first + wave * per; s < min(last, first + (wave + 1) * per); ++s)
      acc += productOf(q, X, work + (size_t)s * WS, s);
  waveSum[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && q <= P)
    partial[(size_t)chunk * (P + 1) + q] =
        ((waveSum[0][lane] + waveSum[1][lane]) + waveSum[2][lane]) + waveSum[3][lane];
}

__global__ void gradientFinishKernel(const double *__restrict__ partial, int chunks, int m,
                                     float *__restrict__ grad, float *loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P)
    return;
  double acc = 0.0;
  for (int c = 0; c < chunks; ++c)
    acc += partial[(size_t)c * (P + 1) + q];
  if (q < P)
    grad[q] = (float)(acc / m);
  else if (loss)
    *loss = (float)(acc / m);
}

__global__ void adamKernel(float *p, const float *g, float *mom, float *vel, int step, float lr) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P)
    return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  mom[q] = b1 * mom[q] + (1.0f - b1) * g[q];
  vel[q] = b2 * vel[q] + (1.0f - b2) * g[q] * g[q];
  const float mh = mom[q] / (1.0f - powf(b1, (float)step));
  const float vh = vel[q] / (1.0f - powf(b2, (float)step));
  p[q] -= lr * mh / (sqrtf(vh) + eps);
}

} // namespace

size_t mlpWorkFloats(int m) {
  // per-sample records, then the FP64 chunk partials (P + 1 per chunk)
  const size_t chunks = (size_t)(m + kChunk - 1) / kChunk;
  return (size_t)m * WS + 2 * chunks * (P + 1) + 2;
}

int mlpGradient(const float *params, const float *X, const float *t, int m, float *work,
                float *grad, float *loss) {
  if (m <= 0)
    return 0;
  const int chunks = (m + kChunk - 1) / kChunk;
  // 8-byte aligned start of the partials after the per-sample records
  double *partial = reinterpret_cast<double *>(
      work + (((size_t)m * WS + 1) & ~(size_t)1));
  hipLaunchKernelGGL(backwardKernel, dim3((m + 255) / 256), dim3(256), 0, stream(), params, X, t,
                     m, work);
  hipLaunchKernelGGL(gradientPartialKernel, dim3((P + 1 + 63) / 64, chunks), dim3(256), 0,
                     stream(), X, work, m, partial);
  hipLaunchKernelGGL(gradientFinishKernel, dim3((P + 1 + 255) / 256), dim3(256), 0, stream(),
                     partial, chunks, m, grad, loss);
  return (int)hipGetLastError();
}

int mlpForward(const float *params, const float *X, int m, float *y) {
  if (m <= 0)
    return 0;
  hipLaunchKernelGGL(forwardKernel, dim3((m + 255) / 256), dim3(256), 0, stream(), params, X, m,
                     y);
  return (int)hipGetLastError();
}

int mlpTrain(float *params, const float *X, const float *t, int m, int steps, float lr) {
  DevArray<float> work, grad, mom, vel;
  Status s;
  s(work.alloc(mlpWorkFloats(m)));
  s(grad.alloc(P));
  s(mom.alloc(P));
  s(vel.alloc(P));
  if (!s.ok())
    return s.code;
  s(fill(mom, 0, P * sizeof(float)));
  s(fill(vel, 0, P * sizeof(float)));
  PhaseTimer timer(kPhaseMlp);
  for (int step = 1; step <= steps && s.ok(); ++step) {
    s(mlpGradient(params, X, t, m, work, grad, nullptr));
    hipLaunchKernelGGL(adamKernel, dim3((P + 255) / 256), dim3(256), 0, stream(), params,
                       grad.get(), mom.get(), vel.get(), step, lr);
    s((int)hipGetLastError());
  }
  return s.code;
}

void mlpInitHost(float *params, unsigned seed) {
  // xorshift32, uniform in [-1/sqrt(fan_in), 1/sqrt(fan_in)]
  unsigned x = seed ? seed : 0x9e3779b9u;
  auto uniform = [&]() {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return (x >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
  };
  for (int q = 0; q < P; ++q) {
    const int fanIn = q < OB1 ? F : (q < OW2 ? F : (q < OB2 ? H : (q < OW3 ? H : H)));
    params[q] = uniform() / std::sqrt((float)fanIn);
  }
}

} // namespace cuhc
