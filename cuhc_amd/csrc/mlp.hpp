// A small multilayer perceptron with hand-written HIP training and inference
// kernels: F inputs -> H tanh -> H tanh -> 1 linear. Used to propose per-point
// search radii (learned.hip). No reference counterpart.
//
// Parameter layout (floats): W1[H][F], b1[H], W2[H][H], b2[H], W3[H], b3.
#pragma once

#include "runtime.hpp"

namespace cuhc {

constexpr int kMlpIn = 4;
constexpr int kMlpHidden = 32;
constexpr int kMlpParams =
    kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;

// Mean-squared-error gradient (0.5 (y - t)^2 averaged over m samples) of the
// parameters at `params` on samples X (m x kMlpIn) with targets t: all device
// pointers; per-sample activations go to `work` (mlpWorkFloats(m) floats).
// Deterministic: every parameter's sum runs over the samples in order.
size_t mlpWorkFloats(int m);
int mlpGradient(const float *params, const float *X, const float *t, int m, float *work,
                float *grad, float *loss /* device, may be null */);

// y[i] = network(X[i]) for m samples.
int mlpForward(const float *params, const float *X, int m, float *y);

// Adam on the MSE loss, full batch, `steps` updates from `params` (in place).
int mlpTrain(float *params, const float *X, const float *t, int m, int steps, float lr);

// Deterministic initial parameters (uniform +-1/sqrt(fan_in), fixed seed).
void mlpInitHost(float *params, unsigned seed);

} // namespace cuhc
