// ORACLE - test infrastructure only (see orc_math.h).
//
// FP64 / FP32 restatement of the radiance-cache network used by the
// HairMSNN renderer (the reference trains a tiny-cuda-nn network through
// TINY_MLP, neural_network.cuh:33-50; that library is not vendored, so the
// network here follows the build's documented definition):
//   input record (12): position / sceneScale (3), wo (3), tangent (3), sigma_a (3)
//   encoding (64):     sin/cos(2^k pi p_c), k = 0..5, c = x,y,z       (36)
//                      (theta/pi, phi/2pi + 1/2) of wo and of tangent  (4)
//                      sigma_a (3), constant 1 (bias feature)          (4)
//                      zero padding                                    (20)
//   5 hidden layers 64 -> 64 with ReLU, no biases; output 64 -> 3, exp.
//   loss: relative L2 on luminance (0), relative L2 (1) or L2 (2),
//         mean over batch x 3; denominators treated as constants.
//   Adam with L2 regularisation folded into the gradient.
#pragma once

#include <cstdint>

namespace orc {

constexpr int kMlpIn = 12;
constexpr int kMlpFeat = 64;
constexpr int kMlpHidden = 5;
constexpr int kMlpOut = 3;
constexpr int kMlpWeights = kMlpHidden * kMlpFeat * kMlpFeat + kMlpOut * kMlpFeat;  // 20672

template <class R>
void mlp_encode(const float* rec, R* feat);

// Forward for n records: out (n x 3). Weights as in hmsnn_mlp_get_weights.
template <class R>
void mlp_forward(const float* w, const float* in, int n, R* out);

// Loss over the batch and dL/dW (FP64 accumulation). The derivative of the
// exp output activation is taken as max(y, exp_floor) (0: the exact exp).
template <class R>
double mlp_loss_grad(const float* w, const float* in, const float* target, int n, int loss, double exp_floor,
                     R* grad);

}  // namespace orc
