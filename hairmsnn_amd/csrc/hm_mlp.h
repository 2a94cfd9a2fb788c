// Radiance-cache MLP (hm_mlp.hip): host-side interface used by the HairMSNN
// frame loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace hm {

constexpr int kMlpWeightCount = 5 * 64 * 64 + 3 * 64;  // 20672, no biases

struct MlpState;
int mlp_create(int precision, uint64_t seed, MlpState** out, std::string& err);
void mlp_destroy(MlpState* s);
int mlp_set_weights(MlpState& s, const float* w_host, std::string& err);
int mlp_get_weights(const MlpState& s, float* w_host, std::string& err);
int mlp_forward(MlpState& s, const float* in, int n, float* out, hipStream_t st, std::string& err);
int mlp_loss_grad_step(MlpState& s, const float* in, const float* target, int n, int loss, float* grad_out,
                       float* loss_out, bool adam, float lr, hipStream_t st, std::string& err);
int mlp_encode(const float* in, int n, float* feat, hipStream_t st, std::string& err);
// Weight files (hm_mlp_io.cpp): kMlpWeightCount floats in the layout of
// mlp_get_weights.
int mlp_weights_write(const std::string& path, const float* w, std::string& err);
int mlp_weights_read(const std::string& path, float* w, std::string& err);
int mlp_save(const MlpState& s, const std::string& path, std::string& err);
int mlp_load(MlpState& s, const std::string& path, std::string& err);

}  // namespace hm
