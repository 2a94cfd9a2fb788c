// Environment importance-sampling tables, built on the GPU.
//
// Same tables as the reference's generateEnvSamplingTables (scene.cpp:349-425):
//   conditional pdf[y][x] = mean(RGB) * sin(pi (y + 1/2) / H), x < W
//   conditional cdf[y][x] = sum_{k<x} pdf[y][k] / W, normalised by the row
//                           total; entry W of the pdf row holds the total,
//                           entry W of the cdf row is 1
//   marginal  pdf[y] = row total, cdf = running sum / H, normalised, with
//                      the grand total in pdf[H] and 1 in cdf[H]
// The reference accumulates these running sums sequentially in FP32; here
// each prefix sum is a block-parallel scan in FP64, rounded to FP32 once.
#include <hip/hip_runtime.h>

#include "hm_env.h"
#include "hm_internal.h"

namespace hm {

namespace {

constexpr int kEnvBlock = 256;

// Block-wide exclusive scan of one double per thread; returns the total.
__device__ double block_exclusive_scan(double v, double* out_excl, double* lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const double o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  double wave_base = 0.0, total = 0.0;
  for (int w = 0; w < kEnvBlock / 64; ++w) {
    if (w < wave) wave_base += lds[w];
    total += lds[w];
  }
  *out_excl = wave_base + incl - v;
  __syncthreads();
  return total;
}

__global__ void k_env_rows(const float4* tex, int W, int H, float* c_pdf, float* c_cdf) {
  __shared__ double lds[kEnvBlock / 64];
  const int y = blockIdx.x;
  const int cw = W + 1;
  const float sin_t = sinf(kHostPi * ((float)y + 0.5f) / (float)H);
  const int chunk = (W + kEnvBlock - 1) / kEnvBlock;
  const int x0 = threadIdx.x * chunk;
  // pass 1: this thread's partial sum of pdf / W
  double part = 0.0;
  for (int k = 0; k < chunk; ++k) {
    const int x = x0 + k;
    if (x < W) {
      const float4 t = tex[(size_t)y * W + x];
      const float pdf = (t.x + t.y + t.z) * (1.0f / 3.0f) * sin_t;
      c_pdf[(size_t)y * cw + x] = pdf;
      part += (double)pdf / (double)W;
    }
  }
  double base;
  const double total = block_exclusive_scan(part, &base, lds);
  // pass 2: running sums, normalised by the row total
  double run = base;
  for (int k = 0; k < chunk; ++k) {
    const int x = x0 + k;
    if (x < W) {
      const float pdf = c_pdf[(size_t)y * cw + x];
      const double cdf = (x == 0) ? 0.0 : (total > 0.0 ? run / total : run);
      c_cdf[(size_t)y * cw + x] = (float)cdf;
      run += (double)pdf / (double)W;
    }
  }
  if (threadIdx.x == 0) {
    c_pdf[(size_t)y * cw + W] = (float)total;
    c_cdf[(size_t)y * cw + W] = 1.f;
  }
}

__global__ void k_env_marginal(int W, int H, const float* c_pdf, float* m_pdf, float* m_cdf) {
  __shared__ double lds[kEnvBlock / 64];
  const int cw = W + 1;
  const int chunk = (H + kEnvBlock - 1) / kEnvBlock;
  const int y0 = threadIdx.x * chunk;
  double part = 0.0;
  for (int k = 0; k < chunk; ++k) {
    const int y = y0 + k;
    if (y < H) {
      const float row = c_pdf[(size_t)y * cw + W];
      m_pdf[y] = row;
      part += (double)row / (double)H;
    }
  }
  double base;
  const double total = block_exclusive_scan(part, &base, lds);
  double run = base;
  for (int k = 0; k < chunk; ++k) {
    const int y = y0 + k;
    if (y < H) {
      const double cdf = (y == 0) ? 0.0 : (total > 0.0 ? run / total : run);
      m_cdf[y] = (float)cdf;
      run += (double)m_pdf[y] / (double)H;
    }
  }
  if (threadIdx.x == 0) {
    m_pdf[H] = (float)total;
    m_cdf[H] = 1.f;
  }
}

}  // namespace

// Guide entries of `rows` CDFs (row r at cdf + r * stride, count entries
// searched): guide[r][k] = lower_bound(cdf_r, count, k / kEnvGuideBins)
// before the max(0, . - 1) step, for k = 0 .. kEnvGuideBins.
__global__ void k_env_guide(const float* cdf, int stride, int count, int rows, int* guide) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k > kEnvGuideBins || r >= rows) return;
  const float u = (float)k / (float)kEnvGuideBins;  // exact: power-of-two bins
  const float* c = cdf + (size_t)r * stride;
  int first = 0, n = count;
  while (n > 0) {
    const int step = n >> 1;
    const int mid = first + step;
    if (c[mid] < u) {
      first = mid + 1;
      n -= step + 1;
    } else {
      n = step;
    }
  }
  guide[(size_t)r * (kEnvGuideBins + 1) + k] = first;
}

int env_build_tables(DeviceEnv& env, hipStream_t stream, std::string& err) {
  const int W = env.width, H = env.height;
  if (W <= 0 || H <= 0 || !env.tex) {
    err = "environment map is empty";
    return kEInval;
  }
  dev_free(env.c_pdf);
  dev_free(env.c_cdf);
  dev_free(env.m_pdf);
  dev_free(env.m_cdf);
  env.c_pdf = env.c_cdf = env.m_pdf = env.m_cdf = nullptr;
  HM_TRY(dev_alloc((void**)&env.c_pdf, sizeof(float) * (size_t)(W + 1) * H, err));
  HM_TRY(dev_alloc((void**)&env.c_cdf, sizeof(float) * (size_t)(W + 1) * H, err));
  HM_TRY(dev_alloc((void**)&env.m_pdf, sizeof(float) * (size_t)(H + 1), err));
  HM_TRY(dev_alloc((void**)&env.m_cdf, sizeof(float) * (size_t)(H + 1), err));
  hipLaunchKernelGGL(k_env_rows, dim3(H), dim3(kEnvBlock), 0, stream, env.tex, W, H, env.c_pdf, env.c_cdf);
  HM_TRY(check_launch("env rows", err));
  hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(kEnvBlock), 0, stream, W, H, env.c_pdf, env.m_pdf,
                     env.m_cdf);
  HM_TRY(check_launch("env marginal", err));
  dev_free(env.m_guide);
  dev_free(env.c_guide);
  env.m_guide = env.c_guide = nullptr;
  HM_TRY(dev_alloc((void**)&env.m_guide, sizeof(int) * (kEnvGuideBins + 1), err));
  HM_TRY(dev_alloc((void**)&env.c_guide, sizeof(int) * (size_t)(kEnvGuideBins + 1) * H, err));
  const int gb = (kEnvGuideBins + 1 + 255) / 256;
  hipLaunchKernelGGL(k_env_guide, dim3(gb, 1), dim3(256), 0, stream, env.m_cdf, 0, H, 1, env.m_guide);
  hipLaunchKernelGGL(k_env_guide, dim3(gb, H), dim3(256), 0, stream, env.c_cdf, W + 1, W, H, env.c_guide);
  HM_TRY(check_launch("env guide tables", err));
  return kOk;
}

}  // namespace hm
