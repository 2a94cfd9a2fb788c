// Batched hair-BSDF kernels behind hmsnn_hair_bsdf_eval_dev /
// hmsnn_hair_bsdf_sample_dev: one query per lane, 40-byte query records.
#include <hip/hip_runtime.h>

#include "hm_hair_bsdf.h"
#include "hm_internal.h"

namespace hm {

struct BsdfQuery {
  float wo[3];
  float wi[3];
  float h;
  float sigma_a[3];
};

namespace {

__global__ void k_bsdf_eval(HairConsts hc, const BsdfQuery* q, int n, float* f_rgb, float* pdf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BsdfQuery x = q[i];
  const HairWo w = hair_prepare(hc, v3(x.wo[0], x.wo[1], x.wo[2]), x.h, v3(x.sigma_a[0], x.sigma_a[1], x.sigma_a[2]));
  float p;
  const V3 f = hair_eval(hc, w, v3(x.wi[0], x.wi[1], x.wi[2]), &p);
  f_rgb[3 * (size_t)i] = f.x;
  f_rgb[3 * (size_t)i + 1] = f.y;
  f_rgb[3 * (size_t)i + 2] = f.z;
  pdf[i] = p;
}

__global__ void k_bsdf_sample(HairConsts hc, const BsdfQuery* q, const float4* u4, int n, float* wi_out,
                              float* f_rgb, float* pdf, int* lobe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BsdfQuery x = q[i];
  const HairWo w = hair_prepare(hc, v3(x.wo[0], x.wo[1], x.wo[2]), x.h, v3(x.sigma_a[0], x.sigma_a[1], x.sigma_a[2]));
  const float4 u = u4[i];
  int p;
  const V3 wi = hair_sample_dir(hc, w, u.x, u.y, u.z, u.w, &p);
  float pd;
  const V3 f = hair_eval(hc, w, wi, &pd);
  wi_out[3 * (size_t)i] = wi.x;
  wi_out[3 * (size_t)i + 1] = wi.y;
  wi_out[3 * (size_t)i + 2] = wi.z;
  f_rgb[3 * (size_t)i] = f.x;
  f_rgb[3 * (size_t)i + 1] = f.y;
  f_rgb[3 * (size_t)i + 2] = f.z;
  pdf[i] = pd;
  if (lobe) lobe[i] = p;
}

}  // namespace

int bsdf_eval_launch(const HairConsts& hc, const void* q, int n, float* f, float* pdf, hipStream_t s,
                     std::string& err) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_bsdf_eval, dim3((n + 255) / 256), dim3(256), 0, s, hc, (const BsdfQuery*)q, n, f, pdf);
  return check_launch("bsdf eval", err);
}

int bsdf_sample_launch(const HairConsts& hc, const void* q, const float* u4, int n, float* wi, float* f,
                       float* pdf, int* lobe, hipStream_t s, std::string& err) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(k_bsdf_sample, dim3((n + 255) / 256), dim3(256), 0, s, hc, (const BsdfQuery*)q,
                     (const float4*)u4, n, wi, f, pdf, lobe);
  return check_launch("bsdf sample", err);
}

}  // namespace hm
