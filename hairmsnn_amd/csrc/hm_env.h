// Environment light on the GPU: radiance lookup, pdf and importance sampling
// over the conditional/marginal tables.
//
// Follows cuda_headers/optix_common.cuh:10-149 (getEnvironmentRadiance,
// getEnvironmentPdf, sampleEnvironmentLight). The reference reads the tables
// through NEAREST-filtered textures with normalised coordinates; with the
// table sizes used (width+1 and height+1 entries addressed by index/width and
// index/height) those lookups resolve to plain indices, which is how they are
// written here. The radiance lookup is an explicit FP32 bilinear filter with
// clamp addressing (the reference's hardware filter uses 8-bit weights).
#pragma once

#include "hm_common.h"

namespace hm {

struct EnvView {
  int enabled;
  int pdf_sampling;
  int width, height;
  float scale, rot_phi;
  const float4* tex;
  const float* c_pdf;  // (width+1) per row
  const float* c_cdf;
  const float* m_pdf;  // height+1
  const float* m_cdf;
  const int* m_guide;  // kEnvGuideBins + 1 entries
  const int* c_guide;  // kEnvGuideBins + 1 entries per row
};

// Guide tables for the CDF inversions: entry k of a table is lower_bound of
// the CDF at k / kEnvGuideBins. With a power-of-two bin count, u * bins is
// exact, so the bin of u brackets lower_bound(u) between two entries and the
// search only runs inside them: the same index as a full binary search,
// with ~2 dependent loads instead of ~11.
constexpr int kEnvGuideBins = 1024;

#if defined(__HIPCC__)

HM_D float spherical_theta(V3 d) { return acosf(clampf(d.z, -1.f, 1.f)); }

HM_D float spherical_phi(V3 d) {
  const float p = atan2f(d.y, d.x);
  return p < 0.f ? p + 2.f * kPi : p;
}

// Bilinear filter with clamp addressing at normalised coordinates.
HM_D V3 env_bilinear(const EnvView& e, float x, float y) {
  const float xb = x * (float)e.width - 0.5f, yb = y * (float)e.height - 0.5f;
  const float fx0 = floorf(xb), fy0 = floorf(yb);
  const float a = xb - fx0, b = yb - fy0;
  const int x0 = (int)fx0, y0 = (int)fy0;
  const int xa = min(max(x0, 0), e.width - 1), xb1 = min(max(x0 + 1, 0), e.width - 1);
  const int ya = min(max(y0, 0), e.height - 1), yb1 = min(max(y0 + 1, 0), e.height - 1);
  const float4 t00 = e.tex[ya * e.width + xa], t10 = e.tex[ya * e.width + xb1];
  const float4 t01 = e.tex[yb1 * e.width + xa], t11 = e.tex[yb1 * e.width + xb1];
  const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
  return v3(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
            w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
            w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z);
}

// optix_common.cuh:10-24
HM_D V3 env_radiance(const EnvView& e, V3 dir) {
  const float theta = spherical_theta(dir);
  float phi = spherical_phi(dir) + e.rot_phi;
  if (phi > kTwoPiEnv) phi = phi - kTwoPiEnv;
  return e.scale * env_bilinear(e, phi / (2.f * kPi), theta / kPi);
}

// optix_common.cuh:26-64
HM_D float env_pdf(const EnvView& e, V3 wi) {
  if (!e.pdf_sampling) return 1.f / (4.f * kPi);
  const float theta = spherical_theta(wi);
  float phi = spherical_phi(wi) + e.rot_phi;
  if (phi > kTwoPiEnv) phi = phi - kTwoPiEnv;
  const float u = phi / (2.f * kPi), v = theta / kPi;
  const int iu = min(max((int)(u * (float)e.width), 0), e.width - 1);
  const int iv = min(max((int)(v * (float)e.height), 0), e.height - 1);
  const int cw = e.width + 1;
  const float last_u = e.c_pdf[iv * cw + e.width];
  const float last_v = e.m_pdf[e.height];
  const float denom = (2.f * kPi * kPi * sin_small(theta)) * last_u * last_v;
  const float cu = e.c_pdf[iv * cw + iu], cv = e.m_pdf[iv];
  return denom != 0.f ? (cu * cv) / denom : 0.f;
}

// std::lower_bound over cdf[0..count) followed by max(0, first - 1)
// (optix_common.cuh:72-89).
HM_D int env_lower_bound(const float* cdf, int count, float u) {
  int first = 0;
  while (count > 0) {
    const int step = count >> 1;
    const int mid = first + step;
    if (cdf[mid] < u) {
      first = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return max(0, first - 1);
}

// env_lower_bound(cdf, count, u) through a guide table of kEnvGuideBins + 1
// entries for this cdf.
HM_D int env_lower_bound_guided(const float* cdf, const int* guide, int count, float u) {
  const int k = min(max((int)(u * (float)kEnvGuideBins), 0), kEnvGuideBins - 1);
  int first = guide[k];
  int n = guide[k + 1] - first;  // lower_bound(u) lies in [first, first + n]
  while (n > 0) {
    const int step = n >> 1;
    const int mid = first + step;
    if (cdf[mid] < u) {
      first = mid + 1;
      n -= step + 1;
    } else {
      n = step;
    }
  }
  (void)count;
  return max(0, first - 1);
}

// Importance-samples a direction from uniforms (ux, uy) drawn in that order
// (optix_common.cuh:66-149). Returns radiance; writes direction and pdf.
HM_D V3 env_sample(const EnvView& e, float ux, float uy, V3* wi, float* pdf) {
  if (!e.pdf_sampling) {
    const float z = 1.f - 2.f * ux;
    const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float ph = 2.f * kPi * uy;
    float sp, cp;
    sincos_small(ph, &sp, &cp);
    *wi = normalize(v3(r * cp, r * sp, z));
    *pdf = 1.f / (4.f * kPi);
    return env_radiance(e, *wi);
  }
  const float W = (float)e.width, H = (float)e.height;
  const int cw = e.width + 1;
  const int iv = env_lower_bound_guided(e.m_cdf, e.m_guide, e.height, uy);
  const float cdf_v = e.m_cdf[iv], cdf_nv = e.m_cdf[min(iv + 1, e.height)];
  const float dv = (cdf_nv - uy) / (cdf_nv - cdf_v);  // reference's within-cell map
  const float v = ((float)iv + dv) / H;
  const float* row_cdf = e.c_cdf + iv * cw;
  const int iu = env_lower_bound_guided(row_cdf, e.c_guide + iv * (kEnvGuideBins + 1), e.width, ux);
  const float cdf_u = row_cdf[iu], cdf_nu = row_cdf[min(iu + 1, e.width)];
  const float du = (cdf_nu - ux) / (cdf_nu - cdf_u);
  const float u = ((float)iu + du) / W;
  const V3 L = e.scale * env_bilinear(e, u, v);
  const float last_u = e.c_pdf[iv * cw + e.width];
  const float last_v = e.m_pdf[e.height];
  const float pu = e.c_pdf[iv * cw + iu], pv = e.m_pdf[iv];
  const float theta = kPi * v;
  float sin_t, cos_t;
  sincos_small(theta, &sin_t, &cos_t);
  const float denom = (2.f * kPi * kPi * sin_t) * last_u * last_v;
  *pdf = denom != 0.f ? (pu * pv) / denom : 0.f;
  float phi = 2.f * kPi * u - e.rot_phi;
  if (phi < 0.f) phi = kTwoPiEnv + phi;
  float sin_p, cos_p;
  sincos_small(phi, &sin_p, &cos_p);
  *wi = v3(sin_t * cos_p, sin_t * sin_p, cos_t);
  return L;
}

#endif  // __HIPCC__

}  // namespace hm
