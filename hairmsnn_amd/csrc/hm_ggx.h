// Surface BSDF of the head proxy: 0.5 * Lambert + 0.5 * isotropic GGX
// (Heitz), with visible-normal sampling. Returns f * |cos theta_i|.
//
// Follows cuda_headers/frostbite_anisotropic.cuh:
//   D, Lambda, G1, G2, Dv        :26-79
//   GGX and frostbite_GGX        :81-100
//   sample_VNDF                  :102-131
//   pdf_GGX, sample_GGX          :133-157
// All vectors are in the local frame with the normal on +z.
#pragma once

#include "hm_common.h"

namespace hm {

HM_HD float ggx_d(float a, V3 n) {
  const float t1 = n.x / a, t2 = n.y / a, t3 = n.z;
  const float q = t1 * t1 + t2 * t2 + t3 * t3;
  return 1.f / (kPi * a * a * (q * q));
}

HM_HD float ggx_lambda(float a, V3 v) {
  const float t1 = v.x * a, t2 = v.y * a, t3 = v.z;
  return 0.5f * (-1.f + sqrtf(1.f + (t1 * t1 + t2 * t2) / (t3 * t3)));
}

HM_HD float ggx_g1(float a, V3 v) { return v.z <= 0.f ? 0.f : 1.f / (1.f + ggx_lambda(a, v)); }

HM_HD float ggx_g2(float a, V3 v, V3 l) {
  if (v.z <= 0.f || l.z <= 0.f) return 0.f;
  return 1.f / (1.f + ggx_lambda(a, v) + ggx_lambda(a, l));
}

// Visible-normal density D_v(Ne) (frostbite_anisotropic.cuh:61-69).
HM_HD float ggx_dv(float a, V3 v, V3 ne) {
  return ggx_g1(a, v) * fmaxf(0.f, dot(v, ne)) * ggx_d(a, ne) / v.z;
}

// f * |cos| of the half-Lambert, half-GGX surface (frostbite_anisotropic.cuh:84-100).
HM_HD V3 surface_eval(V3 wo, V3 wi, V3 albedo, float a) {
  V3 f = v3(0.f);
  if (wo.z > 0.f && wi.z > 0.f) {
    const V3 h = normalize(wo + wi);
    const float spec = ggx_d(a, h) * ggx_g2(a, wo, wi) / 4.f / wo.z / wi.z;
    f = (0.5f / kPi) * albedo + v3(0.5f * spec);
  }
  return f * fabsf(wi.z);
}

// Sample a visible normal (frostbite_anisotropic.cuh:102-131).
HM_HD V3 ggx_sample_vndf(float a, V3 v, float u1, float u2) {
  const V3 vh = normalize(v * v3(a, a, 1.f));
  const float lensq = vh.x * vh.x + vh.y * vh.y;
  const V3 t1v = lensq > 0.f ? v3(-vh.y, vh.x, 0.f) / sqrtf(lensq) : v3(1.f, 0.f, 0.f);
  const V3 t2v = cross(vh, t1v);
  const float r = sqrtf(u1);
  const float phi = 2.f * kPi * u2;
  float sin_p, cos_p;
  sincos_small(phi, &sin_p, &cos_p);
  const float t1 = r * cos_p;
  float t2 = r * sin_p;
  const float s = 0.5f * (1.f + vh.z);
  t2 = (1.f - s) * sqrtf(1.f - t1 * t1) + s * t2;
  const V3 nh = t1 * t1v + t2 * t2v + sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2)) * vh;
  return normalize(v3(a * nh.x, a * nh.y, fmaxf(0.f, nh.z)));
}

// pdf of the reflected direction for visible normal ne (frostbite_anisotropic.cuh:133-141).
HM_HD float surface_pdf(float a, V3 v, V3 ne) { return ggx_dv(a, v, ne) / (4.f * dot(v, ne)); }

// Reflect about a sampled visible normal (frostbite_anisotropic.cuh:143-157).
struct DirPdf {
  V3 dir;
  float pdf;
};
HM_HD DirPdf surface_sample_dp(float a, V3 v, float u1, float u2) {
  const V3 n = ggx_sample_vndf(a, v, u1, u2);
  const V3 l = -v + (2.f * dot(v, n)) * n;
  return DirPdf{normalize(l), surface_pdf(a, v, n)};
}
HM_HD V3 surface_sample(float a, V3 v, float u1, float u2, float* pdf) {
  const DirPdf r = surface_sample_dp(a, v, u1, u2);
  *pdf = r.pdf;
  return r.dir;
}

}  // namespace hm
