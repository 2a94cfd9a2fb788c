// Chiang-style hair BSDF (longitudinal M_p, azimuthal N_p, attenuation A_p,
// lobes p = 0..2 plus the residual p_max lobe), evaluated in FP32.
//
// Math follows the reference's hair BSDF (cuda_headers/disney_hair.cuh):
//   constants      setupHairShading            disney_hair.cuh:25-41
//   Fresnel        Schlick, F0 from eta        disney_hair.cuh:45-51
//   I0 / log I0    10-term series; x > 12      disney_hair.cuh:53-75
//                  asymptotic form
//   M_p            log branch for v <= 0.1     disney_hair.cuh:77-88
//   A_p            disney_hair.cuh:90-109
//   N_p            trimmed logistic            disney_hair.cuh:111-139
//   lobe pdf       disney_hair.cuh:142-182
//   eval (f*cos)   disney_hair.cuh:185-266
//   sampling       disney_hair.cuh:268-386
//
// Layout for the GPU: everything that depends only on the scene (v, s,
// sin/cos 2^k alpha, the logistic and M_p normalisers) is hoisted into
// HairConsts, built once per scene on the host in double precision. Everything
// that depends only on (wo, h, sigma_a) is computed once per hit in
// hair_prepare() and reused by the light-sample evaluation, the MIS BSDF
// sample and the continuation sample at that vertex.
//
// FP32 accuracy: inside M_p the exponent -b + a - 1/v is rewritten as
// -(1 - cos(theta_i + theta_o'))/v with 1 - cos = S^2/(1 + C), which is the
// same quantity without the cancellation of two ~1/v terms.
#pragma once

#include "hm_common.h"

namespace hm {

constexpr float kEta = 1.55f;       // disney_hair.cuh:10
constexpr float kLn2Ref = 0.6931f;  // truncated ln 2 inside M_p (disney_hair.cuh:83)
constexpr float kHalfLog2Pi = 0.9189385091729444f;  // 0.5*log(2*Pi), Pi = 3.1415926f

struct HairConsts {
  float v[4];          // variance per lobe: R, TT, TRT, TRRT (TRRT uses TRT's)
  float inv_v[4];
  // Variance used when *sampling* lobe p. For p = 3 the reference reads
  // si.hair.v[3] past the end of float v[3] (disney_hair.cuh:329), which is the
  // next member of HairInteraction, the logistic scale s (common.cuh:38-39);
  // the pdf keeps v[2] for that lobe (disney_hair.cuh:252).
  float vs[4];
  float inv_vs[4];
  float mp_log_c[4];   // ln2_ref + log(1/(2v)), the v <= 0.1 branch constant
  float mp_sinh_c[4];  // 1 / (sinh(1/v) * 2v), the v > 0.1 branch constant
  float s, inv_s;      // logistic scale and its inverse
  float trim_norm;     // 1 / (LogisticCDF(pi) - LogisticCDF(-pi))
  float cdf_lo;        // LogisticCDF(-pi)
  float cdf_span;      // LogisticCDF(pi) - LogisticCDF(-pi)
  float sin2k[3], cos2k[3];
  float gain[4];       // R, TT, TRT, TRRT
  float f0;            // Schlick F0 for eta
  float eta;
};

// Per-scene constants (host side). The variance and scale polynomials follow
// disney_hair.cuh:28-33; the 2^k alpha rotations follow :35-40.
inline HairConsts make_hair_consts(float beta_m, float beta_n, float alpha, const float gains[4]) {
  HairConsts c{};
  const double bm = beta_m, bn = beta_n;
  double v0 = 0.726 * bm + 0.812 * bm * bm + 3.7 * std::pow(bm, 20.0);
  v0 = v0 * v0;
  const double vv[4] = {v0, 0.25 * v0, 4.0 * v0, 4.0 * v0};
  for (int p = 0; p < 4; ++p) {
    c.v[p] = (float)vv[p];
    c.inv_v[p] = (float)(1.0 / vv[p]);
    c.mp_log_c[p] = (float)((double)kLn2Ref + std::log(1.0 / (2.0 * vv[p])));
    c.mp_sinh_c[p] = (float)(1.0 / (std::sinh(1.0 / vv[p]) * 2.0 * vv[p]));
    c.gain[p] = gains[p];
  }
  const double s = 0.626657069 * (0.265 * bn + 1.194 * bn * bn + 5.372 * std::pow(bn, 22.0));
  c.s = (float)s;
  c.inv_s = (float)(1.0 / s);
  for (int p = 0; p < 4; ++p) {
    const double vs = p < 3 ? vv[p] : s;
    c.vs[p] = (float)vs;
    c.inv_vs[p] = (float)(1.0 / vs);
  }
  const double pi = (double)kPi;
  const double cdf_hi = 1.0 / (1.0 + std::exp(-pi / s));
  const double cdf_lo = 1.0 / (1.0 + std::exp(pi / s));
  c.cdf_lo = (float)cdf_lo;
  c.cdf_span = (float)(cdf_hi - cdf_lo);
  c.trim_norm = (float)(1.0 / (cdf_hi - cdf_lo));
  double sk[3], ck[3];
  sk[0] = std::sin((double)alpha);
  ck[0] = std::sqrt(std::fmax(0.0, 1.0 - sk[0] * sk[0]));
  for (int i = 1; i < 3; ++i) {
    sk[i] = 2.0 * ck[i - 1] * sk[i - 1];
    ck[i] = ck[i - 1] * ck[i - 1] - sk[i - 1] * sk[i - 1];
  }
  for (int i = 0; i < 3; ++i) {
    c.sin2k[i] = (float)sk[i];
    c.cos2k[i] = (float)ck[i];
  }
  const double eta = (double)kEta;
  c.f0 = (float)(((1.0 - eta) * (1.0 - eta)) / ((1.0 + eta) * (1.0 + eta)));
  c.eta = kEta;
  return c;
}

// Ten-term series x^(2i) / (4^i (i!)^2) (disney_hair.cuh:53-67).
HM_HD float bessel_i0_series(float x) {
  const float x2 = x * x;
  float val = 1.f, x2i = x2;
  val += x2i * 0.25f;                     x2i *= x2;
  val += x2i * 0.015625f;                 x2i *= x2;
  val += x2i * 4.3402777777777775e-4f;    x2i *= x2;
  val += x2i * 6.781684027777777e-6f;     x2i *= x2;
  val += x2i * 6.781684027777778e-8f;     x2i *= x2;
  val += x2i * 4.709502797067901e-10f;    x2i *= x2;
  val += x2i * 2.4028075495244395e-12f;   x2i *= x2;
  val += x2i * 9.385966990329842e-15f;    x2i *= x2;
  val += x2i * 2.896903392077112e-17f;
  return val;
}

HM_HD float sel4(int p, float a0, float a1, float a2, float a3) {
  float r = a3;
  r = p == 2 ? a2 : r;
  r = p == 1 ? a1 : r;
  r = p == 0 ? a0 : r;
  return r;
}

// Longitudinal lobe M_p for cos/sin of theta_i and of the tilted theta_o
// (co_t >= 0). Both branches of disney_hair.cuh:77-88.
HM_HD float hair_mp(const HairConsts& hc, int lobe, float ci, float co_t, float si, float so_t) {
  // value selects, not indexing: the lobe may be a loop counter
  const float v = sel4(lobe, hc.v[0], hc.v[1], hc.v[2], hc.v[3]);
  const float inv_v = sel4(lobe, hc.inv_v[0], hc.inv_v[1], hc.inv_v[2], hc.inv_v[3]);
  const float a = ci * co_t * inv_v;
  if (v <= 0.1f) {
    float expo;
    if (a > 12.f) {
      // log I0(a) ~ a + 0.5(-log 2pi + log(1/a) + 1/(8a)), combined with
      // -b - 1/v:  a - b - 1/v = -(1 - cos(theta_i + theta_o'))/v.
      const float S = si * co_t + ci * so_t;
      const float C = ci * co_t - si * so_t;
      const float one_minus_c = (C > 0.f) ? (S * S) / (1.f + C) : (1.f - C);
      expo = -one_minus_c * inv_v + (-kHalfLog2Pi + 0.5f * (logf(1.f / a) + 1.f / (8.f * a)));
    } else {
      const float b = si * so_t * inv_v;
      expo = logf(bessel_i0_series(a)) - b - inv_v;
    }
    return expf(expo + sel4(lobe, hc.mp_log_c[0], hc.mp_log_c[1], hc.mp_log_c[2], hc.mp_log_c[3]));
  }
  const float b = si * so_t * inv_v;
  return expf(-b) * bessel_i0_series(a) * sel4(lobe, hc.mp_sinh_c[0], hc.mp_sinh_c[1], hc.mp_sinh_c[2], hc.mp_sinh_c[3]);
}

// Trimmed logistic over [-pi, pi] (disney_hair.cuh:116-130).
HM_HD float hair_trimmed_logistic(const HairConsts& hc, float x) {
  const float e = expf(-fabsf(x) * hc.inv_s);
  const float d = 1.f + e;
  return e * hc.inv_s / (d * d) * hc.trim_norm;
}

// Lobe centre Phi(p) = 2p gamma_t - 2 gamma_o + p pi (disney_hair.cuh:111-114).
HM_HD float hair_phi(int p, float gamma_o, float gamma_t) {
  return 2.f * (float)p * gamma_t - 2.f * gamma_o + (float)p * kPi;
}

// Azimuthal lobe N_p (disney_hair.cuh:132-139): wrap into [-pi, pi].
HM_HD float hair_np(const HairConsts& hc, float phi, int p, float gamma_o, float gamma_t) {
  float dphi = phi - hair_phi(p, gamma_o, gamma_t);
  // |dphi| < 8 pi for finite inputs, so eight steps always finish the wrap;
  // the bound only guarantees termination for non-finite inputs.
  for (int k = 0; k < 8 && dphi > kPi; ++k) dphi -= 2.f * kPi;
  for (int k = 0; k < 8 && dphi < -kPi; ++k) dphi += 2.f * kPi;
  return hair_trimmed_logistic(hc, dphi);
}

// Everything at a hit that depends on wo, h and sigma_a only. The tilted
// theta_o of each lobe and the lobe probabilities are recomputed where they
// are used (a few multiplies) rather than held in registers through the
// whole shading of a vertex.
struct HairWo {
  float so, co, phi_o;     // sin/cos theta_o, phi_o
  float gamma_o, gamma_t, h;
  V3 ap[4];                // attenuation A_p, RGB
  float inv_ap_sum;        // 1 / sum_p mean-RGB A_p
};

// sin and cos of theta_o tilted for lobe p: lobe 0 rotates by -2 alpha,
// lobe 1 by +alpha, lobe 2 by +4 alpha (disney_hair.cuh:228-246); lobe 3 is
// not tilted. cos is signed (sampling uses the sign, evaluation |cos|).
HM_HD void hair_tilt(const HairConsts& hc, const HairWo& w, int p, float* so_t, float* co_t) {
  const float c = sel4(p, hc.cos2k[1], hc.cos2k[0], hc.cos2k[2], 1.f);
  const float s = sel4(p, -hc.sin2k[1], hc.sin2k[0], hc.sin2k[2], 0.f);
  *so_t = w.so * c + w.co * s;
  *co_t = w.co * c - w.so * s;
}

// Probability of lobe p: its mean-RGB A_p over the sum over the lobes
// (disney_hair.cuh:142-182).
HM_HD float hair_ap_pdf(const HairWo& w, int p) {
  const float x = sel4(p, w.ap[0].x, w.ap[1].x, w.ap[2].x, w.ap[3].x);
  const float y = sel4(p, w.ap[0].y, w.ap[1].y, w.ap[2].y, w.ap[3].y);
  const float z = sel4(p, w.ap[0].z, w.ap[1].z, w.ap[2].z, w.ap[3].z);
  return ((x + y + z) * (1.f / 3.f)) * w.inv_ap_sum;
}

HM_HD HairWo hair_prepare(const HairConsts& hc, V3 wo, float h, V3 sigma_a) {
  HairWo w;
  w.so = wo.x;
  w.co = safe_sqrt((1.f - w.so) * (1.f + w.so));
  w.phi_o = atan2f(wo.z, wo.y);
  w.h = h;
  w.gamma_o = asinf(h);
  // 1 - x^2 is formed as (1 - x)(1 + x), which keeps its relative accuracy
  // as |x| -> 1 (grazing directions, silhouette hits)
  const float cos_gamma_o = safe_sqrt((1.f - h) * (1.f + h));  // = cos(asin(h))
  const float sin_theta_t = w.so / hc.eta;
  const float cos_theta_t = safe_sqrt((1.f - sin_theta_t) * (1.f + sin_theta_t));
  const float etap = sqrtf((hc.eta - w.so) * (hc.eta + w.so)) / w.co;
  const float sin_gamma_t = h / etap;
  const float cos_gamma_t = safe_sqrt((1.f - sin_gamma_t) * (1.f + sin_gamma_t));
  w.gamma_t = asinf(clampf(sin_gamma_t, -1.f, 1.f));
  const float fac = 2.f * cos_gamma_t / cos_theta_t;
  const V3 T = v3(expf(-sigma_a.x * fac), expf(-sigma_a.y * fac), expf(-sigma_a.z * fac));
  const float xd = 1.f - w.co * cos_gamma_o;
  const float xd2 = xd * xd;
  const float F = hc.f0 + (1.f - hc.f0) * (xd2 * xd2 * xd);
  const float omf2 = (1.f - F) * (1.f - F);
  w.ap[0] = v3(F);
  w.ap[1] = omf2 * T;
  w.ap[2] = (omf2 * F) * (T * T);
  w.ap[3] = w.ap[2] * (F * T) / (v3(1.f) - F * T);
  float sum = 0.f;
#pragma unroll
  for (int p = 0; p < 4; ++p) sum += (w.ap[p].x + w.ap[p].y + w.ap[p].z) * (1.f / 3.f);
  w.inv_ap_sum = 1.f / sum;
  return w;
}

// f * |cos theta_i| (RGB) and the lobe-mixture pdf for direction wi, both
// in the hair frame (disney_hair.cuh:185-266). Returned by value: output
// pointers into caller locals defeat register promotion when the optimiser
// merges call sites.
struct FPdf {
  V3 f;
  float pdf;
};

HM_HD FPdf hair_eval_fp(const HairConsts& hc, const HairWo& w, V3 wi) {
  const float si = wi.x;
  const float ci = safe_sqrt((1.f - si) * (1.f + si));
  const float phi = atan2f(wi.z, wi.y) - w.phi_o;
  V3 f = v3(0.f);
  float pd = 0.f;
  // One copy of the lobe code, run for p = 0..3 (the residual lobe p = 3
  // has the untilted theta_o and N = 1 / 2 pi): the shading kernels inline
  // this evaluation at several places, and an unrolled body per lobe made
  // them larger than the instruction cache.
#pragma unroll 1
  for (int p = 0; p < 4; ++p) {
    float so_t, co_t;
    hair_tilt(hc, w, p, &so_t, &co_t);
    const float mp = hair_mp(hc, p, ci, fabsf(co_t), si, so_t);
    const float np = p < 3 ? hair_np(hc, phi, p, w.gamma_o, w.gamma_t) : 1.f / (2.f * kPi);
    const float gain = sel4(p, hc.gain[0], hc.gain[1], hc.gain[2], hc.gain[3]);
    const V3 ap = v3(sel4(p, w.ap[0].x, w.ap[1].x, w.ap[2].x, w.ap[3].x),
                     sel4(p, w.ap[0].y, w.ap[1].y, w.ap[2].y, w.ap[3].y),
                     sel4(p, w.ap[0].z, w.ap[1].z, w.ap[2].z, w.ap[3].z));
    const float ap_pdf = hair_ap_pdf(w, p);
    if (p < 3) {
      const float mn = mp * np;
      f += (gain * mn) * ap;
      pd += mn * ap_pdf;
    } else {
      f += (gain * mp * np) * ap;
      pd += mp * ap_pdf * np;
    }
  }
  return FPdf{f, pd};
}

HM_HD V3 hair_eval(const HairConsts& hc, const HairWo& w, V3 wi, float* pdf) {
  const FPdf r = hair_eval_fp(hc, w, wi);
  *pdf = r.pdf;
  return r.f;
}

// Importance-sample wi from four uniforms (disney_hair.cuh:276-386). Returns
// the sampled direction in the hair frame and writes the chosen lobe.
HM_HD V3 hair_sample_dir(const HairConsts& hc, const HairWo& w, float u0, float u1, float u2,
                         float u3, int* lobe_out) {
  // lobe choice (disney_hair.cuh:300-305), written without a loop so every
  // array index is a constant
  int p;
  float e1 = u0;
  const float q0 = hair_ap_pdf(w, 0);
  if (e1 < q0) {
    p = 0;
  } else {
    e1 -= q0;
    const float q1 = hair_ap_pdf(w, 1);
    if (e1 < q1) {
      p = 1;
    } else {
      e1 -= q1;
      p = e1 < hair_ap_pdf(w, 2) ? 2 : 3;
    }
  }
  // value selects (not lvalue conditionals, not indexing) keep the per-lobe
  // values in registers; sampling uses the signed tilted cosine
  float so_t, co_t;
  hair_tilt(hc, w, p, &so_t, &co_t);
  const float v = sel4(p, hc.vs[0], hc.vs[1], hc.vs[2], hc.vs[3]);
  const float inv_v = sel4(p, hc.inv_vs[0], hc.inv_vs[1], hc.inv_vs[2], hc.inv_vs[3]);
  const float e2 = fmaxf(u1, 1e-5f);
  const float cos_t = 1.f + v * logf(e2 + (1.f - e2) * expf(-2.f * inv_v));
  const float sin_t = safe_sqrt((1.f - cos_t) * (1.f + cos_t));
  const float cos_phi = cos_small(2.f * kPi * u2);
  const float si = -cos_t * so_t + sin_t * cos_phi * co_t;
  const float ci = safe_sqrt((1.f - si) * (1.f + si));
  float dphi;
  if (p < 3) {
    // Trimmed-logistic inversion (disney_hair.cuh:268-274).
    float x = -hc.s * logf(1.f / (u3 * hc.cdf_span + hc.cdf_lo) - 1.f);
    x = clampf(x, -kPi, kPi);
    dphi = hair_phi(p, w.gamma_o, w.gamma_t) + x;
  } else {
    dphi = 2.f * kPi * u3;
  }
  const float phi_i = w.phi_o + dphi;
  *lobe_out = p;
  float sin_pi, cos_pi;
  sincos_small(phi_i, &sin_pi, &cos_pi);
  return normalize(v3(si, ci * cos_pi, ci * sin_pi));
}

// Hair shading frame at a hit (optix_common.cuh:571-583): rows X = tangent,
// Y = normalize(wo x X), Z = normalize(X x Y). h = dot(Y, n) (disney_hair.cuh:206).
HM_HD Frame hair_frame(V3 tangent, V3 wo) {
  const V3 X = tangent;
  const V3 Y = normalize(cross(wo, X));
  const V3 Z = normalize(cross(X, Y));
  return Frame{X, Y, Z};
}

}  // namespace hm
