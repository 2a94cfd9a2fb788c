// Forward-mode derivatives of the hair BSDF value f * |cos| with respect to
// the appearance parameters (beta_m, beta_n, alpha, sigma_a r/g/b), for the
// path-replay adjoint.
//
// Directions, h and every sampling decision are held fixed (detached), so
// only f itself is differentiated: through the lobe variances v (beta_m), the
// logistic scale s (beta_n), the scale-tilt rotations (alpha) and the
// transmittance T (sigma_a). The formulas are those of hm_hair_bsdf.h
// (disney_hair.cuh:25-266) carried on dual numbers.
#pragma once

#include "hm_hair_bsdf.h"

namespace hm {

constexpr int kAdParams = 6;  // beta_m, beta_n, alpha, sigma_r, sigma_g, sigma_b

struct Dual {
  float v;
  float d[kAdParams];
};

HM_HD Dual dconst(float x) {
  Dual r;
  r.v = x;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = 0.f;
  return r;
}
HM_HD Dual operator+(Dual a, Dual b) {
  Dual r;
  r.v = a.v + b.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
HM_HD Dual operator-(Dual a, Dual b) {
  Dual r;
  r.v = a.v - b.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
HM_HD Dual operator-(Dual a) {
  Dual r;
  r.v = -a.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = -a.d[k];
  return r;
}
HM_HD Dual operator*(Dual a, Dual b) {
  Dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
HM_HD Dual operator*(float s, Dual a) {
  Dual r;
  r.v = s * a.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = s * a.d[k];
  return r;
}
HM_HD Dual operator*(Dual a, float s) { return s * a; }
HM_HD Dual operator+(Dual a, float s) {
  a.v += s;
  return a;
}
HM_HD Dual operator+(float s, Dual a) { return a + s; }
HM_HD Dual operator-(float s, Dual a) { return (-a) + s; }
HM_HD Dual operator-(Dual a, float s) {
  a.v -= s;
  return a;
}
HM_HD Dual operator/(Dual a, Dual b) {
  Dual r;
  const float ib = 1.f / b.v;
  r.v = a.v * ib;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
HM_HD Dual operator/(float s, Dual b) { return dconst(s) / b; }
HM_HD Dual operator/(Dual a, float s) { return (1.f / s) * a; }
HM_HD Dual dexp(Dual a) {
  Dual r;
  r.v = expf(a.v);
  for (int k = 0; k < kAdParams; ++k) r.d[k] = r.v * a.d[k];
  return r;
}
HM_HD Dual dlog(Dual a) {
  Dual r;
  r.v = logf(a.v);
  const float ia = 1.f / a.v;
  for (int k = 0; k < kAdParams; ++k) r.d[k] = a.d[k] * ia;
  return r;
}
HM_HD Dual dfabs(Dual a) { return a.v < 0.f ? -a : a; }

// Scene constants with their tangents w.r.t. (beta_m, beta_n, alpha); the
// sigma_a tangents of all of them are zero.
struct HairConstsAD {
  Dual v[3];        // R, TT, TRT (TRRT uses v[2])
  Dual s;
  Dual sin2k[3], cos2k[3];
  float gain[4];
  float f0, eta;
};

// Host-side: derivatives of the polynomials of disney_hair.cuh:28-40 in
// double precision.
inline HairConstsAD make_hair_consts_ad(float beta_m, float beta_n, float alpha, const float gains[4]) {
  HairConstsAD c{};
  const double bm = beta_m, bn = beta_n;
  const double pm = 0.726 * bm + 0.812 * bm * bm + 3.7 * std::pow(bm, 20.0);
  const double dpm = 0.726 + 1.624 * bm + 74.0 * std::pow(bm, 19.0);
  const double v0 = pm * pm, dv0 = 2.0 * pm * dpm;
  const double vs[3] = {v0, 0.25 * v0, 4.0 * v0}, dvs[3] = {dv0, 0.25 * dv0, 4.0 * dv0};
  for (int p = 0; p < 3; ++p) {
    c.v[p] = dconst((float)vs[p]);
    c.v[p].d[0] = (float)dvs[p];
  }
  const double k = 0.626657069;
  const double s = k * (0.265 * bn + 1.194 * bn * bn + 5.372 * std::pow(bn, 22.0));
  const double ds = k * (0.265 + 2.388 * bn + 118.184 * std::pow(bn, 21.0));
  c.s = dconst((float)s);
  c.s.d[1] = (float)ds;
  // sin/cos of alpha, 2 alpha, 4 alpha and their alpha derivatives
  const double a = alpha;
  const double mult[3] = {1.0, 2.0, 4.0};
  for (int i = 0; i < 3; ++i) {
    c.sin2k[i] = dconst((float)std::sin(mult[i] * a));
    c.sin2k[i].d[2] = (float)(mult[i] * std::cos(mult[i] * a));
    c.cos2k[i] = dconst((float)std::cos(mult[i] * a));
    c.cos2k[i].d[2] = (float)(-mult[i] * std::sin(mult[i] * a));
  }
  for (int p = 0; p < 4; ++p) c.gain[p] = gains[p];
  const double eta = (double)kEta;
  c.f0 = (float)(((1.0 - eta) * (1.0 - eta)) / ((1.0 + eta) * (1.0 + eta)));
  c.eta = kEta;
  return c;
}

HM_HD Dual bessel_i0_dual(Dual x) {
  const float c[10] = {1.f, 0.25f, 0.015625f, 4.3402777777777775e-4f, 6.781684027777777e-6f,
                       6.781684027777778e-8f, 4.709502797067901e-10f, 2.4028075495244395e-12f,
                       9.385966990329842e-15f, 2.896903392077112e-17f};
  const Dual x2 = x * x;
  Dual val = dconst(c[9]);
  for (int i = 8; i >= 0; --i) val = val * x2 + c[i];  // Horner in x^2
  return val;
}

// M_p with dual variance and dual tilted angles (disney_hair.cuh:77-88).
HM_HD Dual hair_mp_dual(float ci, Dual co_t, float si, Dual so_t, Dual v) {
  const Dual inv_v = 1.f / v;
  const Dual a = (ci * co_t) * inv_v;
  const Dual b = (si * so_t) * inv_v;
  if (v.v <= 0.1f) {
    Dual lg;
    if (a.v > 12.f) lg = a + 0.5f * (dconst(-2.f * kHalfLog2Pi) + dlog(1.f / a) + 1.f / (8.f * a));
    else lg = dlog(bessel_i0_dual(a));
    return dexp(lg - b - inv_v + kLn2Ref + dlog(1.f / (2.f * v)));
  }
  const Dual e = dexp(inv_v);
  const Dual sinh_iv = 0.5f * (e - 1.f / e);
  return dexp(-b) * bessel_i0_dual(a) / (sinh_iv * 2.f * v);
}

// Trimmed logistic N_p with a dual scale (disney_hair.cuh:116-139); dphi is
// already wrapped into [-pi, pi] and is parameter-free.
HM_HD Dual hair_np_dual(float dphi, Dual s) {
  const Dual inv_s = 1.f / s;
  const Dual e = dexp(-fabsf(dphi) * inv_s);
  const Dual d = 1.f + e;
  const Dual logistic = e * inv_s / (d * d);
  const Dual cdf_hi = 1.f / (1.f + dexp(-kPi * inv_s));
  const Dual cdf_lo = 1.f / (1.f + dexp(kPi * inv_s));
  return logistic / (cdf_hi - cdf_lo);
}

HM_HD float wrap_pi(float dphi) {
  for (int k = 0; k < 8 && dphi > kPi; ++k) dphi -= 2.f * kPi;
  for (int k = 0; k < 8 && dphi < -kPi; ++k) dphi += 2.f * kPi;
  return dphi;
}

// f * |cos theta_i| per channel and the lobe-mixture pdf, as duals. wo/wi in
// the hair frame, h and sigma_a as in hair_prepare. The pdf carries sigma_a
// tangents through the lobe weights A_p / sum A_p (mean over RGB) and the
// roughness / tilt tangents through M_p N_p.
HM_HD void hair_eval_dual(const HairConstsAD& hc, V3 wo, V3 wi, float h, V3 sigma, Dual f[3], Dual* pdf) {
  const float so = wo.x, co = safe_sqrt((1.f - so) * (1.f + so));
  const float phi_o = atan2f(wo.z, wo.y);
  const float si = wi.x, ci = safe_sqrt((1.f - si) * (1.f + si));
  const float phi = atan2f(wi.z, wi.y) - phi_o;
  const float gamma_o = asinf(h);
  const float cos_gamma_o = safe_sqrt((1.f - h) * (1.f + h));
  const float sin_theta_t = so / hc.eta;
  const float cos_theta_t = safe_sqrt((1.f - sin_theta_t) * (1.f + sin_theta_t));
  const float etap = sqrtf((hc.eta - so) * (hc.eta + so)) / co;
  const float sin_gamma_t = h / etap;
  const float cos_gamma_t = safe_sqrt((1.f - sin_gamma_t) * (1.f + sin_gamma_t));
  const float gamma_t = asinf(clampf(sin_gamma_t, -1.f, 1.f));
  const float fac = 2.f * cos_gamma_t / cos_theta_t;
  const float xd = 1.f - co * cos_gamma_o;
  const float xd2 = xd * xd;
  const float F = hc.f0 + (1.f - hc.f0) * (xd2 * xd2 * xd);
  const float omf2 = (1.f - F) * (1.f - F);
  // tilted theta_o per lobe (disney_hair.cuh:228-246)
  Dual so_t[3], co_t[3];
  so_t[0] = so * hc.cos2k[1] - co * hc.sin2k[1];
  co_t[0] = dfabs(co * hc.cos2k[1] + so * hc.sin2k[1]);
  so_t[1] = so * hc.cos2k[0] + co * hc.sin2k[0];
  co_t[1] = dfabs(co * hc.cos2k[0] - so * hc.sin2k[0]);
  so_t[2] = so * hc.cos2k[2] + co * hc.sin2k[2];
  co_t[2] = dfabs(co * hc.cos2k[2] - so * hc.sin2k[2]);
  Dual mn[4];
  for (int p = 0; p < 3; ++p) {
    const Dual mp = hair_mp_dual(ci, co_t[p], si, so_t[p], hc.v[p]);
    const float dphi = wrap_pi(phi - hair_phi(p, gamma_o, gamma_t));
    mn[p] = mp * hair_np_dual(dphi, hc.s);
  }
  mn[3] = hair_mp_dual(ci, dconst(co), si, dconst(so), hc.v[2]) * (1.f / (2.f * kPi));
  // attenuation A_p per channel (disney_hair.cuh:142-183)
  const float sig[3] = {sigma.x, sigma.y, sigma.z};
  Dual m[4];
  for (int p = 0; p < 4; ++p) m[p] = dconst(0.f);
  for (int c = 0; c < 3; ++c) {
    Dual T = dconst(expf(-sig[c] * fac));
    T.d[3 + c] = -fac * T.v;  // d T_c / d sigma_c
    const Dual A0 = dconst(F);
    const Dual A1 = omf2 * T;
    const Dual A2 = (omf2 * F) * (T * T);
    const Dual A3 = A2 * (F * T) / (1.f - F * T);
    f[c] = (hc.gain[0] * mn[0]) * A0 + (hc.gain[1] * mn[1]) * A1 + (hc.gain[2] * mn[2]) * A2 +
           (hc.gain[3] * mn[3]) * A3;
    m[0] = m[0] + A0;
    m[1] = m[1] + A1;
    m[2] = m[2] + A2;
    m[3] = m[3] + A3;
  }
  const Dual inv_sum = 1.f / (m[0] + m[1] + m[2] + m[3]);
  *pdf = mn[0] * (m[0] * inv_sum) + mn[1] * (m[1] * inv_sum) + mn[2] * (m[2] * inv_sum) + mn[3] * (m[3] * inv_sum);
}

}  // namespace hm
