// ORACLE - test infrastructure only. Nothing in the product library links,
// imports or calls this code; tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it as the checker and as the CPU baseline.
//
// CPU restatement of the reference's device math, templated on the real type
// so one source gives the FP64 oracle and an FP32 twin. Each function names
// the reference lines it restates.
#pragma once

#include <cmath>
#include <cstdint>

namespace orc {

template <class R>
struct Vec {
  R x, y, z;
};
template <class R> inline Vec<R> mk(R x, R y, R z) { return Vec<R>{x, y, z}; }
template <class R> inline Vec<R> operator+(Vec<R> a, Vec<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline Vec<R> operator-(Vec<R> a, Vec<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline Vec<R> operator*(Vec<R> a, Vec<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline Vec<R> operator*(R s, Vec<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline Vec<R> operator/(Vec<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline Vec<R> operator/(Vec<R> a, Vec<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline R dot(Vec<R> a, Vec<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline Vec<R> cross(Vec<R> a, Vec<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R len(Vec<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline Vec<R> nrm(Vec<R> a) { return a / len(a); }
template <class R> inline bool has_nan(Vec<R> a) { return std::isnan(a.x) || std::isnan(a.y) || std::isnan(a.z); }
template <class R> inline bool has_inf(Vec<R> a) { return std::isinf(a.x) || std::isinf(a.y) || std::isinf(a.z); }

// utils.cuh:10 Pi = 3.1415926f, utils.cuh:11 TWO_Pi = 2.f * 3.14159f
template <class R> inline R Pi() { return (R)3.1415926f; }
template <class R> inline R TwoPiEnv() { return (R)(2.f * 3.14159f); }

// ---- RNG: lcg_random.cuh:15-66 -------------------------------------------
inline uint32_t mm_mix(uint32_t hash, uint32_t k) {
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  k *= c1;
  k = (k << 15) | (k >> 17);
  k *= c2;
  hash ^= k;
  hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
  return hash;
}
inline uint32_t mm_fin(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
inline uint32_t get_rng(uint32_t frame, uint32_t px, uint32_t py, uint32_t w) {
  uint32_t s = mm_mix(0u, px + py * w);
  s = mm_mix(s, frame);
  return mm_fin(s);
}
struct Lcg {
  uint32_t state;
  uint32_t next_u32() {
    state = state * 1664525u + 1013904223u;
    return state;
  }
  float next() { return std::ldexp((float)next_u32(), -32); }
};

// ---- hair BSDF: disney_hair.cuh -------------------------------------------
template <class R>
struct HairParams {
  R beta_m, beta_n, alpha;
  R gR, gTT, gTRT, gTRRT;
};

template <class R>
struct HairShading {  // setupHairShading, disney_hair.cuh:25-41
  R v[3];
  R s;
  R sin2k[3], cos2k[3];
};

template <class R> inline R Sqr(R v) { return v * v; }
template <class R> inline R SafeSqrt(R x) { return std::sqrt(std::fmax((R)0, x)); }
template <class R> inline R SafeASin(R x) { return std::asin(std::fmin((R)1, std::fmax((R)-1, x))); }

template <class R>
HairShading<R> setup_hair(const HairParams<R>& p) {
  HairShading<R> h;
  const R bm = p.beta_m, bn = p.beta_n;
  h.v[0] = Sqr((R)0.726f * bm + (R)0.812f * Sqr(bm) + (R)3.7f * std::pow(bm, (R)20));
  h.v[1] = (R)0.25 * h.v[0];
  h.v[2] = (R)4 * h.v[0];
  h.s = (R)0.626657069f * ((R)0.265f * bn + (R)1.194f * Sqr(bn) + (R)5.372f * std::pow(bn, (R)22));
  h.sin2k[0] = std::sin(p.alpha);
  h.cos2k[0] = SafeSqrt((R)1 - Sqr(h.sin2k[0]));
  for (int i = 1; i < 3; ++i) {
    h.sin2k[i] = (R)2 * h.cos2k[i - 1] * h.sin2k[i - 1];
    h.cos2k[i] = Sqr(h.cos2k[i - 1]) - Sqr(h.sin2k[i - 1]);
  }
  return h;
}

template <class R> inline R Fresnel(R eta, R cosTheta) {  // :45-51
  const R F0 = Sqr((R)1 - eta) / Sqr((R)1 + eta);
  const R xd = (R)1 - cosTheta;
  return F0 + ((R)1 - F0) * xd * xd * xd * xd * xd;
}
template <class R> inline R I0(R x) {  // :53-67
  R val = 0, x2i = 1;
  int64_t ifact = 1;
  int i4 = 1;
  for (int i = 0; i < 10; ++i) {
    if (i > 1) ifact *= i;
    val += x2i / ((R)i4 * Sqr((R)ifact));
    x2i *= x * x;
    i4 *= 4;
  }
  return val;
}
template <class R> inline R LogI0(R x) {  // :69-75
  if (x > 12) return x + (R)0.5 * (-std::log((R)2 * Pi<R>()) + std::log((R)1 / x) + (R)1 / ((R)8 * x));
  return std::log(I0(x));
}
template <class R> inline R Mp(R cI, R cO, R sI, R sO, R v) {  // :77-88
  const R a = cI * cO / v, b = sI * sO / v;
  if (v <= (R)0.1f) return std::exp(LogI0(a) - b - (R)1 / v + (R)0.6931f + std::log((R)1 / ((R)2 * v)));
  return (std::exp(-b) * I0(a)) / (std::sinh((R)1 / v) * (R)2 * v);
}
template <class R> inline R PhiP(int p, R gO, R gT) { return (R)2 * (R)p * gT - (R)2 * gO + (R)p * Pi<R>(); }
template <class R> inline R Logistic(R x, R s) {
  x = std::fabs(x);
  return std::exp(-x / s) / (s * Sqr((R)1 + std::exp(-x / s)));
}
template <class R> inline R LogisticCDF(R x, R s) { return (R)1 / ((R)1 + std::exp(-x / s)); }
template <class R> inline R TrimmedLogistic(R x, R s, R a, R b) {
  return Logistic(x, s) / (LogisticCDF(b, s) - LogisticCDF(a, s));
}
template <class R> inline R Np(R phi, int p, R s, R gO, R gT) {  // :132-139
  R dphi = phi - PhiP(p, gO, gT);
  while (dphi > Pi<R>()) dphi -= (R)2 * Pi<R>();
  while (dphi < -Pi<R>()) dphi += (R)2 * Pi<R>();
  return TrimmedLogistic(dphi, s, -Pi<R>(), Pi<R>());
}

template <class R>
struct HairQuery {
  Vec<R> wo, wi;
  R h;
  Vec<R> sigma_a;
};

// A_p per lobe (:90-109) and the normalised lobe pdf (:142-182).
template <class R>
void attenuation(R sinThetaO, R cosThetaO, R cosGammaO, R h, Vec<R> sig, Vec<R> ap[4], R ap_pdf[4]) {
  const R eta = (R)1.55f;
  const R sinThetaT = sinThetaO / eta;
  const R cosThetaT = SafeSqrt((R)1 - Sqr(sinThetaT));
  const R etap = std::sqrt(eta * eta - Sqr(sinThetaO)) / cosThetaO;
  const R sinGammaT = h / etap;
  const R cosGammaT = SafeSqrt((R)1 - Sqr(sinGammaT));
  const R fac = (R)2 * cosGammaT / cosThetaT;
  const Vec<R> T = mk(std::exp(-sig.x * fac), std::exp(-sig.y * fac), std::exp(-sig.z * fac));
  const R F = Fresnel(eta, cosThetaO * cosGammaO);
  ap[0] = mk(F, F, F);
  ap[1] = Sqr((R)1 - F) * T;
  ap[2] = (Sqr((R)1 - F) * T * T) * mk(F, F, F);
  ap[3] = ap[2] * mk(F, F, F) * T / (mk((R)1, (R)1, (R)1) - T * mk(F, F, F));
  R sum = 0;
  for (int p = 0; p < 4; ++p) {
    ap_pdf[p] = (ap[p].x + ap[p].y + ap[p].z) / (R)3;
    sum += ap_pdf[p];
  }
  for (int p = 0; p < 4; ++p) ap_pdf[p] /= sum;
}

// f * cos and pdf (disney_hair, :185-266).
template <class R>
Vec<R> hair_eval(const HairParams<R>& prm, const HairShading<R>& hs, const HairQuery<R>& q, R* pdf) {
  const Vec<R> wo = q.wo, wi = q.wi;
  const R sinThetaO = wo.x, cosThetaO = SafeSqrt((R)1 - Sqr(sinThetaO));
  const R phiO = std::atan2(wo.z, wo.y);
  const R sinThetaI = wi.x, cosThetaI = SafeSqrt((R)1 - Sqr(sinThetaI));
  const R phiI = std::atan2(wi.z, wi.y);
  const R eta = (R)1.55f;
  const R h = q.h;
  const R gammaO = std::asin(h);
  const R cosGammaO = std::cos(gammaO);
  const R etap = std::sqrt(eta * eta - Sqr(sinThetaO)) / cosThetaO;
  const R sinGammaT = h / etap;
  const R gammaT = SafeASin(sinGammaT);
  Vec<R> ap[4];
  R ap_pdf[4];
  attenuation(sinThetaO, cosThetaO, cosGammaO, h, q.sigma_a, ap, ap_pdf);
  const R phi = phiI - phiO;
  Vec<R> f = mk((R)0, (R)0, (R)0);
  R sOp, cOp;
  // R lobe: rotate by -2 alpha (:228-234)
  sOp = sinThetaO * hs.cos2k[1] - cosThetaO * hs.sin2k[1];
  cOp = std::fabs(cosThetaO * hs.cos2k[1] + sinThetaO * hs.sin2k[1]);
  const R Mp0 = Mp(cosThetaI, cOp, sinThetaI, sOp, hs.v[0]);
  const R Np0 = Np(phi, 0, hs.s, gammaO, gammaT);
  f = f + (prm.gR * Mp0 * Np0) * ap[0];
  // TT lobe: rotate by +alpha (:236-242)
  sOp = sinThetaO * hs.cos2k[0] + cosThetaO * hs.sin2k[0];
  cOp = std::fabs(cosThetaO * hs.cos2k[0] - sinThetaO * hs.sin2k[0]);
  const R Mp1 = Mp(cosThetaI, cOp, sinThetaI, sOp, hs.v[1]);
  const R Np1 = Np(phi, 1, hs.s, gammaO, gammaT);
  f = f + (prm.gTT * Mp1 * Np1) * ap[1];
  // TRT lobe: rotate by +4 alpha (:244-250)
  sOp = sinThetaO * hs.cos2k[2] + cosThetaO * hs.sin2k[2];
  cOp = std::fabs(cosThetaO * hs.cos2k[2] - sinThetaO * hs.sin2k[2]);
  const R Mp2 = Mp(cosThetaI, cOp, sinThetaI, sOp, hs.v[2]);
  const R Np2 = Np(phi, 2, hs.s, gammaO, gammaT);
  f = f + (prm.gTRT * Mp2 * Np2) * ap[2];
  // residual lobe (:252-254)
  const R MpM = Mp(cosThetaI, cosThetaO, sinThetaI, sinThetaO, hs.v[2]);
  const R NpM = (R)1 / ((R)2 * Pi<R>());
  f = f + (prm.gTRRT * MpM * NpM) * ap[3];
  *pdf = Mp0 * ap_pdf[0] * Np0 + Mp1 * ap_pdf[1] * Np1 + Mp2 * ap_pdf[2] * Np2 + MpM * ap_pdf[3] * NpM;
  return f;
}

// sample_disney_hair (:268-386): returns wi in the hair frame and the lobe.
template <class R>
Vec<R> hair_sample(const HairShading<R>& hs, Vec<R> wo, R h, Vec<R> sig, R u0, R u1, R u2, R u3, int* lobe) {
  const R sinThetaO = wo.x, cosThetaO = SafeSqrt((R)1 - Sqr(sinThetaO));
  const R phiO = std::atan2(wo.z, wo.y);
  const R gammaO = std::asin(h);
  const R cosGammaO = std::cos(gammaO);
  Vec<R> ap[4];
  R ap_pdf[4];
  attenuation(sinThetaO, cosThetaO, cosGammaO, h, sig, ap, ap_pdf);
  int p = 0;
  R e1 = u0;
  for (p = 0; p < 3; ++p) {
    if (e1 < ap_pdf[p]) break;
    e1 -= ap_pdf[p];
  }
  R sOp, cOp;
  if (p == 0) {
    sOp = sinThetaO * hs.cos2k[1] - cosThetaO * hs.sin2k[1];
    cOp = cosThetaO * hs.cos2k[1] + sinThetaO * hs.sin2k[1];
  } else if (p == 1) {
    sOp = sinThetaO * hs.cos2k[0] + cosThetaO * hs.sin2k[0];
    cOp = cosThetaO * hs.cos2k[0] - sinThetaO * hs.sin2k[0];
  } else if (p == 2) {
    sOp = sinThetaO * hs.cos2k[2] + cosThetaO * hs.sin2k[2];
    cOp = cosThetaO * hs.cos2k[2] - sinThetaO * hs.sin2k[2];
  } else {
    sOp = sinThetaO;
    cOp = cosThetaO;
  }
  // p = 3 reads si.hair.v[3], one past float v[3]: the next member of
  // HairInteraction is the logistic scale s (common.cuh:38-39), so the
  // residual lobe is sampled with variance s while its pdf uses v[2].
  const R v = p < 3 ? hs.v[p] : hs.s;
  const R e2 = std::fmax(u1, (R)1e-5f);
  const R cosTheta = (R)1 + v * std::log(e2 + ((R)1 - e2) * std::exp((R)-2 / v));
  const R sinTheta = SafeSqrt((R)1 - Sqr(cosTheta));
  const R cosPhi = std::cos((R)2 * Pi<R>() * u2);
  const R sinThetaI = -cosTheta * sOp + sinTheta * cosPhi * cOp;
  const R cosThetaI = SafeSqrt((R)1 - Sqr(sinThetaI));
  const R eta = (R)1.55f;
  const R etap = std::sqrt(eta * eta - Sqr(sinThetaO)) / cosThetaO;
  const R gammaT = SafeASin(h / etap);
  R dphi;
  if (p < 3) {
    const R s = hs.s, a = -Pi<R>(), b = Pi<R>();
    const R k = LogisticCDF(b, s) - LogisticCDF(a, s);
    R x = -s * std::log((R)1 / (u3 * k + LogisticCDF(a, s)) - (R)1);
    x = std::fmin(b, std::fmax(a, x));
    dphi = PhiP(p, gammaO, gammaT) + x;
  } else {
    dphi = (R)2 * Pi<R>() * u3;
  }
  const R phiI = phiO + dphi;
  *lobe = p;
  return nrm(mk(sinThetaI, cosThetaI * std::cos(phiI), cosThetaI * std::sin(phiI)));
}

// ---- GGX head surface: frostbite_anisotropic.cuh:26-157 --------------------
template <class R> inline R ggxD(R a, Vec<R> N) {
  const R t1 = N.x / a, t2 = N.y / a, t3 = N.z;
  return (R)1 / (Pi<R>() * a * a * std::pow(t1 * t1 + t2 * t2 + t3 * t3, (R)2));
}
template <class R> inline R ggxLambda(R a, Vec<R> V) {
  const R t1 = V.x * a, t2 = V.y * a, t3 = V.z;
  return (R)0.5 * ((R)-1 + std::sqrt((R)1 + (t1 * t1 + t2 * t2) / (t3 * t3)));
}
template <class R> inline R ggxG1(R a, Vec<R> V) { return V.z <= 0 ? (R)0 : (R)1 / ((R)1 + ggxLambda(a, V)); }
template <class R> inline R ggxG2(R a, Vec<R> V, Vec<R> L) {
  if (V.z <= 0 || L.z <= 0) return 0;
  return (R)1 / ((R)1 + ggxLambda(a, V) + ggxLambda(a, L));
}
template <class R> inline Vec<R> frostbite(Vec<R> wo, Vec<R> wi, Vec<R> albedo, R a) {
  Vec<R> brdf = mk((R)0, (R)0, (R)0);
  if (wo.z > 0 && wi.z > 0) {
    const Vec<R> H = nrm(wo + wi);
    const R spec = ggxD(a, H) * ggxG2(a, wo, wi) / (R)4 / wo.z / wi.z;
    brdf = ((R)0.5 / Pi<R>()) * albedo + mk((R)0.5 * spec, (R)0.5 * spec, (R)0.5 * spec);
  }
  return std::fabs(wi.z) * brdf;
}
template <class R> inline R ggx_pdf(R a, Vec<R> V, Vec<R> Ne) {
  const R dv = ggxG1(a, V) * std::fmax((R)0, dot(V, Ne)) * ggxD(a, Ne) / V.z;
  return dv / ((R)4 * dot(V, Ne));
}
template <class R> inline Vec<R> ggx_sample(R a, Vec<R> V, R U1, R U2, R* pdf) {
  const Vec<R> Vh = nrm(V * mk(a, a, (R)1));
  const R lensq = Vh.x * Vh.x + Vh.y * Vh.y;
  const Vec<R> T1 = lensq > 0 ? mk(-Vh.y, Vh.x, (R)0) / std::sqrt(lensq) : mk((R)1, (R)0, (R)0);
  const Vec<R> T2 = cross(Vh, T1);
  const R r = std::sqrt(U1), phi = (R)2 * Pi<R>() * U2;
  const R t1 = r * std::cos(phi);
  R t2 = r * std::sin(phi);
  const R s = (R)0.5 * ((R)1 + Vh.z);
  t2 = ((R)1 - s) * std::sqrt((R)1 - t1 * t1) + s * t2;
  const Vec<R> Nh = t1 * T1 + t2 * T2 + std::sqrt(std::fmax((R)0, (R)1 - t1 * t1 - t2 * t2)) * Vh;
  const Vec<R> Ne = nrm(mk(a * Nh.x, a * Nh.y, std::fmax((R)0, Nh.z)));
  const Vec<R> L = ((R)2 * dot(V, Ne)) * Ne - V;
  *pdf = ggx_pdf(a, V, Ne);
  return nrm(L);
}

}  // namespace orc
