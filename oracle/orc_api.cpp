// ORACLE - test infrastructure only (see orc_math.h).
//
// extern "C" surface of the oracle for the Python test harness (ctypes).
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_mlp.h"
#include "orc_scene.h"

using namespace orc;

namespace {

template <class R>
HairParams<R> hp_from(const float* p) {
  return HairParams<R>{(R)p[0], (R)p[1], (R)p[2], (R)p[3], (R)p[4], (R)p[5], (R)p[6]};
}

template <class R>
void eval_batch(const float* prm, int n, const float* q, double* f, double* pdf) {
  const HairParams<R> hp = hp_from<R>(prm);
  const HairShading<R> hs = setup_hair(hp);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float* x = q + 10 * (size_t)i;
    HairQuery<R> hq{mk((R)x[0], (R)x[1], (R)x[2]), mk((R)x[3], (R)x[4], (R)x[5]), (R)x[6],
                    mk((R)x[7], (R)x[8], (R)x[9])};
    R pd;
    const Vec<R> ff = hair_eval(hp, hs, hq, &pd);
    f[3 * (size_t)i] = (double)ff.x;
    f[3 * (size_t)i + 1] = (double)ff.y;
    f[3 * (size_t)i + 2] = (double)ff.z;
    pdf[i] = (double)pd;
  }
}

template <class R>
void sample_batch(const float* prm, int n, const float* q, const float* u4, double* wi, double* f, double* pdf,
                  int* lobe) {
  const HairParams<R> hp = hp_from<R>(prm);
  const HairShading<R> hs = setup_hair(hp);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float* x = q + 10 * (size_t)i;
    const float* u = u4 + 4 * (size_t)i;
    const Vec<R> wo = mk((R)x[0], (R)x[1], (R)x[2]);
    const Vec<R> sig = mk((R)x[7], (R)x[8], (R)x[9]);
    int p;
    const Vec<R> w = hair_sample(hs, wo, (R)x[6], sig, (R)u[0], (R)u[1], (R)u[2], (R)u[3], &p);
    HairQuery<R> hq{wo, w, (R)x[6], sig};
    R pd;
    const Vec<R> ff = hair_eval(hp, hs, hq, &pd);
    wi[3 * (size_t)i] = (double)w.x;
    wi[3 * (size_t)i + 1] = (double)w.y;
    wi[3 * (size_t)i + 2] = (double)w.z;
    f[3 * (size_t)i] = (double)ff.x;
    f[3 * (size_t)i + 1] = (double)ff.y;
    f[3 * (size_t)i + 2] = (double)ff.z;
    pdf[i] = (double)pd;
    lobe[i] = p;
  }
}

}  // namespace

// Mirror of the Python-side scene description.
struct orc_scene {
  const float* seg;
  int nseg;
  int has_sphere;
  float sphere[4];
  float sphere_albedo[3];
  float sphere_alpha;
  int has_env, env_pdf;
  float env_scale, env_rot;
  int env_w, env_h;
  const float* env_rgba;
  const float* c_pdf;
  const float* c_cdf;
  const float* m_pdf;
  const float* m_cdf;
  int ndl;
  float dl_from[8][3], dl_emit[8][3];
  float cam[12];
  int width, height;
  float beta_m, beta_n, alpha, gains[4];
  float sigma_a[3];
  const float* strand_sigma;
  int mis, path_v1, path_v2;
  int frozen;
  float beta_m0, beta_n0, alpha0;
  const float* strand_sigma0;
  const float* tri_v;
  const float* tri_n;
  const float* tri_kd;
  int ntri;
  const float* crv;
};

namespace {
SceneDesc scene_desc(const orc_scene* sc, EnvTables& env) {
  SceneDesc s{};
  s.seg = sc->seg;
  s.nseg = sc->nseg;
  s.crv = sc->crv;
  s.has_sphere = sc->has_sphere;
  std::memcpy(s.sphere, sc->sphere, sizeof(s.sphere));
  std::memcpy(s.sphere_albedo, sc->sphere_albedo, sizeof(s.sphere_albedo));
  s.sphere_alpha = sc->sphere_alpha;
  s.has_env = sc->has_env;
  s.env_pdf = sc->env_pdf;
  s.env_scale = sc->env_scale;
  s.env_rot = sc->env_rot;
  if (sc->has_env) {
    env.W = sc->env_w;
    env.H = sc->env_h;
    const size_t np = (size_t)env.W * env.H;
    env.rgba.assign(sc->env_rgba, sc->env_rgba + np * 4);
    env.cpdf.assign(sc->c_pdf, sc->c_pdf + (size_t)(env.W + 1) * env.H);
    env.ccdf.assign(sc->c_cdf, sc->c_cdf + (size_t)(env.W + 1) * env.H);
    env.mpdf.assign(sc->m_pdf, sc->m_pdf + env.H + 1);
    env.mcdf.assign(sc->m_cdf, sc->m_cdf + env.H + 1);
  }
  s.env = &env;
  s.ndl = sc->ndl;
  std::memcpy(s.dl_from, sc->dl_from, sizeof(s.dl_from));
  std::memcpy(s.dl_emit, sc->dl_emit, sizeof(s.dl_emit));
  std::memcpy(s.cam, sc->cam, sizeof(s.cam));
  s.width = sc->width;
  s.height = sc->height;
  s.beta_m = sc->beta_m;
  s.beta_n = sc->beta_n;
  s.alpha = sc->alpha;
  std::memcpy(s.gains, sc->gains, sizeof(s.gains));
  std::memcpy(s.sigma_a, sc->sigma_a, sizeof(s.sigma_a));
  s.strand_sigma = sc->strand_sigma;
  s.mis = sc->mis;
  s.path_v1 = sc->path_v1;
  s.path_v2 = sc->path_v2;
  s.frozen = sc->frozen;
  s.beta_m0 = sc->beta_m0;
  s.beta_n0 = sc->beta_n0;
  s.alpha0 = sc->alpha0;
  s.strand_sigma0 = sc->strand_sigma0;
  s.tri_v = sc->tri_v;
  s.tri_n = sc->tri_n;
  s.tri_kd = sc->tri_kd;
  s.ntri = sc->ntri;
  return s;
}
}  // namespace

extern "C" {

uint32_t orc_path_seed(uint32_t frame, uint32_t x, uint32_t y, uint32_t w) { return get_rng(frame, x, y, w); }

void orc_lcg_stream(uint32_t seed, int n, uint32_t* out_u32, float* out_f32) {
  Lcg r{seed};
  for (int i = 0; i < n; ++i) {
    const uint32_t s = r.next_u32();
    if (out_u32) out_u32[i] = s;
    if (out_f32) out_f32[i] = std::ldexp((float)s, -32);
  }
}

// params: beta_m, beta_n, alpha (rad), gain R, TT, TRT, TRRT
// query: wo(3), wi(3), h, sigma_a(3) as 10 floats. prec: 0 = FP64, 1 = FP32.
void orc_hair_eval(const float* params, int n, const float* q, int prec, double* f, double* pdf) {
  if (prec == 0) eval_batch<double>(params, n, q, f, pdf);
  else eval_batch<float>(params, n, q, f, pdf);
}

void orc_hair_sample(const float* params, int n, const float* q, const float* u4, int prec, double* wi, double* f,
                     double* pdf, int* lobe) {
  if (prec == 0) sample_batch<double>(params, n, q, u4, wi, f, pdf, lobe);
  else sample_batch<float>(params, n, q, u4, wi, f, pdf, lobe);
}

// v[3], s, sin2k[3], cos2k[3] in FP64.
void orc_hair_setup(const float* params, double* out) {
  const HairShading<double> hs = setup_hair(hp_from<double>(params));
  for (int i = 0; i < 3; ++i) out[i] = hs.v[i];
  out[3] = hs.s;
  for (int i = 0; i < 3; ++i) out[4 + i] = hs.sin2k[i];
  for (int i = 0; i < 3; ++i) out[7 + i] = hs.cos2k[i];
}

double orc_bessel_i0(double x) { return I0<double>(x); }

double orc_mp(double ci, double co, double si, double so, double v) { return Mp<double>(ci, co, si, so, v); }

// mode 0: FP64 running sums; 1: sequential FP32 (reference order).
void orc_env_tables(int W, int H, const float* rgba, int mode, float* c_pdf, float* c_cdf, float* m_pdf,
                    float* m_cdf) {
  EnvTables e;
  e.W = W;
  e.H = H;
  e.rgba.assign(rgba, rgba + (size_t)W * H * 4);
  env_tables(e, mode);
  std::memcpy(c_pdf, e.cpdf.data(), sizeof(float) * e.cpdf.size());
  std::memcpy(c_cdf, e.ccdf.data(), sizeof(float) * e.ccdf.size());
  std::memcpy(m_pdf, e.mpdf.data(), sizeof(float) * e.mpdf.size());
  std::memcpy(m_cdf, e.mcdf.data(), sizeof(float) * e.mcdf.size());
}

// Brute force (use_bvh = 0) or CPU BVH (1). Rays: 8 floats (o, pad, d, pad).
void orc_intersect(int nseg, const float* seg, int nray, const float* rays, int any_hit, int use_bvh, int* prim,
                   float* t, float* u) {
  Bvh bvh;
  if (use_bvh) bvh.build(seg, nseg);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < nray; ++i) {
    const float* r = rays + 8 * (size_t)i;
    const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    if (any_hit) {
      bool occ = false;
      if (use_bvh) {
        occ = bvh.any(o, d);
      } else {
        for (int s = 0; s < nseg && !occ; ++s) {
          float tt, uu;
          occ = seg_intersect(o, d, seg + 8 * (size_t)s, &tt, &uu);
        }
      }
      prim[i] = occ ? 1 : 0;
      continue;
    }
    float bt = 0.f, bu = 0.f;
    int bid = -1;
    if (use_bvh) {
      if (!bvh.closest(o, d, &bt, &bid, &bu)) bid = -1;
    } else {
      for (int s = 0; s < nseg; ++s) {
        float tt, uu;
        if (seg_intersect(o, d, seg + 8 * (size_t)s, &tt, &uu)) {
          if (bid < 0 || tt < bt) {  // ascending ids: ties keep the lower id
            bt = tt;
            bu = uu;
            bid = s;
          }
        }
      }
    }
    prim[i] = bid;
    if (t) t[i] = bid >= 0 ? bt : 0.f;
    if (u) u[i] = bid >= 0 ? bu : 0.f;
  }
}

void* orc_bvh_create(const float* seg, int n) {
  Bvh* b = new Bvh();
  b->build(seg, n);
  return b;
}

void orc_bvh_destroy(void* b) { delete (Bvh*)b; }

// A BVH over round Catmull-Rom segments (kCrvFloats each).
void* orc_bvh_create_curves(const float* crv, int n) {
  Bvh* b = new Bvh();
  b->build_curves(crv, n);
  return b;
}

int orc_curve_floats() { return kCrvFloats; }

void orc_curve_eval(const float* q, int n, const double* u, int prec, double* out) {
  for (int i = 0; i < n; ++i) crv_eval(q, u[i], prec, out + 8 * (size_t)i);
}

// Closest (or any) hit of each ray over round curves: brute force
// (use_bvh = 0, minimum (t, id)) or the CPU BVH; caps: end caps tested.
// Rays: 8 floats (o, pad, d, pad).
void orc_curve_intersect(int ncrv, const float* crv, int nray, const float* rays, int any_hit, int use_bvh, int caps,
                         int* prim, float* t, float* u) {
  Bvh bvh;
  if (use_bvh) bvh.build_curves(crv, ncrv);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < nray; ++i) {
    const float* r = rays + 8 * (size_t)i;
    const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    float bt = 0.f, bu = 0.f;
    int bid = -1;
    if (use_bvh) {
      if (any_hit) {
        prim[i] = bvh.any(o, d, caps) ? 1 : 0;
        continue;
      }
      if (!bvh.closest(o, d, &bt, &bid, &bu, caps)) bid = -1;
    } else {
      for (int k = 0; k < ncrv; ++k) {
        const float* q = crv + (size_t)kCrvFloats * k;
        float tt, uu;
        if (crv_intersect(o, d, q, caps ? (int)q[17] : 0, &tt, &uu) && (bid < 0 || tt < bt)) {
          bt = tt;
          bu = uu;
          bid = k;
        }
      }
      if (any_hit) {
        prim[i] = bid >= 0 ? 1 : 0;
        continue;
      }
    }
    prim[i] = bid;
    if (t) t[i] = bid >= 0 ? bt : 0.f;
    if (u) u[i] = bid >= 0 ? bu : 0.f;
  }
}

// One ray against one curve segment each: the FP32 iteration (t32, u32,
// hit32) and the FP64 search (t64, u64, hit64, min_clear). flags: caps.
void orc_curve_pairs(int n, const float* crv, const double* rays, int flags, float* t32, float* u32, uint8_t* hit32,
                     double* t64, double* u64, uint8_t* hit64, double* min_clear) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    const float* q = crv + (size_t)kCrvFloats * i;
    const double* r = rays + 6 * (size_t)i;
    const float of[3] = {(float)r[0], (float)r[1], (float)r[2]}, df[3] = {(float)r[3], (float)r[4], (float)r[5]};
    float tf = 0.f, uf = 0.f;
    hit32[i] = crv_intersect(of, df, q, flags, &tf, &uf) ? 1 : 0;
    t32[i] = tf;
    u32[i] = uf;
    // the FP64 search on the FP32-rounded ray, so both see the same ray
    const double od[3] = {of[0], of[1], of[2]}, dd[3] = {df[0], df[1], df[2]};
    double td = 0, ud = 0, mc = 0;
    hit64[i] = crv_exact(od, dd, q, flags, &td, &ud, &mc) ? 1 : 0;
    t64[i] = td;
    u64[i] = ud;
    min_clear[i] = mc;
  }
}

// Renders a crop of the image. mode 0: pathTrace, 1: msnnPathTrace with
// kernel beta kbeta, 2: msnnTrainingPath. prec 0 FP64, 1 FP32. Outputs the
// per-pixel mean over spp passes starting at accum_id0 (RGB doubles).
int orc_render(const orc_scene* sc, void* bvh_handle, int mode, int kbeta, int accum_id0, int spp, int x0, int y0,
               int w, int h, int threads, int prec, double* out_full, double* out_short, uint8_t* out_hit) {
  if (!sc || !bvh_handle || spp <= 0) return 1;
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  const Bvh& bvh = *(const Bvh*)bvh_handle;
  if (prec == 0)
    render_pixels<double>(s, bvh, mode, kbeta, accum_id0, spp, x0, y0, w, h, threads, out_full, out_short, out_hit);
  else
    render_pixels<float>(s, bvh, mode, kbeta, accum_id0, spp, x0, y0, w, h, threads, out_full, out_short, out_hit);
  return 0;
}

// NRC G_BUFFER pass of frame accum_id over the whole image (orc_scene.h
// nrc_gbuffer); out holds width * height NrcPixelOut records.
int orc_nrc_gbuffer(const orc_scene* sc, void* bvh_handle, int accum_id, const uint8_t* flags, float c,
                    float scene_scale, int threads, int prec, NrcPixelOut* out) {
  if (!sc || !bvh_handle || !flags || !out) return 1;
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  const Bvh& bvh = *(const Bvh*)bvh_handle;
  if (prec == 0) nrc_gbuffer<double>(s, bvh, accum_id, flags, c, scene_scale, threads, out);
  else nrc_gbuffer<float>(s, bvh, accum_id, flags, c, scene_scale, threads, out);
  return 0;
}

int orc_nrc_record_size(void) { return (int)sizeof(NrcPixelOut); }

// HairMSNN training records of frame accum_id at n pixels (orc_scene.h
// msnn_records); out holds n MsnnRecordOut records.
int orc_msnn_records(const orc_scene* sc, void* bvh_handle, int accum_id, int kbeta, int n, const int* pixels,
                     float scene_scale, int threads, int prec, MsnnRecordOut* out) {
  if (!sc || !bvh_handle || n < 0 || !pixels || !out) return 1;
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  const Bvh& bvh = *(const Bvh*)bvh_handle;
  if (prec == 0) msnn_records<double>(s, bvh, accum_id, kbeta, n, pixels, scene_scale, threads, out);
  else msnn_records<float>(s, bvh, accum_id, kbeta, n, pixels, scene_scale, threads, out);
  return 0;
}

int orc_msnn_record_size(void) { return (int)sizeof(MsnnRecordOut); }

// Ray k against triangle k (9 floats each; rays as 8 floats o, pad, d, pad):
// hit flag and (t, u, v) per pair.
void orc_tri_intersect(int n, const float* tris, const float* rays, int* hit, float* t, float* u, float* v) {
  for (int k = 0; k < n; ++k) {
    const float* r = rays + 8 * (size_t)k;
    const float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
    float tt = 0.f, uu = 0.f, vv = 0.f;
    hit[k] = tri_intersect(o, d, tris + 9 * (size_t)k, &tt, &uu, &vv) ? 1 : 0;
    t[k] = tt;
    u[k] = uu;
    v[k] = vv;
  }
}

// FP64 path-replay adjoint of the full image (orc_scene.h adjoint_render).
int orc_adjoint(const orc_scene* sc, void* bvh_handle, int spp, const float* dl_dimg, int num_strands, int threads,
                double fd_h, double* grad_sigma, double* grad_glob, double* grad_px) {
  if (!sc || !bvh_handle || spp <= 0 || !dl_dimg || num_strands <= 0) return 1;
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  adjoint_render(s, *(const Bvh*)bvh_handle, spp, dl_dimg, num_strands, threads, fd_h, grad_sigma, grad_glob, grad_px);
  return 0;
}

// Radiance-cache network (orc_mlp.h). prec 0 = FP64, 1 = FP32.
void orc_mlp_encode(const float* in, int n, double* feat) {
  for (int s = 0; s < n; ++s) mlp_encode<double>(in + (size_t)s * kMlpIn, feat + (size_t)s * kMlpFeat);
}

void orc_mlp_forward(const float* w, const float* in, int n, int prec, double* out) {
  if (prec == 0) {
    mlp_forward<double>(w, in, n, out);
  } else {
    std::vector<float> o((size_t)n * kMlpOut);
    mlp_forward<float>(w, in, n, o.data());
    for (size_t i = 0; i < o.size(); ++i) out[i] = o[i];
  }
}

double orc_mlp_loss_grad(const float* w, const float* in, const float* target, int n, int loss, int prec,
                         double exp_floor, double* grad) {
  if (prec == 0) return mlp_loss_grad<double>(w, in, target, n, loss, exp_floor, grad);
  std::vector<float> g(kMlpWeights);
  const double l = mlp_loss_grad<float>(w, in, target, n, loss, exp_floor, g.data());
  for (int i = 0; i < kMlpWeights; ++i) grad[i] = g[i];
  return l;
}

// ---- probes of single functions (orc_scene.h probe_*; GGX from orc_math.h)
// prec 0 = FP64, 1 = FP32 throughout.

void orc_probe_env(const orc_scene* sc, int n, const float* dirs, int prec, double* rgb, double* pdf) {
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  if (prec == 0) probe_env<double>(s, n, dirs, rgb, pdf);
  else probe_env<float>(s, n, dirs, rgb, pdf);
}

void orc_probe_env_sample(const orc_scene* sc, int n, const uint32_t* seeds, int prec, double* wi, double* rgb,
                          double* pdf) {
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  if (prec == 0) probe_env_sample<double>(s, n, seeds, wi, rgb, pdf);
  else probe_env_sample<float>(s, n, seeds, wi, rgb, pdf);
}

int orc_probe_hit_size(void) { return kProbeHit; }

void orc_probe_hit(const orc_scene* sc, void* bvh_handle, int n, const float* rays, const float* t, const int* prim,
                   const float* u, int prec, double* out) {
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  const Bvh& bvh = *(const Bvh*)bvh_handle;
  if (prec == 0) probe_hit<double>(s, bvh, n, rays, t, prim, u, out);
  else probe_hit<float>(s, bvh, n, rays, t, prim, u, out);
}

void orc_probe_direct(const orc_scene* sc, void* bvh_handle, const float* ray, float t, int prim, float u, int n,
                      const uint32_t* seeds, int threads, int prec, double* mean_rgb) {
  EnvTables env;
  const SceneDesc s = scene_desc(sc, env);
  const Bvh& bvh = *(const Bvh*)bvh_handle;
  if (prec == 0) probe_direct<double>(s, bvh, ray, t, prim, u, n, seeds, threads, mean_rgb);
  else probe_direct<float>(s, bvh, ray, t, prim, u, n, seeds, threads, mean_rgb);
}

// GGX microfacet terms in the local frame (z = normal), alpha as used by the
// BSDF (already clamped and squared): D(h) per half vector; frostbite_GGX
// (f cos, RGB) and pdf_GGX of wi at h = normalize(wo + wi) per (wo, wi);
// sample_GGX per (wo, u1, u2): wi, pdf and frostbite_GGX at wi.
void orc_ggx_d(int n, double a, const double* h, double* out) {
  for (int i = 0; i < n; ++i) out[i] = ggxD<double>(a, mk(h[3 * i], h[3 * i + 1], h[3 * i + 2]));
}

void orc_ggx_eval(int n, double a, const double* wo, const double* wi, const double* albedo, double* f,
                  double* pdf) {
  const Vec<double> alb = mk(albedo[0], albedo[1], albedo[2]);
  for (int i = 0; i < n; ++i) {
    const Vec<double> o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    const Vec<double> w = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    const Vec<double> ff = frostbite<double>(o, w, alb, a);
    f[3 * i] = ff.x;
    f[3 * i + 1] = ff.y;
    f[3 * i + 2] = ff.z;
    pdf[i] = ggx_pdf<double>(a, o, nrm(o + w));
  }
}

void orc_ggx_sample(int n, double a, const double* wo, const double* u2, const double* albedo, double* wi,
                    double* pdf, double* f) {
  const Vec<double> alb = mk(albedo[0], albedo[1], albedo[2]);
  for (int i = 0; i < n; ++i) {
    const Vec<double> o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    double p = 0;
    const Vec<double> w = ggx_sample<double>(a, o, u2[2 * i], u2[2 * i + 1], &p);
    const Vec<double> ff = frostbite<double>(o, w, alb, a);
    wi[3 * i] = w.x;
    wi[3 * i + 1] = w.y;
    wi[3 * i + 2] = w.z;
    pdf[i] = p;
    f[3 * i] = ff.x;
    f[3 * i + 1] = ff.y;
    f[3 * i + 2] = ff.z;
  }
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
