// Wavefront path tracer: scene view, per-path state layout and the shading
// routines shared by the plain path tracer and the HairMSNN passes.
//
// Integrator semantics follow cuda_headers/optix_common.cuh:151-409
// (sampleBSDF, sampleLights, directLighting, nextPathVertex, pathTrace) and
// cuda/hair_msnn.cu:17-185 (truncated and training paths), with the
// reference's random-number consumption order per vertex:
//   light pick (1) [+ env sample (2)] [+ MIS BSDF sample (2, +2 on hair)]
//   [Russian roulette (1), from the second vertex on]
//   continuation sample (2, +2 on hair)
// Shadow rays are not traced inside the shading kernel: their unoccluded
// contributions are stored with the ray and resolved by the shadow kernel,
// which applies the reference's NaN scrub to the sum of the visible parts.
#pragma once

#include "hm_bvh.h"
#include "hm_env.h"
#include "hm_ggx.h"
#include "hm_hair_bsdf.h"
#include "hm_rng.h"

namespace hm {

constexpr int kMaxDirLights = 8;
constexpr int kPrimMiss = -1;
constexpr int kPrimSphere = -2;
constexpr int kPrimMesh0 = -3;  // triangle k of the head mesh is prim kPrimMesh0 - k

struct SceneView {
  BvhView bvh;
  const float* seg_by_id;     // 8 floats per segment, original order
  // round Catmull-Rom hair (null for linear segments): 5 float4 per segment,
  // the four control points (x, y, z, radius), then (strand, end flags);
  // the BVH leaves hold their boxes. curve_caps: flat caps at strand ends
  // (on for plain path tracing, off for HairMSNN and NRC, as the reference's
  // forceCaps flags)
  const float4* curves;
  int curve_caps;
  const float* strand_sigma;  // per-strand sigma_a (3 floats) or null
  SphereView sphere;
  MeshView mesh;
  V3 sphere_albedo;
  float sphere_alpha;  // clamped to [0.01, 1] and squared (optix_common.cuh:551-552)
  EnvView env;
  int num_dlights, num_total_lights;
  // per directional light: (direction, 0), (emit, 0); the direction is the
  // parsed (normalised) `from` normalised once more on the device, as
  // sampleLights does at every use (optix_common.cuh:236)
  const float4* dlights;
  V3 cam_pos, cam_d00, cam_du, cam_dv;
  int width, height;
  HairConsts hc;
  V3 sigma_a;
  int mis;
};

constexpr int kDoneCount = 4;  // PathBuffers::counts index of the finished-path queue
// PathBuffers::counts holds each queue length in a 256-byte line of its own:
// counts[kCountStride * k] is the length of queue k (0 and 1 the two
// wavefronts, 2 the shadow records, 3 the shadow rays, kDoneCount the
// finished paths), so the appends to different queues do not contend for
// one line.
constexpr int kCountStride = 64;
constexpr int kCountInts = kCountStride * (kDoneCount + 1);

// Per-slot path state (structure of arrays in HBM). A slot is one pixel of
// the current sample pass.
// Second copy of the per-path state for the streaming schedule, which
// compacts its wavefront into the other copy every iteration.
struct PathState {
  float4* ray_o;
  float4* ray_d;
  float4* beta;
  float4* L;
  uint32_t* rng;
  int* pass;
  int* px;
  int* depth;
  float4* beta_short;  // HairMSNN training paths only
  float4* L_short;
};

// Scheduler words of the job schedule (pt_jobs_*): two copies of [0] next
// job, [2] free path capacity, each copy in its own 64-byte line, then at
// kJobCtlLive the number of iterations that found live paths.
constexpr int kJobCtlHalf = 8;
constexpr int kJobCtlLive = 2 * kJobCtlHalf;
constexpr int kJobCtl = kJobCtlLive + kJobCtlHalf;

struct PathBuffers {
  float4* ray_o;
  float4* ray_d;
  float4* beta;
  float4* beta_short;  // HairMSNN training paths only
  uint32_t* rng;
  float4* L;
  float4* L_short;     // HairMSNN training paths only
  uint32_t* flags;     // bit 0: training pixel
  float* hit_t;
  int* hit_prim;
  float* hit_u;
  // pending direct-lighting records
  float4* sh_o1;  // xyz origin, w = 1 if the ray exists
  float4* sh_d1;
  float4* sh_c1;  // unoccluded contribution (before path throughput)
  float4* sh_o2;
  float4* sh_d2;
  float4* sh_c2;
  float4* sh_beta;        // path throughput at the vertex
  float4* sh_beta_short;  // HairMSNN training paths only
  int* queue[2];
  int* shadow_queue;  // slots with pending direct light
  int* shadow_rays;   // pending shadow rays: slot * 2 + (0: light sample, 1: BSDF sample)
  uint8_t* sh_vis;    // visibility per (slot, ray), same indexing as shadow_rays
  // streaming schedule: per path the job's pass (relative to the call's
  // first) and pixel and the vertex depth
  int* slot_pass;
  int* slot_px;
  int* slot_depth;
  // compacted streaming: the other state copy, where each path of this
  // wavefront went (>= 0: position in the next wavefront, < 0: ~position in
  // the finished list) and the finished paths' radiance, pixel and pass
  PathState alt;
  int* dest;
  float4* done_L;
  float4* done_Ls;  // HairMSNN training paths: the finished paths' short radiance
  int* done_px;
  int* done_pass;
  long long* job_ctl;  // kJobCtl words
  int* counts;  // [0], [1] continuation queues, [2] shadow slots, [3] shadow rays, [4] finished paths
  int* spill;   // traversal stack spill
  int spill_stride;
  int spill_levels;
};

struct ShadeParams {
  int depth;       // vertex index of the hits being shaded (0 = camera hit)
  int dl_zero;     // direct lighting at the camera vertex
  int dl_from;     // direct lighting at vertices >= dl_from (for depth >= 1)
  int max_vertex;  // continuation allowed while depth + 1 <= max_vertex
  int trunc_beta;  // HairMSNN: terminate after the vertex when depth > trunc_beta
  int msnn;        // 1: HairMSNN G-buffer pass (training paths enabled)
  int cur;         // current continuation queue index
};

#if defined(__HIPCC__)

// Shading point after a closest hit.
struct Hit {
  int surface;
  V3 p, n, wo, wo_local;
  Frame fr;
  float radius;
  HairWo hw;
  V3 albedo;
  float alpha;
};

HM_D float draw(uint32_t& s) { return lcg_uniform(s); }

// Closest-hit shading setup (optix_common.cuh:527-599).
// Every field is written on both branches so the optimiser can keep the
// whole record in registers.
HM_D void setup_hit(const SceneView& sc, V3 o, V3 d, float t, int prim, float u, Hit& h) {
  h.wo = -1.f * d;
  const V3 p = o + t * d;
  h.albedo = sc.sphere_albedo;
  h.alpha = sc.sphere_alpha;
  if (prim == kPrimSphere) {
    const V3 c = v3(sc.sphere.cx, sc.sphere.cy, sc.sphere.cz);
    h.n = normalize(p - c);
    h.p = c + sc.sphere.r * h.n;
    h.fr = basis_from_normal(h.n);
    h.wo_local = normalize(to_local(h.fr, h.wo));
    h.surface = 1;
    h.radius = 0.f;
    h.hw = HairWo{};
  } else if (prim <= kPrimMesh0) {
    // triangleMeshCH (optix_common.cuh:527-556): barycentric position and
    // normal, frame from the normal, the material's diffuse colour, alpha 1
    // (mesh->alpha = 1, model.cpp:323) clamped and squared
    const int k = kPrimMesh0 - prim;
    const float4 a = sc.mesh.tri[3 * k], b = sc.mesh.tri[3 * k + 1], c = sc.mesh.tri[3 * k + 2];
    float tt = 0.f, bu = 0.f, bv = 0.f;
    (void)intersect_triangle(o.x, o.y, o.z, d.x, d.y, d.z, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, &tt, &bu,
                             &bv);
    const float bw = 1.f - bu - bv;
    h.p = v3(bw * a.x + bu * b.x + bv * c.x, bw * a.y + bu * b.y + bv * c.y, bw * a.z + bu * b.z + bv * c.z);
    const float4 n0 = sc.mesh.shade[4 * k], n1 = sc.mesh.shade[4 * k + 1], n2 = sc.mesh.shade[4 * k + 2];
    const float4 kd = sc.mesh.shade[4 * k + 3];
    h.n = normalize(v3(bw * n0.x + bu * n1.x + bv * n2.x, bw * n0.y + bu * n1.y + bv * n2.y,
                       bw * n0.z + bu * n1.z + bv * n2.z));
    h.fr = basis_from_normal(h.n);
    h.wo_local = normalize(to_local(h.fr, h.wo));
    h.surface = 1;
    h.radius = 0.f;
    h.hw = HairWo{};
    h.albedo = v3(kd.x, kd.y, kd.z);
    h.alpha = 1.f;
  } else {
    const float* s = sc.seg_by_id + 8 * (size_t)prim;
    V3 tangent;
    if (sc.curves) {
      // hairCH on a round Catmull-Rom curve (computeCurveIntersection,
      // curve_utils.cuh:218-241)
      const float4* q = sc.curves + 5 * (size_t)prim;
      const CurveHitGeom g = curve_hit_geometry(make_curve(q[0], q[1], q[2], q[3]), u, p);
      h.p = g.p;
      h.n = g.n;
      h.radius = g.radius;
      tangent = g.tangent;
    } else {
      const V3 p0 = v3(s[0], s[1], s[2]), p1 = v3(s[4], s[5], s[6]);
      const SegHitGeom g = segment_hit_geometry(p0, p1, s[3], u, p);
      h.p = g.p;
      h.n = g.n;
      h.radius = s[3];
      tangent = g.tangent;
    }
    h.fr = hair_frame(tangent, h.wo);
    h.wo_local = normalize(to_local(h.fr, h.wo));
    h.surface = 0;
    V3 sig = sc.sigma_a;
    if (sc.strand_sigma) {
      const int strand = (int)s[7];
      sig = v3(sc.strand_sigma[3 * strand], sc.strand_sigma[3 * strand + 1], sc.strand_sigma[3 * strand + 2]);
    }
    const float hh = dot(h.fr.r1, h.n);
    h.hw = hair_prepare(sc.hc, h.wo_local, hh, sig);
  }
}

HM_D FPdf eval_bsdf(const SceneView& sc, const Hit& h, V3 wi_local) {
  if (h.surface) {
    return FPdf{surface_eval(h.wo_local, wi_local, h.albedo, h.alpha),
                surface_pdf(h.alpha, h.wo_local, normalize(h.wo_local + wi_local))};
  }
  return hair_eval_fp(sc.hc, h.hw, wi_local);
}

// A BSDF sample: local and world direction, f * |cos|, pdf.
struct BsdfSample {
  V3 wl, wi, f;
  float pdf;
};

// BSDF sample with the reference's draw order: two uniforms, then two more
// on hair (optix_common.cuh:154-170, :314-329).
HM_D BsdfSample sample_bsdf(const SceneView& sc, const Hit& h, uint32_t& rng) {
  const float r0 = draw(rng), r1 = draw(rng);
  BsdfSample s;
  if (!h.surface) {
    const float r2 = draw(rng), r3 = draw(rng);
    int lobe;
    s.wl = hair_sample_dir(sc.hc, h.hw, r0, r1, r2, r3, &lobe);
    s.wi = normalize(to_world(h.fr, s.wl));
    const FPdf e = hair_eval_fp(sc.hc, h.hw, s.wl);
    s.f = e.f;
    s.pdf = e.pdf;
  } else {
    const DirPdf d = surface_sample_dp(h.alpha, h.wo_local, r0, r1);
    s.wl = d.dir;
    s.pdf = d.pdf;
    s.wi = normalize(to_world(h.fr, s.wl));
    s.f = surface_eval(h.wo_local, s.wl, h.albedo, h.alpha);
  }
  return s;
}

struct DirectOut {
  int has1, has2;
  V3 o1, d1, c1, o2, d2, c2;
  // for the adjoint: the BSDF values (f * |cos|) and local directions behind
  // c1 and c2, and the pdfs of their MIS weights (mis1: c1 carries the
  // weight lp1^2 / (lp1^2 + bp1^2); c2 always carries bp2^2 / (bp2^2 + lp2^2))
  V3 f1, f2, wl1, wl2;
  int mis1;
  float lp1, bp1, lp2, bp2;
};

// The light sample of sampleLights (optix_common.cuh:205-246): the light
// pick and, for the environment, its importance sample. It uses only the
// random stream, not the hit, so a shading kernel can take it before the hit
// setup and overlap the table lookups with the hit's loads.
struct LightSample {
  V3 emit, wi;
  float light_pdf;
  bool is_delta;
};

HM_D LightSample sample_light(const SceneView& sc, uint32_t& rng) {
  LightSample ls;
  const float inv_n = 1.f / (float)sc.num_total_lights;
  const float rsel = draw(rng);
  int sel = (int)floorf(rsel * (float)sc.num_total_lights);
  // A draw of exactly 1.0 selects index N; without an environment light that
  // index does not exist (undefined in the reference), so it maps to the last
  // directional light.
  if (!sc.env.enabled && sel >= sc.num_dlights) sel = sc.num_dlights - 1;
  ls.is_delta = false;
  ls.light_pdf = 1.f * inv_n;
  if (sel >= sc.num_dlights) {
    const float ux = draw(rng), uy = draw(rng);
    float env_p;
    ls.emit = env_sample(sc.env, ux, uy, &ls.wi, &env_p);
    ls.light_pdf = ls.light_pdf * env_p;
  } else {
    const float4 f = sc.dlights[2 * sel], e = sc.dlights[2 * sel + 1];
    ls.wi = v3(f.x, f.y, f.z);
    ls.emit = v3(e.x, e.y, e.z);
    ls.is_delta = true;
  }
  return ls;
}

HM_D void direct_lighting(const SceneView& sc, const Hit& h, const LightSample& ls, uint32_t& rng, DirectOut& out);

// directLighting with MIS (optix_common.cuh:205-309) up to the shadow tests.
HM_D void direct_lighting(const SceneView& sc, const Hit& h, uint32_t& rng, DirectOut& out) {
  const LightSample ls = sample_light(sc, rng);
  direct_lighting(sc, h, ls, rng, out);
}

// directLighting with MIS (optix_common.cuh:279-309) up to the shadow tests,
// given the vertex's light sample (drawn first from the same stream).
HM_D void direct_lighting(const SceneView& sc, const Hit& h, const LightSample& ls, uint32_t& rng, DirectOut& out) {
  out.has1 = out.has2 = 0;
  out.c1 = out.c2 = v3(0.f);
  V3 p = h.p;  // the push-through below is shared by both samples (reference quirk)
  const float inv_n = 1.f / (float)sc.num_total_lights;
  const V3 emit = ls.emit, wi = ls.wi;
  const float light_pdf = ls.light_pdf;
  const bool is_delta = ls.is_delta;
  {
    const V3 wi_local = normalize(to_local(h.fr, wi));
    const bool lower = dot(wi, h.n) < 0.f;
    V3 nd = h.n;
    bool skip = false;
    if (h.surface && lower) {
      skip = true;
    } else if (!h.surface && lower) {
      nd = -1.f * h.n;
      p = p + (2.f * h.radius) * nd;
    }
    if (!skip) {
      const FPdf ev = eval_bsdf(sc, h, wi_local);
      const V3 f = ev.f;
      const float bpdf = ev.pdf;
      V3 c;
      if (sc.mis && !is_delta) c = f * emit * power_heuristic(light_pdf, bpdf) / light_pdf;
      else c = f * emit / light_pdf;
      out.has1 = 1;
      out.o1 = p + 1e-3f * nd;
      out.d1 = wi;
      out.c1 = c;
      out.f1 = f;
      out.wl1 = wi_local;
      out.mis1 = sc.mis && !is_delta;
      out.lp1 = light_pdf;
      out.bp1 = bpdf;
    }
  }
  if (!sc.mis || is_delta) return;
  // ---- BSDF sample toward the environment (sampleBSDF) ----
  const BsdfSample bs = sample_bsdf(sc, h, rng);
  const V3 wl = bs.wl, wb = bs.wi, f = bs.f;
  const float bpdf = bs.pdf;
  const bool lower = dot(wb, h.n) < 0.f;
  V3 nd = h.n;
  if (h.surface && lower) return;
  if (!h.surface && lower) {
    nd = -1.f * h.n;
    p = p + (2.f * h.radius) * nd;
  }
  const V3 e = env_radiance(sc.env, wb);
  const float lp = env_pdf(sc.env, wb) * 1.f * inv_n;
  out.has2 = 1;
  out.o2 = p + 1e-3f * nd;
  out.d2 = wb;
  out.c2 = f * e * power_heuristic(bpdf, lp) / bpdf;
  out.f2 = f;
  out.wl2 = wl;
  out.lp2 = lp;
  out.bp2 = bpdf;
}

// The lighting and continuation of one path vertex: directLighting with its
// light sample and MIS BSDF sample (optix_common.cuh:151-309), then Russian
// roulette and the continuation sample (nextPathVertex, :311-360), with the
// reference's draw order. The three BSDF evaluations (light direction, MIS
// sample, continuation) and the two BSDF samples run through one copy of the
// sampling and evaluation code, a loop over k = 0 (light sample), 1 (MIS BSDF
// sample), 2 (continuation): inlined three times over, that code made the
// shading kernels larger than the instruction cache, and the kernels spent
// most of their time refetching instructions.
//
// dl: light this vertex. roulette(eps): when rr, called with the roulette draw
// before the continuation; returns whether the path goes on (and rescales the
// caller's throughput). may_continue: a continuation is allowed.
// light(which, o, d, c): an unoccluded shadow ray (which 0: light sample, 1: MIS
// BSDF sample) with its contribution before the path throughput; called at
// most once per which, in order. The continuation is returned: out.cont, the
// new ray origin and direction, the throughput factor f / pdf (NaN -> 1) and
// the sampling pdf.
struct Continuation {
  int cont;
  V3 o, wi, mul;
  float pdf;
};

template <class Roulette, class LightSink>
HM_D Continuation shade_vertex(const SceneView& sc, const Hit& h, uint32_t& rng, bool dl, bool rr, bool may_continue,
                               Roulette roulette, LightSink light) {
  Continuation out;
  out.cont = 0;
  out.o = out.wi = out.mul = v3(0.f);
  out.pdf = 0.f;
  LightSample ls;
  ls.emit = ls.wi = v3(0.f);
  ls.light_pdf = 1.f;
  ls.is_delta = true;
  if (dl) ls = sample_light(sc, rng);
  const float inv_n = 1.f / (float)sc.num_total_lights;
  V3 pdl = h.p;  // the push-through is shared by both direct-lighting samples (reference quirk)
#pragma unroll 1
  for (int k = 0; k < 3; ++k) {
    if (k == 0 && !dl) continue;
    if (k == 1 && (!dl || !sc.mis || ls.is_delta)) continue;
    if (k == 2) {
      const bool alive = rr ? roulette(draw(rng)) : true;
      if (!alive || !may_continue) break;
    }
    V3 wl, wi;
    float spdf = 0.f;
    if (k == 0) {
      wi = ls.wi;
      wl = normalize(to_local(h.fr, wi));
    } else {
      // sampleBSDF / nextPathVertex draws: two uniforms, two more on hair
      // (optix_common.cuh:154-170, :314-329)
      const float r0 = draw(rng), r1 = draw(rng);
      if (!h.surface) {
        const float r2 = draw(rng), r3 = draw(rng);
        int lobe;
        wl = hair_sample_dir(sc.hc, h.hw, r0, r1, r2, r3, &lobe);
      } else {
        const DirPdf d = surface_sample_dp(h.alpha, h.wo_local, r0, r1);
        wl = d.dir;
        spdf = d.pdf;
      }
      wi = normalize(to_world(h.fr, wl));
    }
    const bool lower = dot(wi, h.n) < 0.f;
    if (k < 2 && h.surface && lower) continue;  // no light from below a surface
    V3 f;
    float pdf;
    if (!h.surface) {
      const FPdf e = hair_eval_fp(sc.hc, h.hw, wl);
      f = e.f;
      pdf = e.pdf;
    } else {
      f = surface_eval(h.wo_local, wl, h.albedo, h.alpha);
      pdf = k == 0 ? surface_pdf(h.alpha, h.wo_local, normalize(h.wo_local + wl)) : spdf;
    }
    if (k < 2) {
      V3 nd = h.n;
      if (!h.surface && lower) {
        nd = -1.f * h.n;
        pdl = pdl + (2.f * h.radius) * nd;
      }
      V3 c;
      if (k == 0) {
        const float lp = ls.light_pdf;
        if (sc.mis && !ls.is_delta) c = f * ls.emit * power_heuristic(lp, pdf) / lp;
        else c = f * ls.emit / lp;
      } else {
        const V3 e = env_radiance(sc.env, wi);
        const float lp = env_pdf(sc.env, wi) * 1.f * inv_n;
        c = f * e * power_heuristic(pdf, lp) / pdf;
      }
      light(k, pdl + 1e-3f * nd, wi, c);
    } else {
      V3 mul = (pdf == 0.f) ? f : f / pdf;
      if (any_nan(mul)) mul = v3(1.f);
      V3 p = h.p, nd = h.n;
      if (lower && !h.surface) {
        nd = -1.f * h.n;
        p = p + (2.f * h.radius) * nd;
      }
      out.cont = 1;
      out.o = p + 1e-3f * nd;
      out.wi = wi;
      out.mul = mul;
      out.pdf = pdf;
    }
  }
  return out;
}

// Writes a vertex's shadow records at index i of the pending-light arrays,
// in the layout k_shadow and the resolve kernels read: (origin, 1 if the ray
// exists), direction, contribution per ray, and the throughput.
struct ShadowSink {
  float4 *o1, *d1, *c1, *o2, *d2, *c2, *beta;
  int i;
  bool has1, has2;
  HM_D ShadowSink(const PathBuffers& pb, int idx)
      : o1(pb.sh_o1), d1(pb.sh_d1), c1(pb.sh_c1), o2(pb.sh_o2), d2(pb.sh_d2), c2(pb.sh_c2), beta(pb.sh_beta),
        i(idx), has1(false), has2(false) {}
  HM_D void operator()(int which, V3 o, V3 d, V3 c) {
    float4* so = which ? o2 : o1;
    float4* sd = which ? d2 : d1;
    float4* scn = which ? c2 : c1;
    so[i] = make_float4(o.x, o.y, o.z, 1.f);
    sd[i] = make_float4(d.x, d.y, d.z, 0.f);
    scn[i] = make_float4(c.x, c.y, c.z, 0.f);
    if (which) has2 = true;
    else has1 = true;
  }
  // after the vertex: clears the flag of a ray the vertex did not make when
  // it made the other, and stores the throughput; returns whether any exists
  HM_D bool finish(V3 b) {
    if (!(has1 || has2)) return false;
    if (!has1) o1[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!has2) o2[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    beta[i] = make_float4(b.x, b.y, b.z, 0.f);
    return true;
  }
};

// Closest hit against the head sphere and the hair BVH. The sphere is tested
// first so its distance prunes the hair traversal; equal distances go to hair.
// kCurves: the hair BVH holds round Catmull-Rom segments (sc.curves), a
// separate instantiation so the linear-segment traversal keeps its registers.
template <bool kCount, bool kCurves = false>
HM_D void closest_hit(const SceneView& sc, V3 o, V3 d, TraceStack& st, float& t, int& prim, float& u,
                      TraceCounters& cnt) {
  t = __builtin_inff();
  prim = kPrimMiss;
  u = 0.f;
  bool sphere_hit = false;
  if (sc.sphere.enabled) {
    float ts;
    if (intersect_sphere(o, d, v3(sc.sphere.cx, sc.sphere.cy, sc.sphere.cz), sc.sphere.r, &ts)) {
      t = ts;
      sphere_hit = true;
    }
  }
  const RayInv r = make_ray(o.x, o.y, o.z, d.x, d.y, d.z);
  int tri = 0x7fffffff;
  if (sc.mesh.num > 0) {
    float bu = 0.f;
    trace_closest<kCount, TriLeaf>(sc.mesh.bvh, r, st, t, tri, bu, cnt, TriLeaf{sc.mesh.tri});
  }
  int id = 0x7fffffff;
  if constexpr (kCurves)
    trace_closest<kCount, CurveLeaf>(sc.bvh, r, st, t, id, u, cnt, CurveLeaf{sc.curves, sc.curve_caps});
  else
    trace_closest<kCount>(sc.bvh, r, st, t, id, u, cnt);
  if (id != 0x7fffffff) prim = id;
  else if (tri != 0x7fffffff) prim = kPrimMesh0 - tri;
  else if (sphere_hit) prim = kPrimSphere;
}

template <bool kCount, bool kCurves = false>
HM_D bool occluded(const SceneView& sc, V3 o, V3 d, TraceStack& st, TraceCounters& cnt) {
  if (sc.sphere.enabled) {
    float ts;
    if (intersect_sphere(o, d, v3(sc.sphere.cx, sc.sphere.cy, sc.sphere.cz), sc.sphere.r, &ts)) return true;
  }
  const RayInv r = make_ray(o.x, o.y, o.z, d.x, d.y, d.z);
  if (sc.mesh.num > 0 && trace_any<kCount, TriLeaf>(sc.mesh.bvh, r, st, cnt, TriLeaf{sc.mesh.tri})) return true;
  if constexpr (kCurves) return trace_any<kCount, CurveLeaf>(sc.bvh, r, st, cnt, CurveLeaf{sc.curves, sc.curve_caps});
  return trace_any<kCount>(sc.bvh, r, st, cnt);
}

#endif  // __HIPCC__

}  // namespace hm
