// ORACLE - test infrastructure only (see orc_math.h).
//
// Environment tables and sampling, hair-segment / sphere intersection, a CPU
// BVH, and the path integrators, restating:
//   scene.cpp:349-425           generateEnvSamplingTables
//   optix_common.cuh:10-149     environment radiance / pdf / sampling
//   optix_common.cuh:151-409    sampleBSDF, sampleLights, directLighting,
//                               nextPathVertex, pathTrace
//   optix_common.cuh:527-599    closest-hit programs (hit frames)
//   curve_utils.cuh:159-197     hit-point refinement and normal (linear)
//   cuda/path_tracing.cu:19-67  camera ray and per-pixel seed
//   cuda/hair_msnn.cu:17-185    truncated and training paths
#pragma once

#include <cstdint>
#include <vector>

#include "orc_math.h"

namespace orc {

// ---- environment ----------------------------------------------------------
struct EnvTables {
  int W = 0, H = 0;
  std::vector<float> rgba;               // W*H*4
  std::vector<float> cpdf, ccdf;         // (W+1)*H
  std::vector<float> mpdf, mcdf;         // H+1
};
// mode 0: running sums in FP64, rounded once (the GPU build's scheme);
// mode 1: the reference's sequential FP32 sums (scene.cpp:349-425).
void env_tables(EnvTables& e, int mode);

// ---- geometry ---------------------------------------------------------------
// Ray / finite open cylinder around p0->p1 (constant radius r): nearest root
// t > 0 with axial u in [0,1]. IEEE FP32, no contraction (build flag), with
// the root solved around the ray point nearest the segment centre.
bool seg_intersect(const float o[3], const float d[3], const float* seg, float* t, float* u);
bool sphere_intersect(const float o[3], const float d[3], const float c[3], float r, float* t);
// Ray / triangle (Moller-Trumbore, double-sided), nearest t > 0 and the
// barycentrics (u, v) of vertices 1 and 2; IEEE FP32 like seg_intersect.
bool tri_intersect(const float o[3], const float d[3], const float* v, float* t, float* u, float* w);

// Round Catmull-Rom hair segments (orc_curve.cpp), kCrvFloats per segment:
// four control points (x, y, z, radius), strand, end flags (1: first
// segment of its strand, 2: last), 0, 0. crv_intersect: the GPU build's
// iteration in the same FP32 operations (nearest t > 0; u = 0 / 1 on a cap;
// flags: caps to test). crv_exact: FP64 first crossing of the swept-circle
// surface along the ray (min_clear: the smallest |distance - radius| seen,
// which tells grazing rays apart). crv_box: the box the BVH bounds it by.
constexpr int kCrvFloats = 20;
float crv_radius_bound(const float* q);
void crv_box(const float* q, float lo[3], float hi[3]);
// Position (x, y, z, radius) and its derivative at u: prec 1 the FP32
// evaluator the iteration uses, 0 the FP64 one of the brute-force search.
void crv_eval(const float* q, double u, int prec, double out[8]);
bool crv_intersect(const float o[3], const float d[3], const float* q, int flags, float* t, float* u);
bool crv_exact(const double o[3], const double d[3], const float* q, int flags, double* t, double* u,
               double* min_clear);

struct Bvh {
  struct Node {
    float lo[3], hi[3];
    int left, right;  // children when count == 0
    int first, count; // leaf range into idx
  };
  std::vector<Node> nodes;
  std::vector<int> idx;
  const float* seg = nullptr;  // 8 floats per segment (curves: their boxes)
  int nseg = 0;
  const float* crv = nullptr;  // round curves, or null for segments
  std::vector<float> boxes;
  void build(const float* seg, int n);
  void build_curves(const float* crv, int n);
  // minimum (t, id) over all primitives, like a brute-force scan; caps: the
  // curves' end caps are tested (plain path tracing) or not (HairMSNN, NRC)
  bool closest(const float o[3], const float d[3], float* t, int* id, float* u, int caps = 1) const;
  bool any(const float o[3], const float d[3], int caps = 1) const;
  bool leaf(int id, const float o[3], const float d[3], int caps, float* t, float* u) const;
};

// ---- scene and integrators --------------------------------------------------
struct SceneDesc {
  // geometry
  const float* seg;  // 8 floats per segment (with curves: each one's chord)
  int nseg;
  int has_sphere;
  float sphere[4];
  float sphere_albedo[3];
  float sphere_alpha;  // already clamped and squared
  // environment
  int has_env, env_pdf;
  float env_scale, env_rot;
  const EnvTables* env;
  // lights
  int ndl;
  float dl_from[8][3], dl_emit[8][3];
  // camera
  float cam[12];
  int width, height;
  // material and integrator
  float beta_m, beta_n, alpha, gains[4];
  float sigma_a[3];
  const float* strand_sigma;  // optional, 3 per strand
  int mis;
  int path_v1, path_v2;
  // Frozen sampling, a test aid for the adjoint: when set, directions, the
  // sampling pdfs in the estimators' denominators and Russian roulette use
  // (beta_m0, beta_n0, alpha0, strand_sigma0 or sigma_a), while f and the MIS
  // weights use the parameters above. Only pathTrace (mode 0) honours it.
  int frozen;
  float beta_m0, beta_n0, alpha0;
  const float* strand_sigma0;
  // OBJ head mesh: per triangle 3 vertices (9 floats), 3 vertex normals (9)
  // and the diffuse colour (3); ntri = 0 without a mesh
  const float* tri_v;
  const float* tri_n;
  const float* tri_kd;
  int ntri;
  // round Catmull-Rom hair (kCrvFloats per segment, same ids as seg), or null
  const float* crv;
};

// mode 0: pathTrace; mode 1: msnnPathTrace truncated at kernel beta;
// mode 2: msnnTrainingPath (full and short colours).
template <class R>
void render_pixels(const SceneDesc& s, const Bvh& bvh, int mode, int kbeta, int accum_id0, int spp, int x0,
                   int y0, int w, int h, int threads, double* out_full, double* out_short, uint8_t* out_hit);

// HairMSNN G_BUFFER training records (cuda/hair_msnn.cu:203-276) of frame
// accum_id at the given pixels: the 12-float network input at the primary
// hit (position / scene_scale, wo, tangent, absorption), the target
// full - short after the reference's scrub (NaN full -> 0, Inf full -> 1e5,
// NaN short -> 0.01, Inf short -> 1e5; both 0 without a hit), the full and
// short colours after the scrub, and the hit flags (bit 0 hit, bit 1
// surface).
struct MsnnRecordOut {
  float input[12] = {0};
  double target[3] = {0, 0, 0}, full[3] = {0, 0, 0}, shortc[3] = {0, 0, 0};
  uint8_t hit = 0;
};
template <class R>
void msnn_records(const SceneDesc& s, const Bvh& bvh, int accum_id, int kbeta, int n, const int* pixels,
                  float scene_scale, int threads, MsnnRecordOut* out);

// Neural radiance caching G_BUFFER pass (cuda/nrc.cu:135-343, 357-370) for
// one frame: per pixel, the radiance before the cache query, the throughput
// at it and the query record; for training pixels (flags bit 0; bit 1:
// unbiased suffix) the per-vertex records and the suffix query.
constexpr int kNrcMaxBounces = 40;
struct NrcPixelOut {
  uint8_t gb_hit = 0, query = 0, tb_hit = 0;
  int bounces = 0;
  double rad[3] = {0, 0, 0}, beta[3] = {0, 0, 0};
  float q[9] = {0}, sq[9] = {0};
  double rec[kNrcMaxBounces * 15] = {0};  // per vertex: pos / scale, wo, n, direct light, f cos / pdf
};
template <class R>
void nrc_gbuffer(const SceneDesc& s, const Bvh& bvh, int accum_id, const uint8_t* flags, float c, float scene_scale,
                 int threads, NrcPixelOut* out);

// Path-replay adjoint of pathTrace in FP64 (BSDF derivatives by central
// differences of step fd_h): per-strand d/d sigma_a, d/d (beta_m, beta_n,
// alpha) and each pixel's share of the latter, for dL/d(image) dl_dimg
// (W x H x 4) of the mean of passes 0..spp-1.
void adjoint_render(const SceneDesc& s, const Bvh& bvh, int spp, const float* dl_dimg, int num_strands, int threads,
                    double fd_h, double* grad_sigma, double* grad_glob, double* grad_px);

// Probes of single integrator functions, for the oracle's own tests:
// env radiance and pdf per world direction (optix_common.cuh:10-64); env
// samples from one LCG per seed (:66-149); the hit setup of a closest hit
// (hairCH / triangleMeshCH, optix_common.cuh:527-599), kProbeHit doubles per
// hit: p, n, frame rows X, Y, Z, wo_local, h = dot(Y, n), radius; the mean of
// directLighting (optix_common.cuh:279-309) over one LCG per seed at one hit
// (mean_rgb: mean RGB, then mean square RGB).
constexpr int kProbeHit = 20;
template <class R>
void probe_env(const SceneDesc& s, int n, const float* dirs, double* rgb, double* pdf);
template <class R>
void probe_env_sample(const SceneDesc& s, int n, const uint32_t* seeds, double* wi, double* rgb, double* pdf);
template <class R>
void probe_hit(const SceneDesc& s, const Bvh& bvh, int n, const float* rays, const float* t, const int* prim,
               const float* u, double* out);
template <class R>
void probe_direct(const SceneDesc& s, const Bvh& bvh, const float ray[8], float t, int prim, float u, int n,
                  const uint32_t* seeds, int threads, double* mean_rgb);

}  // namespace orc
