// Scene description: the reference's scene-config JSON schema
// (scene.cpp:119-339, headers/scene.h:21-82) plus additive keys for
// procedurally generated hair, head proxy and environment maps.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hm_common.h"

namespace hm {

struct SceneCamera {
  V3 from{0, 0, 0}, at{0, 0, 1}, up{0, 0, 1};
  float cos_fovy = 0.66f;
};

// hair.procedural {...}: seeded strand generator.
struct ProceduralHair {
  std::string kind = "ball";  // "ball": curl random walk; "curly": helices
  int strands = 2000;
  int segments = 16;
  float segment_length = 0.05f;  // ball only
  float radius = 0.005f;         // fiber radius
  float curl_sigma = 0.05f;      // ball only: Gaussian curl per segment
  float helix_radius = 0.02f;    // curly only
  float helix_pitch = 0.1f;      // curly only: rise per turn
  int segments_per_turn = 16;    // curly only
  float root_sphere_radius = 1.0f;
  uint64_t seed = 0;
};

// lights.environment.procedural {...}: log-normal texels plus Gaussian spots.
struct EnvHotspot {
  float theta, phi, width;
  V3 color;
};
struct ProceduralEnv {
  int width = 2048, height = 1024;
  float mu = 0.f, sigma = 1.5f;
  uint64_t seed = 1;
  std::vector<EnvHotspot> hotspots;
};

struct Scene {
  std::string json_path;
  std::string base_dir;

  // geometry
  bool has_surface = false;
  bool surface_sphere = false;
  V3 sphere_center{0, 0, 0};
  float sphere_radius = 1.f;
  V3 surface_albedo{0, 0, 0};  // head.mtl has no Kd
  float surface_alpha = 1.f;   // roughness before clamp + square (model.cpp:323)
  std::string surface_geometry;

  bool has_hair = false;
  bool hair_procedural = false;
  ProceduralHair hair_gen;
  std::string hair_geometry;
  int hair_tessellation = 4;  // linear pieces per Catmull-Rom segment of a .hair file ("linear")
  // .hair strands as round cubic Catmull-Rom curves (the reference's
  // primitive; hair.primitive "curve", the default) or as hair_tessellation
  // linear pieces per curve segment ("linear")
  bool hair_curves = true;

  // camera (required)
  SceneCamera camera;

  // hair material (defaults per scene.h:41-45 / scene.cpp:198-227)
  int bsdf = 0;
  V3 sig_a{0.06f, 0.1f, 0.2f};
  float beta_m = 0.3f, beta_n = 0.3f;
  float alpha = 0.03f;  // radians
  float R_G = 1.f, TT_G = 1.f, TRT_G = 1.f, TRRT_G = 1.f;
  float pixar_factor = 1.f;
  std::string dual_scatter_precomp;

  // environment light
  bool has_env = false;
  std::string env_exr;
  bool env_constant = false;
  V3 env_constant_rgb{0, 0, 0};
  int env_constant_w = 64, env_constant_h = 32;
  bool env_procedural = false;
  ProceduralEnv env_gen;
  float env_scale = 0.f, env_rot_phi = 0.f;

  // directional lights
  bool has_directional = false;
  std::vector<V3> dlight_from, dlight_emit;

  // integrator (required)
  bool mis = true, env_pdf = true;
  int spp = 1, path_v1 = 0, path_v2 = 0;
  int width = 1280, height = 1280;
  std::string image_output, stats_output;

  // tcnn block
  std::string nrc_config, nrc_config2, nrc_weights;
  bool nrc_train = true;

  // hairmsnn extension block
  int msnn_train_records = 16384;  // reference: 128 x 128 (render_hair_msnn.h:127-129)
  int msnn_warmup_steps = 64;      // fixed, replaces the 1 s wall-clock warm-up
  float msnn_lr = 1e-2f;
  std::string msnn_loss = "RelativeL2Luminance";
  std::string msnn_precision = "fp16";
  uint64_t msnn_seed = 7;

  // acceleration structure extension block: "lbvh" (Karras radix tree) or
  // "ploc" (locally-ordered clustering) binary tree, collapsed 4-wide
  std::string bvh_builder = "sah";  // accel.builder: "sah" (host, binned SAH), "lbvh", "ploc"
  int ploc_radius = 16;
  int bvh_max_leaf = 1;  // accel.max_leaf: primitives per leaf run of the SAH tree (1..8)
  int bvh_split = 1;     // accel.split: boxes per linear hair segment in the SAH tree (1..4)

  // neural radiance caching extension block (render_nrc.h:112-134 defaults)
  float nrc_c = 0.01f;              // spread threshold multiplier
  bool nrc_all_unbiased = false;
  int nrc_train_records = 65536;
};

// Parses a scene file. Returns false and fills `err` like the reference's
// parseScene returns false with a LOG line (scene.cpp:114-339).
bool parse_scene_file(const std::string& path, Scene& scene, std::string& err);
bool parse_scene_text(const std::string& text, const std::string& base_dir, Scene& scene,
                      std::string& err);

// Host-side generated content.
struct HairGeometry {
  // 8 floats per primitive: p0 xyz, r, p1 xyz, strand. Linear pieces, or for
  // curves each segment's chord (its two interpolated points, the radius at
  // u = 0.5), which per-strand lookups and point sampling use.
  std::vector<float> seg;
  // curves only: kCurveFloats per segment: the four Catmull-Rom control
  // points (x, y, z, radius), then strand, end flags (1: first segment of its
  // strand, 2: last), 0, 0
  std::vector<float> crv;
  std::vector<int32_t> strand_of_seg;
  int num_strands = 0;
  V3 bmin{1e30f, 1e30f, 1e30f}, bmax{-1e30f, -1e30f, -1e30f};
};
void generate_hair(const ProceduralHair& g, HairGeometry& out);
// Cem Yuksel .hair file (the format cyHairFile reads; scene.cpp:10-75 builds
// the curves): each strand's points, with a phantom point mirrored past each
// end, form uniform Catmull-Rom segments of radius 0.2 x thickness; every
// segment is split into `tess` linear pieces (sampled on the curve, radius
// at the piece midpoint).
// With curves, every segment is kept as one round Catmull-Rom curve instead.
constexpr int kCurveFloats = 20;
int load_hair_file(const std::string& path, int tess, bool curves, HairGeometry& out, std::string& err);

// Triangle mesh of an OBJ head (hm_obj.cpp): per triangle its three
// vertices, three vertex normals and the material's diffuse colour.
struct TriMesh {
  std::vector<float> v;   // 9 floats per triangle
  std::vector<float> n;   // 9 floats per triangle
  std::vector<float> kd;  // 3 floats per triangle
  V3 bmin{1e30f, 1e30f, 1e30f}, bmax{-1e30f, -1e30f, -1e30f};
  int num() const { return (int)(v.size() / 9); }
};
int load_obj(const std::string& path, TriMesh& out, std::string& err);

struct EnvImage {
  int width = 0, height = 0;
  std::vector<float> rgba;  // width*height*4
};
void generate_env(const Scene& s, EnvImage& out);
// Scanline OpenEXR reader (hm_image.cpp): RGBA floats, file row order.
int read_exr(const std::string& path, int& width, int& height, std::vector<float>& rgba, std::string& err);
// A scene-relative path (the scene file's directory) or an absolute one.
std::string scene_path(const Scene& s, const std::string& p);

}  // namespace hm
