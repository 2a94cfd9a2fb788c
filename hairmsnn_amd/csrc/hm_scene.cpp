// Scene-config parsing with the reference's defaults and failure rules, and
// the seeded procedural generators for hair, head proxy and environment.
#include "hm_scene.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/hmsnn.h"
#include "hm_json.h"

namespace hm {

namespace {

V3 vec3_of(const Json& a) { return v3(a.at(0).f(), a.at(1).f(), a.at(2).f()); }

std::string dir_of(const std::string& path) {
  size_t k = path.find_last_of("/\\");
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

// Deterministic generator for procedural content: splitmix64 stream.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double gauss() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.141592653589793 * u2);
  }
};

void parse_procedural_hair(const Json& j, ProceduralHair& g) {
  if (j.has("kind")) g.kind = j.at("kind").str();
  if (j.has("strands")) g.strands = j.at("strands").i();
  if (j.has("segments")) g.segments = j.at("segments").i();
  if (j.has("segment_length")) g.segment_length = j.at("segment_length").f();
  if (j.has("radius")) g.radius = j.at("radius").f();
  if (j.has("curl_sigma")) g.curl_sigma = j.at("curl_sigma").f();
  if (j.has("helix_radius")) g.helix_radius = j.at("helix_radius").f();
  if (j.has("helix_pitch")) g.helix_pitch = j.at("helix_pitch").f();
  if (j.has("segments_per_turn")) g.segments_per_turn = j.at("segments_per_turn").i();
  if (j.has("root_sphere_radius")) g.root_sphere_radius = j.at("root_sphere_radius").f();
  if (j.has("seed")) g.seed = (uint64_t)j.at("seed").num();
  if (g.kind != "ball" && g.kind != "curly") throw JsonError("hair.procedural.kind must be 'ball' or 'curly'");
  if (g.strands <= 0 || g.segments <= 0) throw JsonError("hair.procedural needs strands > 0 and segments > 0");
  if (g.segments_per_turn <= 0) throw JsonError("hair.procedural.segments_per_turn must be > 0");
}

void parse_procedural_env(const Json& j, ProceduralEnv& g) {
  if (j.has("width")) g.width = j.at("width").i();
  if (j.has("height")) g.height = j.at("height").i();
  if (j.has("mu")) g.mu = j.at("mu").f();
  if (j.has("sigma")) g.sigma = j.at("sigma").f();
  if (j.has("seed")) g.seed = (uint64_t)j.at("seed").num();
  g.hotspots.clear();
  if (j.has("hotspots")) {
    for (const Json& h : j.at("hotspots").items()) {
      EnvHotspot s;
      s.theta = h.at("theta").f();
      s.phi = h.at("phi").f();
      s.width = h.at("width").f();
      s.color = vec3_of(h.at("color"));
      g.hotspots.push_back(s);
    }
  }
  if (g.width <= 0 || g.height <= 0) throw JsonError("environment.procedural needs a positive size");
}

bool parse_json(const Json& cfg, Scene& scene, std::string& err) {
  // ---- geometry (scene.cpp:137-169) ----
  bool is_surface = false, is_hair = false;
  if (cfg.has("surface") && cfg.at("surface").is_object()) {
    const Json& s = cfg.at("surface");
    if (s.has("sphere")) {
      const Json& sp = s.at("sphere");
      scene.surface_sphere = true;
      scene.sphere_center = sp.has("center") ? vec3_of(sp.at("center")) : v3(0.f);
      scene.sphere_radius = sp.has("radius") ? sp.at("radius").f() : 1.f;
      scene.surface_albedo = sp.has("albedo") ? vec3_of(sp.at("albedo")) : v3(0.f);
      scene.surface_alpha = sp.has("alpha") ? sp.at("alpha").f() : 1.f;
      is_surface = true;
    } else if (s.has("geometry") && s.at("geometry").is_string()) {
      scene.surface_geometry = s.at("geometry").str();
      is_surface = true;
    }
  }
  scene.has_surface = is_surface;
  if (cfg.has("hair") && cfg.at("hair").is_object()) {
    const Json& h = cfg.at("hair");
    if (h.has("procedural")) {
      parse_procedural_hair(h.at("procedural"), scene.hair_gen);
      scene.hair_procedural = true;
      is_hair = true;
    } else if (h.has("geometry") && h.at("geometry").is_string()) {
      scene.hair_geometry = h.at("geometry").str();
      is_hair = true;
    }
    if (h.has("tessellation")) scene.hair_tessellation = std::max(1, std::min(64, h.at("tessellation").i()));
    if (h.has("primitive") && h.at("primitive").is_string()) {
      const std::string prim = h.at("primitive").str();
      if (prim != "curve" && prim != "linear") {
        err = "hair.primitive must be \"curve\" or \"linear\"";
        return false;
      }
      scene.hair_curves = prim == "curve";
    }
  }
  scene.has_hair = is_hair;
  if (!is_surface && !is_hair) {
    err = "Either hair or surface must be defined!";
    return false;
  }

  // ---- camera (scene.cpp:174-188), required ----
  try {
    const Json& c = cfg.at("camera");
    scene.camera.from = vec3_of(c.at("from"));
    scene.camera.at = vec3_of(c.at("to"));
    scene.camera.up = vec3_of(c.at("up"));
    scene.camera.cos_fovy = c.at("cos_fovy").f();
  } catch (const JsonError&) {
    err = "Camera must be defined!";
    return false;
  }

  // ---- hair material (scene.cpp:193-231) ----
  // A missing "type" abandons the whole block and keeps the struct defaults
  // (alpha = 0.03 rad); otherwise each key falls back on its own default.
  try {
    const Json& b = cfg.at("hair");
    scene.bsdf = b.at("type").i();
    auto getf = [&](const char* k, float dflt) {
      try {
        return b.at(k).f();
      } catch (const JsonError&) {
        return dflt;
      }
    };
    try {
      scene.sig_a = vec3_of(b.at("sigma_a"));
    } catch (const JsonError&) {
      scene.sig_a = v3(0.06f, 0.1f, 0.2f);
    }
    scene.beta_m = getf("beta_m", 0.3f);
    scene.beta_n = getf("beta_n", 0.3f);
    try {
      scene.alpha = kHostPi * b.at("alpha").f() / 180.f;
    } catch (const JsonError&) {
      scene.alpha = 0.f;
    }
    scene.R_G = getf("Gain R", 1.f);
    scene.TT_G = getf("Gain TT", 1.f);
    scene.TRT_G = getf("Gain TRT", 1.f);
    scene.TRRT_G = getf("Gain TRRT", 1.f);
    scene.pixar_factor = getf("pixar_factor", 1.f);
    try {
      scene.dual_scatter_precomp = b.at("dual_scatter_precomputation").str();
    } catch (const JsonError&) {
      scene.dual_scatter_precomp = "";
    }
  } catch (const JsonError&) {
    // "Hair material not defined. Using default values..."
  }

  // ---- environment light (scene.cpp:236-253) ----
  bool is_env = false;
  scene.has_env = false;
  try {
    const Json& e = cfg.at("lights").at("environment");
    if (e.has("constant")) {
      scene.env_constant = true;
      scene.env_constant_rgb = vec3_of(e.at("constant"));
      if (e.has("resolution")) {
        scene.env_constant_w = e.at("resolution").at(0).i();
        scene.env_constant_h = e.at("resolution").at(1).i();
      }
    } else if (e.has("procedural")) {
      scene.env_procedural = true;
      parse_procedural_env(e.at("procedural"), scene.env_gen);
    } else {
      scene.env_exr = e.at("exr").str();
    }
    is_env = true;
    scene.has_env = true;
    scene.env_scale = e.at("scale").f();  // a missing scale leaves 0 (env still on)
    try {
      scene.env_rot_phi = e.at("rotation").f();
    } catch (const JsonError&) {
      scene.env_rot_phi = 0.f;
    }
  } catch (const JsonError&) {
    // "No environment light found"
  }

  // ---- directional lights (scene.cpp:258-276) ----
  // An absent list iterates as empty and still counts as "defined".
  bool is_directional = false;
  scene.has_directional = false;
  try {
    if (cfg.has("lights") && cfg.at("lights").has("directional")) {
      const Json& dl = cfg.at("lights").at("directional");
      if (dl.is_array()) {
        for (const Json& d : dl.items()) {
          V3 from = normalize(vec3_of(d.at("from")));
          V3 emit = vec3_of(d.at("emit"));
          scene.dlight_from.push_back(from);
          scene.dlight_emit.push_back(emit);
        }
      } else if (!dl.is_null()) {
        throw JsonError("directional lights must be a list");
      }
    }
    scene.has_directional = true;
    is_directional = true;
  } catch (const JsonError&) {
    // "No directional lights found"
  }
  if (!is_directional && !is_env) {
    err = "Either directional or environment light must be defined!";
    return false;
  }

  // ---- integrator (scene.cpp:286-311), required ----
  try {
    const Json& in = cfg.at("integrator");
    scene.spp = in.at("spp").i();
    scene.path_v1 = in.at("path_v1").i();
    scene.path_v2 = in.at("path_v2").i();
    scene.width = in.at("width").i();
    scene.height = in.at("height").i();
    scene.image_output = in.at("image_output").str();
    scene.stats_output = in.at("stats_output").str();
    scene.mis = in.at("MIS").b();
    scene.env_pdf = in.at("ENV_PDF").b();
  } catch (const JsonError&) {
    err = "Integrator must be defined!";
    return false;
  }
  const long long px = (long long)scene.width * (long long)scene.height;
  if (scene.width <= 0 || scene.height <= 0 || px > 2048LL * 2048LL || px % 128 != 0) {
    err = "Image size has a hard limit of 2K*2K!";
    return false;
  }

  // ---- tcnn (scene.cpp:316-335) ----
  try {
    const Json& t = cfg.at("tcnn");
    scene.nrc_config = t.at("config").str();
    try {
      scene.nrc_config2 = t.at("config2").str();
    } catch (const JsonError&) {
      scene.nrc_config2 = scene.nrc_config;
    }
    try {
      scene.nrc_train = t.at("init_train").b();
    } catch (const JsonError&) {
      scene.nrc_train = false;
    }
    try {
      scene.nrc_weights = t.at("init_weights").str();
    } catch (const JsonError&) {
      scene.nrc_weights = "";
    }
  } catch (const JsonError&) {
    // "TCNN not defined."
  }

  // ---- hairmsnn extension ----
  if (cfg.has("hairmsnn") && cfg.at("hairmsnn").is_object()) {
    const Json& m = cfg.at("hairmsnn");
    if (m.has("train_records")) scene.msnn_train_records = m.at("train_records").i();
    if (m.has("warmup_steps")) scene.msnn_warmup_steps = m.at("warmup_steps").i();
    if (m.has("learning_rate")) scene.msnn_lr = m.at("learning_rate").f();
    if (m.has("loss")) scene.msnn_loss = m.at("loss").str();
    if (m.has("precision")) scene.msnn_precision = m.at("precision").str();
    if (m.has("seed")) scene.msnn_seed = (uint64_t)m.at("seed").num();
  }
  // ---- acceleration structure extension ----
  if (cfg.has("accel") && cfg.at("accel").is_object()) {
    const Json& a = cfg.at("accel");
    if (a.has("builder")) scene.bvh_builder = a.at("builder").str();
    if (a.has("ploc_radius")) scene.ploc_radius = a.at("ploc_radius").i();
    if (a.has("max_leaf")) scene.bvh_max_leaf = a.at("max_leaf").i();
    if (scene.bvh_max_leaf < 1 || scene.bvh_max_leaf > 8) {
      err = "accel.max_leaf must be 1..8";
      return false;
    }
    if (a.has("split")) scene.bvh_split = a.at("split").i();
    if (scene.bvh_split < 1 || scene.bvh_split > 4) {
      err = "accel.split must be 1..4";
      return false;
    }
    if (scene.bvh_builder != "lbvh" && scene.bvh_builder != "ploc" && scene.bvh_builder != "sah") {
      err = "accel.builder must be \"lbvh\", \"ploc\" or \"sah\"";
      return false;
    }
  }
  // ---- neural radiance caching extension ----
  if (cfg.has("nrc") && cfg.at("nrc").is_object()) {
    const Json& m = cfg.at("nrc");
    if (m.has("c")) scene.nrc_c = m.at("c").f();
    if (m.has("all_unbiased")) scene.nrc_all_unbiased = m.at("all_unbiased").b();
    if (m.has("train_records")) scene.nrc_train_records = m.at("train_records").i();
  }
  return true;
}

}  // namespace

bool parse_scene_text(const std::string& text, const std::string& base_dir, Scene& scene,
                      std::string& err) {
  Json cfg;
  try {
    cfg = Json::parse(text);
  } catch (const JsonError& e) {
    err = std::string("Could not load scene .json file! (") + e.what() + ")";
    return false;
  }
  scene.base_dir = base_dir;
  try {
    return parse_json(cfg, scene, err);
  } catch (const JsonError& e) {
    err = e.what();
    return false;
  }
}

bool parse_scene_file(const std::string& path, Scene& scene, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "Could not load scene .json file! (cannot open " + path + ")";
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  scene.json_path = path;
  return parse_scene_text(ss.str(), dir_of(path), scene, err);
}

// ---------------------------------------------------------------------------
// Procedural hair. Roots are area-uniform on a sphere; "ball" strands grow
// along the normal with a Gaussian curl of sigma per segment, "curly"
// strands are helices around the normal with a random phase.
void generate_hair(const ProceduralHair& g, HairGeometry& out) {
  SplitMix rng{g.seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
  out.seg.assign((size_t)g.strands * g.segments * 8, 0.f);
  out.strand_of_seg.assign((size_t)g.strands * g.segments, 0);
  out.num_strands = g.strands;
  out.bmin = v3(1e30f);
  out.bmax = v3(-1e30f);
  const double twopi = 2.0 * 3.141592653589793;
  std::vector<double> px(g.segments + 1), py(g.segments + 1), pz(g.segments + 1);
  for (int s = 0; s < g.strands; ++s) {
    const double z = 1.0 - 2.0 * rng.uniform();
    const double ph = twopi * rng.uniform();
    const double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    const double nx = rr * std::cos(ph), ny = rr * std::sin(ph), nz = z;
    const double R = g.root_sphere_radius;
    px[0] = R * nx;
    py[0] = R * ny;
    pz[0] = R * nz;
    if (g.kind == "ball") {
      double dx = nx, dy = ny, dz = nz;
      for (int k = 0; k < g.segments; ++k) {
        dx += g.curl_sigma * rng.gauss();
        dy += g.curl_sigma * rng.gauss();
        dz += g.curl_sigma * rng.gauss();
        const double inv = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        dx *= inv;
        dy *= inv;
        dz *= inv;
        px[k + 1] = px[k] + g.segment_length * dx;
        py[k + 1] = py[k] + g.segment_length * dy;
        pz[k + 1] = pz[k] + g.segment_length * dz;
      }
    } else {
      // orthonormal pair around the normal
      double ax = std::fabs(nx) < 0.9 ? 1.0 : 0.0, ay = ax == 1.0 ? 0.0 : 1.0, az = 0.0;
      double e1x = ay * nz - az * ny, e1y = az * nx - ax * nz, e1z = ax * ny - ay * nx;
      double inv = 1.0 / std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
      e1x *= inv;
      e1y *= inv;
      e1z *= inv;
      const double e2x = ny * e1z - nz * e1y, e2y = nz * e1x - nx * e1z, e2z = nx * e1y - ny * e1x;
      const double phase = twopi * rng.uniform();
      const double c0 = std::cos(phase), s0 = std::sin(phase);
      for (int k = 1; k <= g.segments; ++k) {
        const double ang = phase + twopi * (double)k / (double)g.segments_per_turn;
        const double rise = g.helix_pitch * (double)k / (double)g.segments_per_turn;
        const double ca = std::cos(ang) - c0, sa = std::sin(ang) - s0;
        px[k] = px[0] + rise * nx + g.helix_radius * (ca * e1x + sa * e2x);
        py[k] = py[0] + rise * ny + g.helix_radius * (ca * e1y + sa * e2y);
        pz[k] = pz[0] + rise * nz + g.helix_radius * (ca * e1z + sa * e2z);
      }
    }
    for (int k = 0; k < g.segments; ++k) {
      const size_t id = (size_t)s * g.segments + k;
      float* q = &out.seg[id * 8];
      q[0] = (float)px[k];
      q[1] = (float)py[k];
      q[2] = (float)pz[k];
      q[3] = g.radius;
      q[4] = (float)px[k + 1];
      q[5] = (float)py[k + 1];
      q[6] = (float)pz[k + 1];
      q[7] = (float)s;
      out.strand_of_seg[id] = s;
    }
    for (int k = 0; k <= g.segments; ++k) {
      const V3 p = v3((float)px[k], (float)py[k], (float)pz[k]);
      out.bmin = vmin(out.bmin, p);
      out.bmax = vmax(out.bmax, p);
    }
  }
}

int load_hair_file(const std::string& path, int tess, bool curves, HairGeometry& out, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "Error: Cannot open hair file! (" + path + ")";
    return HMSNN_EIO;
  }
  std::vector<uint8_t> b;
  {
    uint8_t tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) b.insert(b.end(), tmp, tmp + k);
    std::fclose(f);
  }
  // header: signature, hair_count, point_count, arrays, d_segments,
  // d_thickness, d_transparency, d_color[3], info[88]  (128 bytes)
  if (b.size() < 128) {
    err = "Error: Cannot read hair file header! (" + path + ")";
    return HMSNN_EIO;
  }
  if (std::memcmp(b.data(), "HAIR", 4) != 0) {
    err = "Error: File has wrong signature! (" + path + ")";
    return HMSNN_EIO;
  }
  auto u32 = [&](size_t o) {
    uint32_t v;
    std::memcpy(&v, b.data() + o, 4);
    return v;
  };
  auto f32 = [&](size_t o) {
    float v;
    std::memcpy(&v, b.data() + o, 4);
    return v;
  };
  const uint32_t hair_count = u32(4), point_count = u32(8), arrays = u32(12), d_segments = u32(16);
  const float d_thickness = f32(20);
  constexpr uint32_t kSegments = 1, kPoints = 2, kThickness = 4, kTransparency = 8, kColors = 16;
  if (!(arrays & kPoints)) {
    err = "Error: Cannot read hair points! (" + path + ": no point array)";
    return HMSNN_EIO;
  }
  size_t o = 128;
  const uint16_t* segs = nullptr;
  if (arrays & kSegments) {
    if (o + 2 * (size_t)hair_count > b.size()) {
      err = "Error: Cannot read hair segments! (" + path + ")";
      return HMSNN_EIO;
    }
    segs = reinterpret_cast<const uint16_t*>(b.data() + o);
    o += 2 * (size_t)hair_count;
  }
  if (o + 12 * (size_t)point_count > b.size()) {
    err = "Error: Cannot read hair points! (" + path + ")";
    return HMSNN_EIO;
  }
  std::vector<float> pts((size_t)point_count * 3);
  std::memcpy(pts.data(), b.data() + o, 12 * (size_t)point_count);
  o += 12 * (size_t)point_count;
  std::vector<float> thick;
  if (arrays & kThickness) {
    if (o + 4 * (size_t)point_count > b.size()) {
      err = "Error: Cannot read hair thickness! (" + path + ")";
      return HMSNN_EIO;
    }
    thick.resize(point_count);
    std::memcpy(thick.data(), b.data() + o, 4 * (size_t)point_count);
    o += 4 * (size_t)point_count;
  }
  if ((arrays & kTransparency) && o + 4 * (size_t)point_count > b.size()) {
    err = "Error: Cannot read hair transparency! (" + path + ")";
    return HMSNN_EIO;
  }
  if (arrays & kTransparency) o += 4 * (size_t)point_count;
  if ((arrays & kColors) && o + 12 * (size_t)point_count > b.size()) {
    err = "Error: Cannot read hair colors! (" + path + ")";
    return HMSNN_EIO;
  }
  out = HairGeometry{};
  out.bmin = v3(1e30f);
  out.bmax = v3(-1e30f);
  tess = std::max(1, tess);
  size_t first = 0;
  std::vector<V3> cp;
  std::vector<float> cw;
  for (uint32_t sidx = 0; sidx < hair_count; ++sidx) {
    const int nseg = segs ? (int)segs[sidx] : (int)d_segments;
    if (first + (size_t)nseg + 1 > point_count) {
      err = "Error: Cannot read hair points! (" + path + ": strands need more points than the file has)";
      return HMSNN_EIO;
    }
    auto P = [&](size_t i) { return v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
    if (nseg < 1) {
      // a strand of 0 segments still owns one point (the reference's point
      // loop consumes segments + 1 points per strand, scene.cpp:24-66) and
      // that point counts toward the bounds; it contributes no curve but
      // keeps its strand index, so strand ids stay the file's
      out.bmin = vmin(out.bmin, P(first));
      out.bmax = vmax(out.bmax, P(first));
      out.num_strands++;
      first += 1;
      continue;
    }
    cp.clear();
    cw.clear();
    auto Wd = [&](size_t i) { return 0.2f * (thick.empty() ? d_thickness : thick[i]); };
    // phantom points mirrored past the ends (scene.cpp:47-66)
    cp.push_back(P(first) + (P(first) - P(first + 1)));
    cw.push_back(Wd(first));
    for (int k = 0; k <= nseg; ++k) {
      cp.push_back(P(first + k));
      cw.push_back(Wd(first + k));
      out.bmin = vmin(out.bmin, P(first + k));
      out.bmax = vmax(out.bmax, P(first + k));
    }
    cp.push_back(P(first + nseg) + (P(first + nseg) - P(first + nseg - 1)));
    cw.push_back(Wd(first + nseg));
    const int strand = out.num_strands++;
    for (int k = 0; k < nseg; ++k) {
      // uniform Catmull-Rom through cp[k+1] .. cp[k+2]
      const V3 p0 = cp[k], p1 = cp[k + 1], p2 = cp[k + 2], p3 = cp[k + 3];
      auto curve = [&](float t) {
        const float t2 = t * t, t3 = t2 * t;
        return 0.5f * ((2.f * p1) + (p2 - p0) * t + (2.f * p0 - 5.f * p1 + 4.f * p2 - p3) * t2 +
                       (3.f * p1 - p0 - 3.f * p2 + p3) * t3);
      };
      auto width = [&](float t) {
        const float t2 = t * t, t3 = t2 * t;
        return 0.5f * ((2.f * cw[k + 1]) + (cw[k + 2] - cw[k]) * t +
                       (2.f * cw[k] - 5.f * cw[k + 1] + 4.f * cw[k + 2] - cw[k + 3]) * t2 +
                       (3.f * cw[k + 1] - cw[k] - 3.f * cw[k + 2] + cw[k + 3]) * t3);
      };
      if (curves) {
        // the segment itself (render_path_tracing.cu:236-244 hands the
        // control points and widths to the curve builder), and its chord
        const int flags = (k == 0 ? 1 : 0) | (k == nseg - 1 ? 2 : 0);
        const float rec[kCurveFloats] = {p0.x, p0.y, p0.z, cw[k],     p1.x,          p1.y,          p1.z,
                                         cw[k + 1], p2.x, p2.y, p2.z, cw[k + 2],     p3.x,          p3.y,
                                         p3.z, cw[k + 3], (float)strand, (float)flags, 0.f,           0.f};
        out.crv.insert(out.crv.end(), rec, rec + kCurveFloats);
        const float chord[8] = {p1.x, p1.y, p1.z, std::max(0.f, width(0.5f)), p2.x, p2.y, p2.z, (float)strand};
        out.seg.insert(out.seg.end(), chord, chord + 8);
        out.strand_of_seg.push_back(strand);
        continue;
      }
      V3 a = p1;
      for (int j = 1; j <= tess; ++j) {
        const V3 e = j == tess ? p2 : curve((float)j / (float)tess);
        const float r = std::max(0.f, width(((float)j - 0.5f) / (float)tess));
        const float rec[8] = {a.x, a.y, a.z, r, e.x, e.y, e.z, (float)strand};
        out.seg.insert(out.seg.end(), rec, rec + 8);
        out.strand_of_seg.push_back(strand);
        a = e;
      }
    }
    first += (size_t)nseg + 1;
  }
  if (out.seg.empty()) {
    err = "Error: hair file " + path + " holds no segments";
    return HMSNN_EIO;
  }
  return HMSNN_OK;
}

// Procedural / constant environment maps (RGBA32F, row y = theta bin).
std::string scene_path(const Scene& s, const std::string& p) {
  if (p.empty() || p[0] == '/' || s.base_dir.empty()) return p;
  return s.base_dir + "/" + p;
}

void generate_env(const Scene& s, EnvImage& out) {
  if (s.env_constant) {
    out.width = s.env_constant_w;
    out.height = s.env_constant_h;
    out.rgba.assign((size_t)out.width * out.height * 4, 1.f);
    for (size_t i = 0; i < (size_t)out.width * out.height; ++i) {
      out.rgba[i * 4 + 0] = s.env_constant_rgb.x;
      out.rgba[i * 4 + 1] = s.env_constant_rgb.y;
      out.rgba[i * 4 + 2] = s.env_constant_rgb.z;
    }
    return;
  }
  const ProceduralEnv& g = s.env_gen;
  out.width = g.width;
  out.height = g.height;
  out.rgba.assign((size_t)g.width * g.height * 4, 1.f);
  SplitMix rng{g.seed * 0x9E3779B97F4A7C15ull + 0xABCDEFull};
  const double pi = 3.141592653589793;
  std::vector<double> hx, hy, hz;
  for (const auto& h : g.hotspots) {
    hx.push_back(std::sin(h.theta) * std::cos(h.phi));
    hy.push_back(std::sin(h.theta) * std::sin(h.phi));
    hz.push_back(std::cos(h.theta));
  }
  for (int y = 0; y < g.height; ++y) {
    const double th = pi * (y + 0.5) / g.height;
    for (int x = 0; x < g.width; ++x) {
      const double ph = 2.0 * pi * (x + 0.5) / g.width;
      const double L = std::exp(g.mu + g.sigma * rng.gauss());
      double r = L, gg = L, b = L;
      const double dx = std::sin(th) * std::cos(ph), dy = std::sin(th) * std::sin(ph), dz = std::cos(th);
      for (size_t k = 0; k < g.hotspots.size(); ++k) {
        const double c = std::fmin(1.0, std::fmax(-1.0, dx * hx[k] + dy * hy[k] + dz * hz[k]));
        const double ang = std::acos(c);
        const double w = g.hotspots[k].width;
        const double a = std::exp(-0.5 * ang * ang / (w * w));
        r += a * g.hotspots[k].color.x;
        gg += a * g.hotspots[k].color.y;
        b += a * g.hotspots[k].color.z;
      }
      float* q = &out.rgba[((size_t)y * g.width + x) * 4];
      q[0] = (float)r;
      q[1] = (float)gg;
      q[2] = (float)b;
      q[3] = 1.f;
    }
  }
}

}  // namespace hm
