// C ABI (include/hmsnn.h): library plumbing, scene contexts, path tracing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"
#include "hm_context.h"

namespace hm {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int bsdf_eval_launch(const HairConsts& hc, const void* q, int n, float* f, float* pdf, hipStream_t s,
                     std::string& err);
int bsdf_sample_launch(const HairConsts& hc, const void* q, const float* u4, int n, float* wi, float* f,
                       float* pdf, int* lobe, hipStream_t s, std::string& err);

// Directional lights: (from, 0), (emit, 0) per light; the direction is
// normalised here with the device's own arithmetic, as sampleLights
// normalises the stored direction at every use (optix_common.cuh:236).
__global__ void k_normalize_dlights(float4* dl, int n) {
  const int i = threadIdx.x;
  if (i >= n) return;
  const float4 f = dl[2 * i];
  const V3 d = normalize(v3(f.x, f.y, f.z));
  dl[2 * i] = make_float4(d.x, d.y, d.z, 0.f);
}

static HairConsts consts_from(const hmsnn_hair_params* p) {
  return make_hair_consts(p->beta_m, p->beta_n, p->alpha, p->gain);
}

// Pinhole camera (render_path_tracing.cu:767-799): the image plane spans
// cos_fovy vertically at unit distance; pixel row 0 is at the bottom.
void camera_vectors(const Scene& scene, V3& pos, V3& d00_out, V3& du, V3& dv) {
  const SceneCamera& cam = scene.camera;
  pos = cam.from;
  V3 d00 = normalize(cam.at - cam.from);
  const float aspect = (float)scene.width / (float)scene.height;
  du = cam.cos_fovy * aspect * normalize(cross(d00, cam.up));
  dv = cam.cos_fovy * normalize(cross(du, d00));
  d00 = d00 - 0.5f * du;
  d00 = d00 - 0.5f * dv;
  d00_out = d00;
}

// Scene-derived fields of hmsnn_info (no generated content).
static void fill_scene_info(const Scene& s, hmsnn_info* info) {
  info->width = s.width;
  info->height = s.height;
  info->spp = s.spp;
  info->path_v1 = s.path_v1;
  info->path_v2 = s.path_v2;
  info->num_dir_lights = (int)s.dlight_from.size();
  info->has_env = s.has_env;
  info->has_sphere = s.has_surface && s.surface_sphere;
  info->mis = s.mis;
  info->env_pdf = s.env_pdf;
  info->sigma_a[0] = s.sig_a.x;
  info->sigma_a[1] = s.sig_a.y;
  info->sigma_a[2] = s.sig_a.z;
  info->beta_m = s.beta_m;
  info->beta_n = s.beta_n;
  info->alpha = s.alpha;
  info->gain[0] = s.R_G;
  info->gain[1] = s.TT_G;
  info->gain[2] = s.TRT_G;
  info->gain[3] = s.TRRT_G;
  info->env_scale = s.env_scale;
  info->env_rot_phi = s.env_rot_phi;
  info->sphere[0] = s.sphere_center.x;
  info->sphere[1] = s.sphere_center.y;
  info->sphere[2] = s.sphere_center.z;
  info->sphere[3] = s.sphere_radius;
  info->sphere_albedo[0] = s.surface_albedo.x;
  info->sphere_albedo[1] = s.surface_albedo.y;
  info->sphere_albedo[2] = s.surface_albedo.z;
  const float a = std::min(std::max(s.surface_alpha, 0.01f), 1.f);
  info->sphere_alpha = a * a;
  V3 cam[4];
  camera_vectors(s, cam[0], cam[1], cam[2], cam[3]);
  for (int k = 0; k < 4; ++k) {
    info->camera[3 * k] = cam[k].x;
    info->camera[3 * k + 1] = cam[k].y;
    info->camera[3 * k + 2] = cam[k].z;
  }
  for (int i = 0; i < (int)s.dlight_from.size() && i < 8; ++i) {
    info->dir_from[i][0] = s.dlight_from[i].x;
    info->dir_from[i][1] = s.dlight_from[i].y;
    info->dir_from[i][2] = s.dlight_from[i].z;
    info->dir_emit[i][0] = s.dlight_emit[i].x;
    info->dir_emit[i][1] = s.dlight_emit[i].y;
    info->dir_emit[i][2] = s.dlight_emit[i].z;
  }
  if (s.has_hair && s.hair_procedural) {
    info->num_strands = s.hair_gen.strands;
    info->num_segments = s.hair_gen.strands * s.hair_gen.segments;
  }
}

Context::~Context() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (Flight& f : extra) {
    if (f.stream) (void)hipStreamSynchronize(f.stream);
    for (void* p : f.allocs) dev_free(p);
    f.allocs.clear();
    if (f.stream) (void)hipStreamDestroy(f.stream);
    f.stream = nullptr;
  }
  if (acc_event) (void)hipEventDestroy(acc_event);
  for (hipEvent_t& e : live_ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (live_host) (void)hipHostFree(live_host);
  dev_free(stream_ring);
  dev_free(stream_ctl);
  msnn_release(msnn);
  msnn = nullptr;
  nrc_release(nrc);
  nrc = nullptr;
  free_path_buffers();
  dev_free(scene_points);
  dev_free(seg_dev);
  dev_free(crv_dev);
  dev_free(strand_sigma_dev);
  dev_free(stats_dev);
  dev_free(dlights_dev);
  bvh.release();
  mesh_bvh.release();
  dev_free(mesh_tri);
  dev_free(mesh_shade);
  env.release();
  if (stream) (void)hipStreamDestroy(stream);
}

void Context::free_path_buffers() {
  for (void* p : pb_allocs) dev_free(p);
  pb_allocs.clear();
  pb = PathBuffers{};
  dev_free(accum);
  dev_free(average);
  accum = average = nullptr;
}

// Allocates one set of path buffers for n slots (recording every allocation
// in allocs) and clears its counters on stream st.
int alloc_path_buffers(PathBuffers& pb, std::vector<void*>& allocs, size_t n, bool msnn_mode,
                              int spill_levels, int spill_stride, hipStream_t st, std::string& err) {
  pb = PathBuffers{};
  auto A = [&](void** p, size_t bytes) -> int {
    HM_TRY(dev_alloc(p, bytes, err));
    allocs.push_back(*p);
    return kOk;
  };
  const size_t f4 = sizeof(float4) * n;
  HM_TRY(A((void**)&pb.ray_o, f4));
  HM_TRY(A((void**)&pb.ray_d, f4));
  HM_TRY(A((void**)&pb.beta, f4));
  HM_TRY(A((void**)&pb.rng, sizeof(uint32_t) * n));
  HM_TRY(A((void**)&pb.L, f4));
  HM_TRY(A((void**)&pb.flags, sizeof(uint32_t) * n));
  HM_TRY(A((void**)&pb.hit_t, sizeof(float) * n));
  HM_TRY(A((void**)&pb.hit_prim, sizeof(int) * n));
  HM_TRY(A((void**)&pb.hit_u, sizeof(float) * n));
  HM_TRY(A((void**)&pb.sh_o1, f4));
  HM_TRY(A((void**)&pb.sh_d1, f4));
  HM_TRY(A((void**)&pb.sh_c1, f4));
  HM_TRY(A((void**)&pb.sh_o2, f4));
  HM_TRY(A((void**)&pb.sh_d2, f4));
  HM_TRY(A((void**)&pb.sh_c2, f4));
  HM_TRY(A((void**)&pb.sh_beta, f4));
  HM_TRY(A((void**)&pb.queue[0], sizeof(int) * n));
  HM_TRY(A((void**)&pb.queue[1], sizeof(int) * n));
  HM_TRY(A((void**)&pb.shadow_queue, sizeof(int) * n));
  HM_TRY(A((void**)&pb.shadow_rays, sizeof(int) * 2 * n));
  HM_TRY(A((void**)&pb.sh_vis, 2 * n));
  HM_TRY(A((void**)&pb.counts, sizeof(int) * kCountInts));
  HM_TRY(A((void**)&pb.slot_pass, sizeof(int) * n));
  HM_TRY(A((void**)&pb.slot_px, sizeof(int) * n));
  HM_TRY(A((void**)&pb.slot_depth, sizeof(int) * n));
  HM_TRY(A((void**)&pb.alt.ray_o, f4));
  HM_TRY(A((void**)&pb.alt.ray_d, f4));
  HM_TRY(A((void**)&pb.alt.beta, f4));
  HM_TRY(A((void**)&pb.alt.L, f4));
  HM_TRY(A((void**)&pb.alt.rng, sizeof(uint32_t) * n));
  HM_TRY(A((void**)&pb.alt.pass, sizeof(int) * n));
  HM_TRY(A((void**)&pb.alt.px, sizeof(int) * n));
  HM_TRY(A((void**)&pb.alt.depth, sizeof(int) * n));
  HM_TRY(A((void**)&pb.dest, sizeof(int) * n));
  HM_TRY(A((void**)&pb.done_L, f4));
  HM_TRY(A((void**)&pb.done_px, sizeof(int) * n));
  HM_TRY(A((void**)&pb.done_pass, sizeof(int) * n));
  HM_TRY(A((void**)&pb.job_ctl, sizeof(long long) * kJobCtl));
  if (msnn_mode) {
    HM_TRY(A((void**)&pb.beta_short, f4));
    HM_TRY(A((void**)&pb.L_short, f4));
    HM_TRY(A((void**)&pb.sh_beta_short, f4));
    HM_TRY(A((void**)&pb.alt.beta_short, f4));
    HM_TRY(A((void**)&pb.alt.L_short, f4));
    HM_TRY(A((void**)&pb.done_Ls, f4));
  }
  // traversal stack spill: levels beyond the LDS part, one column per thread
  pb.spill_stride = spill_stride;
  pb.spill_levels = spill_levels;
  HM_TRY(A((void**)&pb.spill, sizeof(int) * (size_t)spill_levels * spill_stride));
  HM_HIP(hipMemsetAsync(pb.counts, 0, sizeof(int) * kCountInts, st), err);
  return kOk;
}

int Context::ensure_path_buffers(bool msnn_mode, std::string& err, int min_slots) {
  const int want = std::max(scene.width * scene.height, min_slots);
  if (pb.ray_o && (!msnn_mode || pb_msnn) && pb_slots >= want) return kOk;
  free_path_buffers();
  const size_t n = (size_t)want;
  pb_slots = want;
  spill_levels = stack_levels();
  HM_TRY(alloc_path_buffers(pb, pb_allocs, n, msnn_mode, spill_levels, trace_blocks * kTraceBlock, stream, err));
  pb_msnn = msnn_mode;
  const size_t f4 = sizeof(float4) * n;
  HM_TRY(dev_alloc((void**)&accum, f4, err));
  HM_TRY(dev_alloc((void**)&average, f4, err));
  HM_HIP(hipMemsetAsync(accum, 0, f4, stream), err);
  HM_HIP(hipMemsetAsync(average, 0, f4, stream), err);
  accum_id = 0;
  return kOk;
}

// Path-tracing buffers for flights 1.. (flight 0 is pb).
int Context::ensure_flights(std::string& err) {
  if (!acc_event) HM_HIP(hipEventCreateWithFlags(&acc_event, hipEventDisableTiming), err);
  const int want = scene.width * scene.height;
  for (int f = 1; f < num_flights; ++f) {
    Flight& fl = extra[f - 1];
    if (!fl.stream) HM_HIP(hipStreamCreateWithFlags(&fl.stream, hipStreamNonBlocking), err);
    if (fl.slots >= want) continue;
    for (void* p : fl.allocs) dev_free(p);
    fl.allocs.clear();
    HM_TRY(alloc_path_buffers(fl.pb, fl.allocs, (size_t)want, false, stack_levels(),
                              trace_blocks * kTraceBlock, fl.stream, err));
    fl.slots = want;
  }
  return kOk;
}

int Context::stack_levels() const {
  return std::max(1, std::max(bvh.stack_bound(), mesh_bvh.stack_bound()) - kStackLds + 1);
}

PassConfig Context::pt_pass_config() const {
  PassConfig c;
  c.accum_id = accum_id;
  c.seed_offset = seed_offset;
  const int v1 = scene.path_v1 - 1, v2 = scene.path_v2 - 1;
  c.dl_zero = (v1 == 0);
  c.dl_from = v1;
  c.max_vertex = v2;
  if (v2 < v1) {  // pathTrace returns black (optix_common.cuh:365-366)
    c.dl_zero = 0;
    c.dl_from = 1 << 30;
  }
  c.trace_blocks = trace_blocks;
  c.shade_blocks = std::max(64, num_cus * 8);
  return c;
}

static int bvh_builder_id(const std::string& name) {
  return name == "ploc" ? kBvhPloc : name == "sah" ? kBvhSah : kBvhLbvh;
}

int Context::init(std::string& err) {
  HM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), err);
  int dev = 0;
  HM_HIP(hipGetDevice(&dev), err);
  hipDeviceProp_t prop;
  HM_HIP(hipGetDeviceProperties(&prop, dev), err);
  num_cus = prop.multiProcessorCount;
  // traversal blocks resident per CU: 8 (8 waves per SIMD), fewer when the
  // blocks' LDS stacks do not fit the CU's 160 KB together; the traversal
  // grid is that many blocks per CU unless HM_TRACE_GRID_PER_CU says more
  const int lds_blocks = (160 * 1024) / (kStackLds * kTraceBlock * (int)sizeof(int));
  const int resident = std::max(1, std::min(8, lds_blocks));
#ifdef HM_TRACE_GRID_PER_CU
  trace_blocks = std::max(64, num_cus * HM_TRACE_GRID_PER_CU);
#else
  trace_blocks = std::max(64, num_cus * resident);
#endif

  // ---- geometry ----
  if ((int)scene.dlight_from.size() > kMaxDirLights) {
    err = "at most " + std::to_string(kMaxDirLights) + " directional lights are supported";
    return kEUnsupported;
  }
  if (!scene.has_env && scene.dlight_from.empty()) {
    err = "the scene has no light: sampling would read an undefined environment";
    return kEInval;
  }
  bmin = v3(1e30f);
  bmax = v3(-1e30f);
  float sscale = 0.f;
  if (scene.has_surface && scene.surface_sphere) {
    const V3 c = scene.sphere_center;
    const float r = scene.sphere_radius;
    bmin = vmin(bmin, c - v3(r));
    bmax = vmax(bmax, c + v3(r));
    sscale = length(bmax - bmin);
  } else if (scene.has_surface) {
    // loadOBJ (model.cpp:233-358); the scene scale takes the mesh bounds like
    // the reference's surface scale (render_path_tracing.cu:212-214)
    HM_TRY(load_obj(scene_path(scene, scene.surface_geometry), mesh, err));
    bmin = vmin(bmin, mesh.bmin);
    bmax = vmax(bmax, mesh.bmax);
    sscale = length(mesh.bmax - mesh.bmin);
    const int nt = mesh.num();
    std::vector<float4> tri((size_t)3 * nt), shade((size_t)4 * nt);
    std::vector<float> boxes((size_t)8 * nt);  // the triangles' boxes as zero-radius segments for the BVH builder
    for (int k = 0; k < nt; ++k) {
      V3 lo = v3(1e30f), hi = v3(-1e30f);
      for (int j = 0; j < 3; ++j) {
        const float* p = &mesh.v[9 * (size_t)k + 3 * j];
        const float* q = &mesh.n[9 * (size_t)k + 3 * j];
        tri[3 * (size_t)k + j] = make_float4(p[0], p[1], p[2], 0.f);
        shade[4 * (size_t)k + j] = make_float4(q[0], q[1], q[2], 0.f);
        lo = vmin(lo, v3(p[0], p[1], p[2]));
        hi = vmax(hi, v3(p[0], p[1], p[2]));
      }
      shade[4 * (size_t)k + 3] = make_float4(mesh.kd[3 * k], mesh.kd[3 * k + 1], mesh.kd[3 * k + 2], 0.f);
      const float rec[8] = {lo.x, lo.y, lo.z, 0.f, hi.x, hi.y, hi.z, (float)k};
      std::copy(rec, rec + 8, &boxes[8 * (size_t)k]);
    }
    HM_TRY(dev_alloc((void**)&mesh_tri, sizeof(float4) * tri.size(), err));
    HM_TRY(dev_alloc((void**)&mesh_shade, sizeof(float4) * shade.size(), err));
    HM_HIP(hipMemcpy(mesh_tri, tri.data(), sizeof(float4) * tri.size(), hipMemcpyHostToDevice), err);
    HM_HIP(hipMemcpy(mesh_shade, shade.data(), sizeof(float4) * shade.size(), hipMemcpyHostToDevice), err);
    float* boxes_dev = nullptr;
    HM_TRY(dev_alloc((void**)&boxes_dev, sizeof(float) * boxes.size(), err));
    HM_HIP(hipMemcpy(boxes_dev, boxes.data(), sizeof(float) * boxes.size(), hipMemcpyHostToDevice), err);
    const int rc = bvh_build(boxes_dev, nt, mesh_bvh, stream, err, bvh_builder_id(scene.bvh_builder),
                             scene.ploc_radius, scene.bvh_max_leaf);
    (void)hipStreamSynchronize(stream);
    dev_free(boxes_dev);
    HM_TRY(rc);
  }
  if (scene.has_hair) {
    if (scene.hair_procedural) generate_hair(scene.hair_gen, hair);
    else
      HM_TRY(load_hair_file(scene_path(scene, scene.hair_geometry), scene.hair_tessellation, scene.hair_curves, hair,
                            err));
    bmin = vmin(bmin, hair.bmin);
    bmax = vmax(bmax, hair.bmax);
    sscale = std::max(sscale, length(hair.bmax - hair.bmin));  // render_path_tracing.cu:212-214
    const size_t bytes = sizeof(float) * hair.seg.size();
    HM_TRY(dev_alloc((void**)&seg_dev, bytes, err));
    HM_HIP(hipMemcpyAsync(seg_dev, hair.seg.data(), bytes, hipMemcpyHostToDevice, stream), err);
    const int builder = bvh_builder_id(scene.bvh_builder);
    if (hair.crv.empty()) {
      HM_TRY(bvh_build(seg_dev, (int)(hair.seg.size() / 8), bvh, stream, err, builder, scene.ploc_radius,
                       scene.bvh_max_leaf, scene.bvh_split));
    } else {
      // round curves: the BVH over each segment's box (the Bezier hull of
      // the Catmull-Rom segment, padded by the radius bound), stored as
      // zero-radius segment records for the builder; the curves themselves
      // as 5 float4 per segment
      const int nc = (int)(hair.crv.size() / kCurveFloats);
      std::vector<float> boxes((size_t)8 * nc);
      for (int k = 0; k < nc; ++k) {
        const float* q = &hair.crv[(size_t)kCurveFloats * k];
        const float rmax = curve_radius_bound(q[3], q[7], q[11], q[15]);
        float rec[8];
        for (int a = 0; a < 3; ++a) {
          const float b0 = q[4 + a], b3 = q[8 + a];
          const float b1 = b0 + (q[8 + a] - q[a]) * (1.f / 6.f), b2 = b3 - (q[12 + a] - q[4 + a]) * (1.f / 6.f);
          const float lo = std::min(std::min(b0, b1), std::min(b2, b3));
          const float hi = std::max(std::max(b0, b1), std::max(b2, b3));
          const float pad = 1.01f * rmax + 1e-5f * (std::fabs(lo) + std::fabs(hi)) + 1e-7f;
          rec[a] = lo - pad;
          rec[4 + a] = hi + pad;
        }
        rec[3] = 0.f;
        rec[7] = (float)k;
        std::copy(rec, rec + 8, &boxes[8 * (size_t)k]);
      }
      HM_TRY(dev_alloc((void**)&crv_dev, sizeof(float) * hair.crv.size(), err));
      HM_HIP(hipMemcpyAsync(crv_dev, hair.crv.data(), sizeof(float) * hair.crv.size(), hipMemcpyHostToDevice,
                            stream),
             err);
      float* boxes_dev = nullptr;
      HM_TRY(dev_alloc((void**)&boxes_dev, sizeof(float) * boxes.size(), err));
      HM_HIP(hipMemcpy(boxes_dev, boxes.data(), sizeof(float) * boxes.size(), hipMemcpyHostToDevice), err);
      const int rc = bvh_build(boxes_dev, nc, bvh, stream, err, builder, scene.ploc_radius, scene.bvh_max_leaf);
      (void)hipStreamSynchronize(stream);
      dev_free(boxes_dev);
      HM_TRY(rc);
    }
  }
  scene_scale = sscale > 0.f ? sscale : 1.f;

  // ---- environment ----
  if (scene.has_env) {
    if (scene.env_constant || scene.env_procedural) {
      generate_env(scene, env_host);
    } else {
      // loadEnvTexture (model.cpp:158-180)
      HM_TRY(read_exr(scene_path(scene, scene.env_exr), env_host.width, env_host.height, env_host.rgba, err));
    }
    env.width = env_host.width;
    env.height = env_host.height;
    const size_t bytes = sizeof(float) * env_host.rgba.size();
    HM_TRY(dev_alloc((void**)&env.tex, bytes, err));
    HM_HIP(hipMemcpyAsync(env.tex, env_host.rgba.data(), bytes, hipMemcpyHostToDevice, stream), err);
    HM_TRY(env_build_tables(env, stream, err));
  }

  camera_vectors(scene, cam_pos, cam_d00, cam_du, cam_dv);

  // ---- view ----
  SceneView& v = view;
  v = SceneView{};
  v.bvh.nodes = bvh.nodes_w;
  v.bvh.segs = bvh.segs;
  v.bvh.num_segs = bvh.num_segs;
  v.seg_by_id = seg_dev;
  v.curves = reinterpret_cast<const float4*>(crv_dev);
  v.curve_caps = 1;  // render_path_tracing's forceCaps (render_path_tracing.cu:223)
  v.strand_sigma = nullptr;
  v.sphere.enabled = scene.has_surface && scene.surface_sphere ? 1 : 0;
  v.mesh.bvh.nodes = mesh_bvh.nodes_w;
  v.mesh.bvh.segs = mesh_bvh.segs;
  v.mesh.bvh.num_segs = mesh_bvh.num_segs;
  v.mesh.tri = mesh_tri;
  v.mesh.shade = mesh_shade;
  v.mesh.num = mesh.num();
  v.sphere.cx = scene.sphere_center.x;
  v.sphere.cy = scene.sphere_center.y;
  v.sphere.cz = scene.sphere_center.z;
  v.sphere.r = scene.sphere_radius;
  v.sphere_albedo = scene.surface_albedo;
  const float a = std::min(std::max(scene.surface_alpha, 0.01f), 1.f);
  v.sphere_alpha = a * a;
  v.env.enabled = scene.has_env ? 1 : 0;
  v.env.pdf_sampling = scene.env_pdf ? 1 : 0;
  v.env.width = env.width;
  v.env.height = env.height;
  v.env.scale = scene.env_scale;
  v.env.rot_phi = scene.env_rot_phi;
  v.env.tex = env.tex;
  v.env.c_pdf = env.c_pdf;
  v.env.c_cdf = env.c_cdf;
  v.env.m_pdf = env.m_pdf;
  v.env.m_guide = env.m_guide;
  v.env.c_guide = env.c_guide;
  v.env.m_cdf = env.m_cdf;
  v.num_dlights = (int)scene.dlight_from.size();
  v.num_total_lights = v.num_dlights + (scene.has_env ? 1 : 0);
  {
    std::vector<float4> dl(2 * kMaxDirLights, make_float4(0.f, 0.f, 0.f, 0.f));
    for (int i = 0; i < v.num_dlights; ++i) {
      const V3 f = scene.dlight_from[i], e = scene.dlight_emit[i];
      dl[2 * i] = make_float4(f.x, f.y, f.z, 0.f);
      dl[2 * i + 1] = make_float4(e.x, e.y, e.z, 0.f);
    }
    HM_TRY(dev_alloc((void**)&dlights_dev, sizeof(float4) * dl.size(), err));
    HM_HIP(hipMemcpyAsync(dlights_dev, dl.data(), sizeof(float4) * dl.size(), hipMemcpyHostToDevice, stream), err);
    hipLaunchKernelGGL(k_normalize_dlights, dim3(1), dim3(kMaxDirLights), 0, stream, dlights_dev, v.num_dlights);
    HM_HIP(hipGetLastError(), err);
  }
  v.dlights = dlights_dev;
  v.cam_pos = cam_pos;
  v.cam_d00 = cam_d00;
  v.cam_du = cam_du;
  v.cam_dv = cam_dv;
  v.width = scene.width;
  v.height = scene.height;
  const float gains[4] = {scene.R_G, scene.TT_G, scene.TRT_G, scene.TRRT_G};
  v.hc = make_hair_consts(scene.beta_m, scene.beta_n, scene.alpha, gains);
  v.sigma_a = scene.sig_a;
  v.mis = scene.mis ? 1 : 0;
  HM_TRY(dev_alloc((void**)&stats_dev, sizeof(unsigned long long) * 8, err));
  HM_HIP(hipStreamSynchronize(stream), err);
  return kOk;
}

}  // namespace hm

using namespace hm;

struct hmsnn_ctx {
  Context c;
};

// ---------------------------------------------------------------------------
extern "C" {

const char* hmsnn_last_error(void) { return g_last_error.c_str(); }

int hmsnn_version(void) { return 10000; }

int hmsnn_device_count(int* count, char* name, int name_len) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    if (count) *count = 0;
    return fail(HMSNN_EDEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  if (count) *count = n;
  if (name && name_len > 0) {
    name[0] = 0;
    if (n > 0) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, 0) == hipSuccess) {
        std::snprintf(name, (size_t)name_len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
      }
    }
  }
  return HMSNN_OK;
}

int hmsnn_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(HMSNN_EINVAL, "ptr is NULL");
  std::string err;
  int rc = dev_alloc(ptr, bytes, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}
int hmsnn_dev_free(void* ptr) {
  if (ptr) {
    hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return fail(HMSNN_EDEVICE, std::string("hipFree: ") + hipGetErrorString(e));
  }
  return HMSNN_OK;
}
int hmsnn_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("hipMemcpy H2D: ") + hipGetErrorString(e));
}
int hmsnn_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("hipMemcpy D2H: ") + hipGetErrorString(e));
}
int hmsnn_dev_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
}

int hmsnn_hair_bsdf_eval_dev(const hmsnn_hair_params* params, const hmsnn_bsdf_query* q, int n, float* f,
                             float* pdf, void* stream) {
  if (!params || n < 0 || (n > 0 && (!q || !f || !pdf))) return fail(HMSNN_EINVAL, "bsdf eval: bad arguments");
  std::string err;
  int rc = bsdf_eval_launch(consts_from(params), q, n, f, pdf, (hipStream_t)stream, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_hair_bsdf_sample_dev(const hmsnn_hair_params* params, const hmsnn_bsdf_query* q, const float* u4, int n,
                               float* wi, float* f, float* pdf, int* lobe, void* stream) {
  if (!params || n < 0 || (n > 0 && (!q || !u4 || !wi || !f || !pdf)))
    return fail(HMSNN_EINVAL, "bsdf sample: bad arguments");
  std::string err;
  int rc = bsdf_sample_launch(consts_from(params), q, u4, n, wi, f, pdf, lobe, (hipStream_t)stream, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

static int create_common(Scene&& scene, hmsnn_ctx** out) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(HMSNN_EDEVICE, "no GPU visible");
  hmsnn_ctx* c = new hmsnn_ctx();
  c->c.scene = std::move(scene);
  std::string err;
  int rc = c->c.init(err);
  if (rc) {
    delete c;
    return fail(rc, err);
  }
  *out = c;
  return HMSNN_OK;
}

int hmsnn_create(const char* path, hmsnn_ctx** out) {
  if (!path || !out) return fail(HMSNN_EINVAL, "hmsnn_create: NULL argument");
  *out = nullptr;
  Scene s;
  std::string err;
  if (!parse_scene_file(path, s, err)) {
    const bool io = err.find("cannot open") != std::string::npos;
    return fail(io ? HMSNN_EIO : HMSNN_EINVAL, err);
  }
  return create_common(std::move(s), out);
}

int hmsnn_create_from_json(const char* text, const char* base_dir, hmsnn_ctx** out) {
  if (!text || !out) return fail(HMSNN_EINVAL, "hmsnn_create_from_json: NULL argument");
  *out = nullptr;
  Scene s;
  std::string err;
  if (!parse_scene_text(text, base_dir ? base_dir : ".", s, err)) return fail(HMSNN_EINVAL, err);
  return create_common(std::move(s), out);
}

int hmsnn_destroy(hmsnn_ctx* ctx) {
  delete ctx;
  return HMSNN_OK;
}

int hmsnn_get_info(const hmsnn_ctx* ctx, hmsnn_info* info) {
  if (!ctx || !info) return fail(HMSNN_EINVAL, "hmsnn_get_info: NULL argument");
  const Context& c = ctx->c;
  std::memset(info, 0, sizeof(*info));
  fill_scene_info(c.scene, info);
  info->num_segments = c.bvh.num_prims;
  info->num_curves = (int)(c.hair.crv.size() / kCurveFloats);
  info->num_strands = c.hair.num_strands;
  info->num_triangles = c.mesh.num();
  info->bvh_nodes = c.bvh.num_nodes_w;
  info->bvh_depth = c.bvh.depth_w;
  info->env_width = c.env.width;
  info->env_height = c.env.height;
  info->scene_scale = c.scene_scale;
  return HMSNN_OK;
}

int hmsnn_get_string(const hmsnn_ctx* ctx, int which, char* buf, int len) {
  if (!ctx || !buf || len <= 0 || which < 0 || which > 2) return fail(HMSNN_EINVAL, "hmsnn_get_string: bad arguments");
  const Scene& s = ctx->c.scene;
  std::string v = which == 0 ? s.image_output : which == 1 ? s.stats_output : s.json_path;
  if (which < 2 && !v.empty() && v[0] != '/' && !(v.size() > 1 && v[1] == ':')) v = s.base_dir + "/" + v;
  if ((int)v.size() >= len) return fail(HMSNN_EINVAL, "hmsnn_get_string: buffer too small");
  std::memcpy(buf, v.c_str(), v.size() + 1);
  return HMSNN_OK;
}

int hmsnn_parse_scene(const char* text, hmsnn_info* info) {
  if (!text || !info) return fail(HMSNN_EINVAL, "hmsnn_parse_scene: NULL argument");
  Scene s;
  std::string err;
  if (!parse_scene_text(text, ".", s, err)) return fail(HMSNN_EINVAL, err);
  std::memset(info, 0, sizeof(*info));
  fill_scene_info(s, info);
  return HMSNN_OK;
}

int hmsnn_generate_host(const char* text, float* seg_host, size_t seg_cap, int* num_segments, float* env_rgba,
                        size_t env_cap, int* env_w, int* env_h) {
  if (!text) return fail(HMSNN_EINVAL, "hmsnn_generate_host: NULL json");
  Scene s;
  std::string err;
  if (!parse_scene_text(text, ".", s, err)) return fail(HMSNN_EINVAL, err);
  HairGeometry hg;
  if (s.has_hair && s.hair_procedural) generate_hair(s.hair_gen, hg);
  else if (s.has_hair) {
    const int rc = load_hair_file(scene_path(s, s.hair_geometry), s.hair_tessellation, s.hair_curves, hg, err);
    if (rc) return fail(rc, err);
  }
  const size_t nseg = hg.seg.size() / 8;
  if (num_segments) *num_segments = (int)nseg;
  if (seg_host) {
    if (seg_cap < hg.seg.size()) return fail(HMSNN_EINVAL, "segment buffer too small");
    std::memcpy(seg_host, hg.seg.data(), sizeof(float) * hg.seg.size());
  }
  EnvImage env;
  if (s.has_env && (s.env_constant || s.env_procedural)) generate_env(s, env);
  else if (s.has_env && !s.env_exr.empty()) {
    const int rc = read_exr(scene_path(s, s.env_exr), env.width, env.height, env.rgba, err);
    if (rc) return fail(rc, err);
  }
  if (env_w) *env_w = env.width;
  if (env_h) *env_h = env.height;
  if (env_rgba) {
    if (env_cap < env.rgba.size()) return fail(HMSNN_EINVAL, "environment buffer too small");
    std::memcpy(env_rgba, env.rgba.data(), sizeof(float) * env.rgba.size());
  }
  return HMSNN_OK;
}

int hmsnn_generate_curves_host(const char* text, float* crv_host, size_t crv_cap, int* num_curves) {
  if (!text) return fail(HMSNN_EINVAL, "hmsnn_generate_curves_host: NULL json");
  Scene s;
  std::string err;
  if (!parse_scene_text(text, ".", s, err)) return fail(HMSNN_EINVAL, err);
  HairGeometry hg;
  if (s.has_hair && !s.hair_procedural) {
    const int rc = load_hair_file(scene_path(s, s.hair_geometry), s.hair_tessellation, s.hair_curves, hg, err);
    if (rc) return fail(rc, err);
  }
  if (num_curves) *num_curves = (int)(hg.crv.size() / kCurveFloats);
  if (crv_host) {
    if (crv_cap < hg.crv.size()) return fail(HMSNN_EINVAL, "curve buffer too small");
    std::memcpy(crv_host, hg.crv.data(), sizeof(float) * hg.crv.size());
  }
  return HMSNN_OK;
}

int hmsnn_export_curves(const hmsnn_ctx* ctx, float* crv_host) {
  if (!ctx || !crv_host) return fail(HMSNN_EINVAL, "hmsnn_export_curves: NULL argument");
  std::memcpy(crv_host, ctx->c.hair.crv.data(), sizeof(float) * ctx->c.hair.crv.size());
  return HMSNN_OK;
}

int hmsnn_export_segments(const hmsnn_ctx* ctx, float* seg_host) {
  if (!ctx || !seg_host) return fail(HMSNN_EINVAL, "hmsnn_export_segments: NULL argument");
  std::memcpy(seg_host, ctx->c.hair.seg.data(), sizeof(float) * ctx->c.hair.seg.size());
  return HMSNN_OK;
}

int hmsnn_export_mesh(const hmsnn_ctx* ctx, float* vertices, float* normals, float* diffuse) {
  if (!ctx) return fail(HMSNN_EINVAL, "hmsnn_export_mesh: NULL context");
  const TriMesh& m = ctx->c.mesh;
  if (vertices) std::memcpy(vertices, m.v.data(), sizeof(float) * m.v.size());
  if (normals) std::memcpy(normals, m.n.data(), sizeof(float) * m.n.size());
  if (diffuse) std::memcpy(diffuse, m.kd.data(), sizeof(float) * m.kd.size());
  return HMSNN_OK;
}

int hmsnn_export_env(const hmsnn_ctx* ctx, float* rgba, float* c_pdf, float* c_cdf, float* m_pdf, float* m_cdf) {
  if (!ctx) return fail(HMSNN_EINVAL, "hmsnn_export_env: NULL context");
  const Context& c = ctx->c;
  if (!c.scene.has_env) return fail(HMSNN_EINVAL, "scene has no environment light");
  const size_t W = c.env.width, H = c.env.height;
  auto cp = [&](float* dst, const float* src, size_t n) -> int {
    if (!dst) return HMSNN_OK;
    hipError_t e = hipMemcpy(dst, src, sizeof(float) * n, hipMemcpyDeviceToHost);
    return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("export env: ") + hipGetErrorString(e));
  };
  HM_TRY(cp(rgba, (const float*)c.env.tex, W * H * 4));
  HM_TRY(cp(c_pdf, c.env.c_pdf, (W + 1) * H));
  HM_TRY(cp(c_cdf, c.env.c_cdf, (W + 1) * H));
  HM_TRY(cp(m_pdf, c.env.m_pdf, H + 1));
  HM_TRY(cp(m_cdf, c.env.m_cdf, H + 1));
  return HMSNN_OK;
}

int hmsnn_set_strand_sigma_a(hmsnn_ctx* ctx, const float* sigma, int num_strands) {
  if (!ctx) return fail(HMSNN_EINVAL, "hmsnn_set_strand_sigma_a: NULL context");
  Context& c = ctx->c;
  if (!sigma) {
    c.view.strand_sigma = nullptr;
    return HMSNN_OK;
  }
  if (num_strands != c.hair.num_strands)
    return fail(HMSNN_EINVAL, "sigma_a array must hold num_strands x 3 floats");
  std::string err;
  if (!c.strand_sigma_dev) {
    int rc = dev_alloc((void**)&c.strand_sigma_dev, sizeof(float) * 3 * (size_t)num_strands, err);
    if (rc) return fail(rc, err);
  }
  hipError_t e = hipMemcpyAsync(c.strand_sigma_dev, sigma, sizeof(float) * 3 * (size_t)num_strands,
                                hipMemcpyHostToDevice, c.stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  (void)hipStreamSynchronize(c.stream);
  c.view.strand_sigma = c.strand_sigma_dev;
  return HMSNN_OK;
}

int hmsnn_set_pass_concurrency(hmsnn_ctx* ctx, int passes_in_flight) {
  if (!ctx || passes_in_flight < 1 || passes_in_flight > Context::kMaxFlights)
    return fail(HMSNN_EINVAL, "hmsnn_set_pass_concurrency: expects 1..4");
  Context& c = ctx->c;
  (void)hipStreamSynchronize(c.stream);  // every flight's work is behind the context stream
  c.num_flights = passes_in_flight;
  return HMSNN_OK;
}

int hmsnn_set_pt_schedule(hmsnn_ctx* ctx, int schedule) {
  if (!ctx || (schedule != HMSNN_PT_STREAM && schedule != HMSNN_PT_PASSES))
    return fail(HMSNN_EINVAL, "hmsnn_set_pt_schedule: expects HMSNN_PT_STREAM or HMSNN_PT_PASSES");
  Context& c = ctx->c;
  (void)hipStreamSynchronize(c.stream);
  c.pt_schedule = schedule;
  return HMSNN_OK;
}

int hmsnn_set_seed_offset(hmsnn_ctx* ctx, int offset) {
  if (!ctx || offset < 0) return fail(HMSNN_EINVAL, "hmsnn_set_seed_offset: bad arguments");
  ctx->c.seed_offset = offset;
  return HMSNN_OK;
}

int hmsnn_pt_reset(hmsnn_ctx* ctx) {
  if (!ctx) return fail(HMSNN_EINVAL, "NULL context");
  ctx->c.accum_id = 0;
  return HMSNN_OK;
}

}  // extern "C"

namespace hm {

// Streaming schedule: passes first_pass .. first_pass + passes - 1 of every
// pixel accumulated into acc / avg (the context's image, or scratch buffers of
// a caller that only wants the radiance ring); the context stream waits for
// every group. The ring then holds the last (up to) 16 passes' radiance.
int stream_passes(Context& c, int first_pass, int passes, float4* acc, float4* avg, PassTimer* timer,
                  std::string& err) {
  HM_TRY(c.ensure_path_buffers(false, err));
  HM_TRY(c.ensure_flights(err));
  const int n = c.scene.width * c.scene.height;
  const int K = std::max(1, std::min(c.num_flights, Context::kMaxFlights));
  if (passes <= 0) return kOk;
  if (!c.live_host) {
    HM_HIP(hipHostMalloc((void**)&c.live_host, 2 * Context::kMaxFlights * sizeof(int)), err);
    for (hipEvent_t& e : c.live_ev) HM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), err);
  }
  if (!c.stream_ring) {
    HM_TRY(dev_alloc((void**)&c.stream_ring, sizeof(float4) * kStreamRing * (size_t)n, err));
    HM_TRY(dev_alloc((void**)&c.stream_ctl, sizeof(long long) * kStreamCtl * Context::kMaxFlights, err));
  }
  // every group's stream starts after the work queued so far (including
  // the other schedule's last accumulation)
  if (c.acc_pending) HM_HIP(hipStreamWaitEvent(c.stream, c.acc_event, 0), err);
  c.acc_pending = false;
  HM_HIP(hipEventRecord(c.acc_event, c.stream), err);
  PathBuffers* pbs[Context::kMaxFlights];
  hipStream_t sts[Context::kMaxFlights];
  for (int g = 0; g < K; ++g) {
    pbs[g] = &c.flight_pb(g);
    sts[g] = c.flight_stream(g);
    if (g > 0) HM_HIP(hipStreamWaitEvent(sts[g], c.acc_event, 0), err);
  }
  PassConfig cfg = c.pt_pass_config();
  cfg.accum_id = first_pass;
  cfg.timer = timer;
  HM_TRY(pt_run_stream(c.view, pbs, sts, K, cfg, passes, acc, avg, c.stream_ring, c.stream_ctl, c.live_host,
                       c.live_ev, err));
  // the context stream waits for every group
  for (int g = 1; g < K; ++g) {
    HM_HIP(hipEventRecord(c.acc_event, sts[g]), err);
    HM_HIP(hipStreamWaitEvent(c.stream, c.acc_event, 0), err);
  }
  return kOk;
}

}  // namespace hm

extern "C" {

// Runs passes round-robin over the flights. Every flight first waits for
// the work already queued on the context stream; accumulations are chained
// through acc_event so they happen in pass order (the image is the one a
// single stream makes), and the context stream waits for the last one.
static int run_pt_passes(Context& c, int passes, std::string& err, PassTimer* timer = nullptr) {
  HM_TRY(c.ensure_path_buffers(false, err));
  SceneView v = c.view;
  v.bvh.nodes = c.bvh.nodes_w;
  const int n = c.scene.width * c.scene.height;
  HM_TRY(c.ensure_flights(err));
  const int K = std::max(1, std::min(c.num_flights, Context::kMaxFlights));
  if (c.pt_schedule == 0) {
    if (passes <= 0) return kOk;
    HM_TRY(stream_passes(c, c.accum_id, passes, c.accum, c.average, timer, err));
    c.last_ring_entry = (passes - 1) % kStreamRing;
    c.accum_id += passes;
    c.last_streamed = true;
    c.stream_groups = K;
    return kOk;
  }
  c.last_streamed = false;
  hipEvent_t start = nullptr;
  HM_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming), err);
  HM_HIP(hipEventRecord(start, c.stream), err);
  for (int f = 1; f < K; ++f) HM_HIP(hipStreamWaitEvent(c.flight_stream(f), start, 0), err);
  int rc = kOk;
  for (int k = 0; k < passes && rc == kOk; ++k) {
    const int f = c.accum_id % K;
    PathBuffers& pb = c.flight_pb(f);
    hipStream_t st = c.flight_stream(f);
    PassConfig cfg = c.pt_pass_config();
    cfg.timer = timer;
    rc = pt_run_pass(v, pb, cfg, st, err);
    if (rc) break;
    if (c.acc_pending && hipStreamWaitEvent(st, c.acc_event, 0) != hipSuccess) {
      rc = kEDevice;
      err = "stream wait failed";
      break;
    }
    rc = pt_accumulate(pb.L, c.accum, c.average, n, c.accum_id, st, err);
    if (rc) break;
    (void)hipEventRecord(c.acc_event, st);
    c.acc_pending = true;
    c.accum_id++;
  }
  if (c.acc_pending) (void)hipStreamWaitEvent(c.stream, c.acc_event, 0);
  (void)hipEventDestroy(start);
  return rc;
}

int hmsnn_pt_passes(hmsnn_ctx* ctx, int passes) {
  if (!ctx || passes < 0) return fail(HMSNN_EINVAL, "hmsnn_pt_passes: bad arguments");
  std::string err;
  int rc = run_pt_passes(ctx->c, passes, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_ctx_synchronize(hmsnn_ctx* ctx) {
  if (!ctx) return fail(HMSNN_EINVAL, "NULL context");
  hipError_t e = hipStreamSynchronize(ctx->c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("stream sync: ") + hipGetErrorString(e));
}

int hmsnn_pt_accum_count(const hmsnn_ctx* ctx, int* count) {
  if (!ctx || !count) return fail(HMSNN_EINVAL, "NULL argument");
  *count = ctx->c.accum_id;
  return HMSNN_OK;
}

int hmsnn_get_image(hmsnn_ctx* ctx, int which, float* rgba) {
  if (!ctx || !rgba || which < 0 || which > 1) return fail(HMSNN_EINVAL, "hmsnn_get_image: bad arguments");
  Context& c = ctx->c;
  if (!c.accum) return fail(HMSNN_EINVAL, "nothing rendered yet");
  const size_t bytes = sizeof(float4) * (size_t)c.scene.width * c.scene.height;
  hipError_t e = hipMemcpyAsync(rgba, which == 0 ? c.average : c.accum, bytes, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("get image: ") + hipGetErrorString(e));
}

int hmsnn_get_last_pass(hmsnn_ctx* ctx, float* rgba) {
  if (!ctx || !rgba) return fail(HMSNN_EINVAL, "hmsnn_get_last_pass: bad arguments");
  Context& c = ctx->c;
  if (!c.pb.L || c.accum_id == 0) return fail(HMSNN_EINVAL, "nothing rendered yet");
  const int npx = c.scene.width * c.scene.height;
  hipError_t e = hipSuccess;
  if (c.last_streamed) {
    // the ring entry of the last pass holds every pixel's radiance
    e = hipMemcpyAsync(rgba, c.stream_ring + (size_t)c.last_ring_entry * npx, sizeof(float4) * (size_t)npx,
                       hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  } else {
    const int K = std::max(1, std::min(c.num_flights, Context::kMaxFlights));
    const PathBuffers& last = c.flight_pb((c.accum_id - 1) % K);
    e = hipMemcpyAsync(rgba, last.L, sizeof(float4) * (size_t)npx, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  }
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("get last pass: ") + hipGetErrorString(e));
}

int hmsnn_render_path_tracing(hmsnn_ctx* ctx, int spp, float* rgba) {
  if (!ctx || spp <= 0) return fail(HMSNN_EINVAL, "hmsnn_render_path_tracing: bad arguments");
  ctx->c.accum_id = 0;
  std::string err;
  int rc = run_pt_passes(ctx->c, spp, err);
  if (rc) return fail(rc, err);
  if (rgba) return hmsnn_get_image(ctx, 0, rgba);
  return hmsnn_ctx_synchronize(ctx);
}

int hmsnn_pt_trace_stats(hmsnn_ctx* ctx, int passes, double out[6]) {
  if (!ctx || passes <= 0 || !out) return fail(HMSNN_EINVAL, "hmsnn_pt_trace_stats: bad arguments");
  Context& c = ctx->c;
  std::string err;
  int rc = c.ensure_path_buffers(false, err);
  if (rc) return fail(rc, err);
  hipError_t e = hipMemsetAsync(c.stats_dev, 0, sizeof(unsigned long long) * 8, c.stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  for (int p = 0; p < passes; ++p) {
    PassConfig cfg = c.pt_pass_config();
    cfg.accum_id = c.accum_id + p;
    cfg.stats = c.stats_dev;
    rc = pt_run_pass(c.view, c.pb, cfg, c.stream, err);
    if (rc) return fail(rc, err);
  }
  unsigned long long h[8];
  e = hipMemcpyAsync(h, c.stats_dev, sizeof(h), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  out[0] = (double)h[4];
  out[1] = (double)h[0];
  out[2] = (double)h[1];
  out[3] = (double)h[5];
  out[4] = (double)h[2];
  out[5] = (double)h[3];
  return HMSNN_OK;
}

int hmsnn_pt_time_passes(hmsnn_ctx* ctx, int passes, double out[5]) {
  if (!ctx || passes <= 0 || !out) return fail(HMSNN_EINVAL, "hmsnn_pt_time_passes: bad arguments");
  Context& c = ctx->c;
  std::string err;
  int rc = c.ensure_path_buffers(false, err);
  if (rc) return fail(rc, err);
  // records per vertex iteration: extend, shade, shadow (+ finish when streaming)
  const int per_pass = 4 * (c.scene.path_v2 + 1);
  PassTimer tm;
  tm.capacity = per_pass * passes + 4 * 64;
  std::vector<hipEvent_t> ev(2 * (size_t)tm.capacity);
  std::vector<int> kinds((size_t)tm.capacity);
  for (auto& x : ev) (void)hipEventCreate(&x);
  tm.events = ev.data();
  tm.kinds = kinds.data();
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  (void)hipEventRecord(t0, c.stream);
  rc = run_pt_passes(c, passes, err, &tm);
  (void)hipEventRecord(t1, c.stream);
  hipError_t e = hipEventSynchronize(t1);
  double sums[3] = {0, 0, 0};
  int launches = 0;
  if (rc == 0 && e == hipSuccess) {
    for (int k = 0; k < tm.used; ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
      if (kinds[k] >= 0 && kinds[k] < 3) sums[kinds[k]] += ms;
      if (kinds[k] == 0) launches++;
    }
    float tot = 0.f;
    (void)hipEventElapsedTime(&tot, t0, t1);
    out[0] = tot;
    out[1] = sums[0];
    out[2] = sums[1];
    out[3] = sums[2];
    out[4] = launches;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) return fail(rc, err);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  return HMSNN_OK;
}

int hmsnn_trace_rays_dev(hmsnn_ctx* ctx, const float* rays, int n, int any_hit, int* prim, float* t, float* u) {
  if (!ctx || n < 0 || (n > 0 && (!rays || !prim))) return fail(HMSNN_EINVAL, "hmsnn_trace_rays_dev: bad arguments");
  Context& c = ctx->c;
  std::string err;
  int rc = c.ensure_path_buffers(false, err);
  if (rc) return fail(rc, err);
  rc = trace_rays(c.view, rays, n, any_hit, prim, t, u, c.pb.spill, c.pb.spill_stride, c.pb.spill_levels,
                  c.trace_blocks, nullptr, c.stream, err);
  if (rc) return fail(rc, err);
  hipError_t e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, std::string("trace rays: ") + hipGetErrorString(e));
}

}  // extern "C"
