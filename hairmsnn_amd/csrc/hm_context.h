// Scene context: owns everything a render needs on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "hm_internal.h"
#include "hm_pt_kernels.h"
#include "hm_scene.h"

namespace hm {

struct MlpState;  // radiance-cache network (hm_mlp.hip)
struct MsnnState;  // HairMSNN frame state (hm_msnn.hip)
void msnn_release(MsnnState* m);
struct NrcState;  // neural radiance caching frame state (hm_nrc.hip)
void nrc_release(NrcState* s);

// An extra set of path buffers with its own stream, so that independent
// passes overlap on the GPU (the last depths of a pass launch kernels whose
// duration is set by their slowest ray, not by their ray count).
struct Flight {
  hipStream_t stream = nullptr;
  PathBuffers pb{};
  std::vector<void*> allocs;
  int slots = 0;
};

// One set of path buffers for n slots (msnn_mode adds the training-path
// arrays); every allocation is appended to allocs, counters cleared on st.
int alloc_path_buffers(PathBuffers& pb, std::vector<void*>& allocs, size_t n, bool msnn_mode, int spill_levels,
                       int spill_stride, hipStream_t st, std::string& err);

struct Context {
  Scene scene;
  hipStream_t stream = nullptr;
  int num_cus = 256;

  // geometry
  HairGeometry hair;
  float* seg_dev = nullptr;  // 8 floats per segment, original order
  float* crv_dev = nullptr;  // round curves: kCurveFloats per segment
  float* strand_sigma_dev = nullptr;
  DeviceBvh bvh;
  // OBJ head mesh (surface.geometry)
  TriMesh mesh;
  float4* mesh_tri = nullptr;    // 3 per triangle
  float4* mesh_shade = nullptr;  // 4 per triangle
  DeviceBvh mesh_bvh;
  float scene_scale = 1.f;
  V3 bmin{0, 0, 0}, bmax{0, 0, 0};

  // environment
  DeviceEnv env;
  EnvImage env_host;

  // camera (render_path_tracing.cu:767-799)
  V3 cam_pos, cam_d00, cam_du, cam_dv;

  SceneView view{};
  // The view HairMSNN and NRC trace with: their round curves have no end
  // caps (owlCurvesSetDegree(…, 3, false), render_hair_msnn.cu:346,
  // render_nrc.cu:409), plain path tracing's have (render_path_tracing.cu:223).
  SceneView view_no_caps() const {
    SceneView v = view;
    v.curve_caps = 0;
    return v;
  }
  float4* dlights_dev = nullptr;  // SceneView::dlights

  // path-tracer state
  PathBuffers pb{};
  std::vector<void*> pb_allocs;
  bool pb_msnn = false;
  float4* accum = nullptr;
  float4* average = nullptr;
  int accum_id = 0;
  int seed_offset = 0;  // hmsnn_set_seed_offset
  int trace_blocks = 2048;
  int spill_levels = 0;
  unsigned long long* stats_dev = nullptr;

  MsnnState* msnn = nullptr;
  NrcState* nrc = nullptr;
  // path-tracing passes in flight at once: flight 0 is (pb, stream)
  static constexpr int kMaxFlights = 4;
  int num_flights = 4;
  Flight extra[kMaxFlights - 1];
  hipEvent_t acc_event = nullptr;  // last accumulate issued (any flight)
  bool acc_pending = false;
  // plain path tracing schedule: 0 streaming (path regeneration, one set of
  // path buffers), 1 one wavefront per pass with passes in flight
  int pt_schedule = 0;
  bool last_streamed = false;   // the last pt call used the streaming schedule
  int stream_groups = 1;        // groups of that call
  int* live_host = nullptr;     // pinned, 2 per group: live-path count polls
  float4* stream_ring = nullptr;  // streaming schedule: kStreamRing passes of radiance per pixel
  long long* stream_ctl = nullptr;  // streaming schedule: scheduler counters per group
  int last_ring_entry = 0;          // ring entry of the last streamed pass
  hipEvent_t live_ev[2 * kMaxFlights] = {};
  int ensure_flights(std::string& err);
  PathBuffers& flight_pb(int f) { return f == 0 ? pb : extra[f - 1].pb; }
  hipStream_t flight_stream(int f) { return f == 0 ? stream : extra[f - 1].stream; }
  // HairMSNN scene samples on the hair (3 floats each), made once per scene
  float* scene_points = nullptr;
  int num_scene_points = 0;

  ~Context();
  int init(std::string& err);
  // Path state for max(width * height, min_slots) slots; msnn_mode adds the
  // training-path arrays.
  int ensure_path_buffers(bool msnn_mode, std::string& err, int min_slots = 0);
  int pb_slots = 0;
  void free_path_buffers();
  PassConfig pt_pass_config() const;
  // global stack-spill levels beyond the LDS part for the deepest tree traversed
  int stack_levels() const;
};

// Streaming-schedule passes first_pass .. first_pass + passes - 1 into acc /
// avg; afterwards the context stream is ordered after them and c.stream_ring
// holds the radiance of the last (up to) kStreamRing passes by pass - first_pass.
int stream_passes(Context& c, int first_pass, int passes, float4* acc, float4* avg, PassTimer* timer,
                  std::string& err);

}  // namespace hm
