// Host-callable launchers of the wavefront kernels and the G-buffer writers
// used by the HairMSNN G_BUFFER pass.
#pragma once

#include <algorithm>
#include <string>

#include "hm_internal.h"
#include "hm_pt.h"

namespace hm {

// Per-pixel outputs of the HairMSNN G_BUFFER pass at the primary vertex
// (hair_msnn.cu:290-311): the 12-float network input record and hit flags.
// Record layout: position / sceneScale (3), wo (3), tangent (3), sigma_a (3).
// The reference writes channels 0-8 and leaves 9-11 uninitialised; here they
// carry the hit's absorption coefficient.
struct GBufferOut {
  int enabled;
  float* nn_input;     // 12 floats per pixel
  uint8_t* hit_flags;  // bit 0: hit, bit 1: surface
  float scene_scale;
};

constexpr int kNnInputCh = 12;

#if defined(__HIPCC__)
HM_D void gbuffer_write_hit(const GBufferOut& gb, const SceneView& sc, int slot, const Hit& h, int prim) {
  float* r = gb.nn_input + (size_t)slot * kNnInputCh;
  const V3 p = h.p / gb.scene_scale;
  r[0] = p.x; r[1] = p.y; r[2] = p.z;
  r[3] = h.wo.x; r[4] = h.wo.y; r[5] = h.wo.z;
  r[6] = h.fr.r0.x; r[7] = h.fr.r0.y; r[8] = h.fr.r0.z;
  V3 sig = sc.sigma_a;
  if (!h.surface && sc.strand_sigma) {
    const int strand = (int)sc.seg_by_id[8 * (size_t)prim + 7];
    sig = v3(sc.strand_sigma[3 * strand], sc.strand_sigma[3 * strand + 1], sc.strand_sigma[3 * strand + 2]);
  }
  r[9] = sig.x; r[10] = sig.y; r[11] = sig.z;
  gb.hit_flags[slot] = (uint8_t)(1u | (h.surface ? 2u : 0u));
}

HM_D void gbuffer_write_miss(const GBufferOut& gb, int slot, V3 dir) {
  float* r = gb.nn_input + (size_t)slot * kNnInputCh;
  r[0] = 0.f; r[1] = 0.f; r[2] = 0.f;
  r[3] = -dir.x; r[4] = -dir.y; r[5] = -dir.z;
  for (int k = 6; k < kNnInputCh; ++k) r[k] = 0.f;
  gb.hit_flags[slot] = 0;
}
#endif

// Neural radiance caching G_BUFFER pass (cuda/nrc.cu:135-398): paths end
// where their spread reaches the camera footprint (NRC sect. 3.4) and query
// the cache there; training pixels keep tracing a suffix whose vertices
// become training records. Slots are pixels.
constexpr int kNrcMaxBounces = 40;  // MAX_BOUNCES (nrc.cuh:13)
constexpr uint32_t kNrcTrain = 1u, kNrcUnbiased = 2u, kNrcSuffix = 4u, kNrcRecalcA0 = 8u;

struct NrcPassArgs {
  int enabled;
  int all_unbiased;
  float c;             // spread threshold multiplier
  float scene_scale;
  int frame_size;      // pixels; suffix cache queries follow at frame_size + training index
  int n_tpx;           // training pixels
  int every_nth;       // training pixel of pixel p: index p / every_nth
  float* query;        // cache-query records, 12 floats each
  float4* gb_beta;     // per pixel: throughput the cache value is multiplied by (0: no query)
  uint8_t* gb_hit;     // per pixel: camera ray hit a front-facing surface or hair
  float4* state;       // per pixel: a0, spread, c, pdf of the last sampled direction
  float4* prev;        // per pixel: previous vertex
  // per training pixel x kNrcMaxBounces vertices
  float4* tb_pos;      // vertex / scene scale
  float4* tb_wo;
  float4* tb_n;
  float4* tb_rad;      // direct light at the vertex
  float4* tb_beta;     // f * cos / pdf of the sampled continuation
  int* tb_bounces;
  uint8_t* tb_hit;     // the suffix ended in a cache query
};

// Optional per-kernel timing: event pairs recorded around each launch.
struct PassTimer {
  hipEvent_t* events = nullptr;  // 2 per record
  int* kinds = nullptr;          // 0 extend, 1 shade, 2 shadow, 3 other
  int capacity = 0;              // records
  int used = 0;
};

struct PassConfig {
  int accum_id = 0;
  int seed_offset = 0;  // added to accum_id for the per-pixel seeds only
  int dl_zero = 1;        // direct lighting at the camera vertex
  int dl_from = 0;        // direct lighting at vertices >= dl_from (>= 1)
  int max_vertex = 39;    // path_v2 - 1
  int trunc_beta = 1 << 30;
  int msnn = 0;
  const int* train_idx = nullptr;
  int every_nth = 1;
  GBufferOut gb{};
  // non-null: counting traversal. [0] closest-hit nodes, [1] closest-hit
  // segments, [2] any-hit nodes, [3] any-hit segments, [4] closest-hit rays,
  // [5] any-hit rays
  unsigned long long* stats = nullptr;
  PassTimer* timer = nullptr;
  int chunk = 0;  // camera_mode 1: paths per pass index (0 = all slots one pass)
  int trace_blocks = 2048;
  int shade_blocks = 2048;
  // camera_mode 0: one jittered camera ray per pixel (cuda/path_tracing.cu:32-42)
  // camera_mode 1: rays from the camera to sampled hair points, all slots
  //                training paths (TRAIN_DATA_GEN, cuda/hair_msnn.cu:222-233)
  int camera_mode = 0;
  int n_slots = 0;             // camera_mode 1: number of records
  int grid_w = 0;              // camera_mode 1: launch width used for seeding
  const float* points = nullptr;  // camera_mode 1: 3 floats per record
  // camera_mode 2: the training pixels of frames accum_id.. (chunk records
  // each, train_idx per record); their pixel indices are written to pix
  int* pix = nullptr;
  NrcPassArgs nrc{};
};

// Runs one full sample pass (camera -> all vertices). Leaves per-pixel
// radiance in pb.L (and pb.L_short for training pixels).
int pt_run_pass(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, hipStream_t stream,
                std::string& err);

// Job schedule: the records of a camera_mode 1 or 2 pass (cfg.n_slots
// HairMSNN training paths) traced on a compacted wavefront that is refilled
// with new records, in record order, up to `capacity` paths in flight. The
// caller launches iterations (pt_jobs_iterate) and learns from
// pt_jobs_status whether the wavefront has drained; pt_jobs_drain does both
// until it has. Each path's operations are those of pt_run_pass, so the
// records are the same bits; out_L / out_Ls receive each record's full and
// short radiance, cfg.gb its network input at the primary hit.
struct JobRun {
  int capacity = 0;
  int cur = 0;           // current wavefront (state copy and count index)
  long long calls = 0;   // scheduler steps so far
  int iters = 0;         // iterations launched
};
int pt_jobs_start(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, int capacity, JobRun& run,
                  hipStream_t stream, std::string& err);
int pt_jobs_iterate(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, JobRun& run, int iters,
                    float4* out_L, float4* out_Ls, hipStream_t stream, std::string& err);
// host[0]: paths in the current wavefront, host[1]: iterations that found
// live paths (pinned memory); ev is recorded after the copies.
int pt_jobs_status(const PathBuffers& pb, const JobRun& run, long long* host, hipEvent_t ev, hipStream_t stream,
                   std::string& err);
// Iterates until the wavefront is empty, reading the status one batch of
// iterations late (host: 4 pinned words, ev: two events).
// live_iters (optional) receives the number of iterations that found live
// paths over the whole run.
int pt_jobs_drain(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, JobRun& run, float4* out_L,
                  float4* out_Ls, long long* host, const hipEvent_t* ev, hipStream_t stream, std::string& err,
                  int* live_iters = nullptr);
// Iterations after which the wavefront is certainly empty: every max_vertex
// + 1 iterations retire all paths that were in flight at their start.
inline int pt_jobs_bound(const PassConfig& cfg, int capacity) {
  const long long rounds = ((long long)cfg.n_slots + capacity - 1) / std::max(1, capacity) + 1;
  return (int)std::min<long long>(1 << 30, rounds * (std::max(0, cfg.max_vertex) + 1));
}

// Passes the streaming schedule keeps in flight at most (its radiance ring).
constexpr int kStreamRing = 16;
// Scheduler counters per stream group (64-bit).
constexpr int kStreamCtl = 64;
constexpr int kStreamHalf = 32;  // two copies of the scheduler state per group
static_assert(8 + kStreamRing <= kStreamHalf && 2 * kStreamHalf <= kStreamCtl, "scheduler state layout");

// Runs `passes` passes from cfg.accum_id with the streaming schedule: slots
// take (pass, pixel) jobs in pass order as their paths end, finished jobs
// leave their radiance in ring (kStreamRing x npx float4, by pass % ring and
// pixel) and each pass is accumulated into accum / average, in pass order,
// once all its jobs are done. The pixels are split into K groups (64-pixel
// chunks dealt round-robin, the last one possibly partial), group g on pbs[g]
// and streams[g] with ctl +
// g * kStreamCtl; live_host (pinned, 2K ints) and live_ev (2K events) poll the
// live-path counts. A timer records group 0's launches. The last pass's
// radiance is left in ring entry (passes - 1) % kStreamRing.
int pt_run_stream(const SceneView& sc, PathBuffers* const* pbs, const hipStream_t* streams, int K,
                  const PassConfig& cfg, int passes, float4* accum, float4* average, float4* ring, long long* ctl,
                  int* live_host, hipEvent_t* live_ev, std::string& err);
// Pixels of stream group g of K over npx pixels, and the pixel of a group's
// i-th pixel.
int stream_group_slots(int npx, int g, int K);
HM_HD int stream_group_pixel(int slot, int g, int K) { return ((slot >> 6) * K + g) * 64 + (slot & 63); }

// Adds pb.L into the running sum and mean buffers (accumulation id `accum_id`).
int pt_accumulate(const float4* L, float4* accum, float4* average, int n, int accum_id, hipStream_t stream,
                  std::string& err);

// Batched closest-hit / any-hit over an explicit ray list (8 floats per ray:
// origin xyz, pad, direction xyz, pad).
int trace_rays(const SceneView& sc, const float* rays, int n, int any_hit, int* prim, float* t, float* u,
               int* spill, int spill_stride, int spill_levels, int blocks, unsigned long long* stats,
               hipStream_t stream, std::string& err);

}  // namespace hm
