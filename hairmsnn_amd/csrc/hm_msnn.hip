// HairMSNN frame loop (render_hair_msnn.cu:98-771, cuda/hair_msnn.cu:17-357):
//   warm-up : training-data passes (rays to sampled hair points) + training
//   per frame:
//     shuffle the training-pixel indices
//     G_BUFFER : path pass truncated after the vertex where depth > beta;
//                1 pixel in every_nth runs the full + short training path
//     train    : one Adam step of the radiance cache on those records
//     infer    : the cache at every pixel's primary hit
//     RENDER   : short path + cache on hair primary hits, accumulated
// The reference's warm-up runs for at least one second of wall-clock time
// (render_hair_msnn.cu:633-643); here it is a fixed number of steps so runs
// are reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"
#include "hm_context.h"
#include "hm_mlp.h"

namespace hm {

// Per-frame state of one pipeline slot: its path buffers and G-buffer /
// training arrays, and the stream its path-tracing passes run on. Consecutive
// frames (and warm-up steps) use consecutive slots, so their passes overlap;
// training, inference and compositing run in frame order on the context
// stream.
struct MsnnSlot {
  PathBuffers pb{};
  std::vector<void*> allocs;
  hipStream_t st = nullptr;
  float* nn_input = nullptr;     // W*H*12
  uint8_t* hit_flags = nullptr;  // W*H
  float4* short_color = nullptr; // W*H
  float* train_input = nullptr;  // n_train*12
  float* train_target = nullptr; // n_train*3
  uint8_t* train_hit = nullptr;  // n_train
  int* train_idx = nullptr;      // n_train
  float* gen_points = nullptr;   // points of one training-data pass, n_train*3
  int* idx_host = nullptr;       // pinned staging of train_idx
  hipEvent_t ev_gb = nullptr;    // the slot's pass and finalize are done
  hipEvent_t ev_done = nullptr;  // the context stream has consumed the slot
  hipEvent_t ev_copy = nullptr;  // idx_host has been read
  bool used = false;
  void release() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : allocs) dev_free(p);
    allocs.clear();
    if (idx_host) (void)hipHostFree(idx_host);
    idx_host = nullptr;
    for (hipEvent_t* e : {&ev_gb, &ev_done, &ev_copy})
      if (*e) (void)hipEventDestroy(*e), *e = nullptr;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
  }
};

// Training paths of up to kWarmBatch consecutive frames, traced as one pass
// (they do not depend on the network). Double-buffered: a batch is filled
// while the previous one is being consumed.
struct TrainBatch {
  PathBuffers pb{};
  std::vector<void*> allocs;
  int* idx_dev = nullptr;       // shuffled training indices, frame-major
  int* idx_host = nullptr;      // pinned staging of idx_dev
  int* pix = nullptr;           // pixel of each record
  float* nn_input = nullptr;    // G-buffer record of each training path
  uint8_t* hit = nullptr;
  float* train_input = nullptr;
  float* train_target = nullptr;
  float4* full = nullptr;       // full-path colour (the training pixel's G-buffer colour)
  hipEvent_t ev_ready = nullptr, ev_free = nullptr, ev_copy = nullptr;
  bool used = false;
  void release() {
    for (void* p : allocs) dev_free(p);
    allocs.clear();
    if (idx_host) (void)hipHostFree(idx_host);
    idx_host = nullptr;
    for (hipEvent_t* e : {&ev_ready, &ev_free, &ev_copy})
      if (*e) (void)hipEventDestroy(*e), *e = nullptr;
  }
};

struct MsnnState {
  MlpState* mlp = nullptr;
  int BETA = 1;
  int kbeta = 0;       // kernel beta = BETA - 1
  int n_train = 0;     // training records per frame
  int every_nth = 1;
  int grid_w = 256;    // training-data launch width
  int loss = HMSNN_LOSS_RELATIVE_L2_LUMINANCE;
  float lr = 1e-2f;
  int train_steps = 0;
  int frame = 0;       // accumulated frames (the reference's accumId)
  int gen_id = 0;      // accumId used by the warm-up passes
  int next_slot = 0;
  uint64_t rng_state = 0;
  double setup_ms = 0.0, warmup_ms = 0.0;  // host wall clock of msnn_begin's two parts
  // frames in flight: 64 C1 frames took 0.33 s with 3 and 0.41 s with 4 in
  // the benchmark process; a process running only HairMSNN favoured 4
  // (0.26 s against 0.32 s): the GPU has 4 hardware queues, and which streams
  // share one depends on every stream the process created
  static constexpr int kSlots = 3;
  // warm-up training-data passes batched per path-tracing pass (the passes do
  // not depend on the network); slot arrays hold this many batches
  static constexpr int kWarmBatch = 8;
  int warm_batch = 1;
  MsnnSlot slots[kSlots];
  TrainBatch tb[2];
  float* nn_output = nullptr;    // W*H*3 (used in frame order on the context stream)
  float4* acc[3] = {nullptr, nullptr, nullptr};  // short, network, final
  float4* avg[3] = {nullptr, nullptr, nullptr};
  double loss_last = 0.0;
  hipStream_t ctx_stream = nullptr;
  hipStream_t batch_stream = nullptr;  // training batches traced one batch ahead
  int batches_made = 0;                // training batches generated so far
  int call_end = 0;                    // frame index one past the current hmsnn_msnn_frames call
  int slot_paths = 0;                      // paths each slot's path buffers hold
  long long* job_status = nullptr;         // pinned status words of warm-up job runs
  hipEvent_t job_ev[2] = {nullptr, nullptr};
  ~MsnnState() {
    if (ctx_stream) (void)hipStreamSynchronize(ctx_stream);
    if (batch_stream) (void)hipStreamSynchronize(batch_stream);
    if (batch_stream) (void)hipStreamDestroy(batch_stream);
    for (MsnnSlot& sl : slots) sl.release();
    for (TrainBatch& b : tb) b.release();
    if (job_status) (void)hipHostFree(job_status);
    for (hipEvent_t& e : job_ev)
      if (e) (void)hipEventDestroy(e);
    if (mlp) mlp_destroy(mlp);
    dev_free(nn_output);
    for (int k = 0; k < 3; ++k) {
      dev_free(acc[k]);
      dev_free(avg[k]);
    }
  }
  uint64_t next_u64() {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  MsnnSlot& take_slot() {
    MsnnSlot& sl = slots[next_slot];
    next_slot = (next_slot + 1) % kSlots;
    return sl;
  }
};

void msnn_release(MsnnState* m) { delete m; }

namespace {

// Training records and the G-buffer colour after a pass
// (cuda/hair_msnn.cu:243-311). train_slot < 0: pixel pass, records of the
// flagged pixels go to index pixel / every_nth; otherwise every slot is a
// record.
__global__ void k_msnn_finalize(const float4* Lf, const float4* Lsh, const uint32_t* flags, int n, int all_slots,
                                int every_nth, int n_train, const float* rec_in, const uint8_t* hit_in,
                                float* train_input, float* train_target, float4* short_color) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 L = Lf[i];
  V3 color = v3(L.x, L.y, L.z);
  const bool training = all_slots || (flags[i] & 1u);
  if (training) {
    const int tr = all_slots ? i : i / every_nth;
    const bool hit = hit_in[i] & 1u;
    V3 full = v3(0.f), sh = v3(0.f);
    if (hit) {
      full = color;
      const float4 S = Lsh[i];
      sh = v3(S.x, S.y, S.z);
    }
    if (any_nan(full)) full = v3(0.f);
    if (any_inf(full)) full = v3(1e5f);
    if (any_nan(sh)) sh = v3(0.01f);
    if (any_inf(sh)) sh = v3(1e5f);
    if (tr < n_train) {
      for (int k = 0; k < 12; ++k) train_input[(size_t)tr * 12 + k] = rec_in[(size_t)i * 12 + k];
      train_target[(size_t)tr * 3 + 0] = full.x - sh.x;
      train_target[(size_t)tr * 3 + 1] = full.y - sh.y;
      train_target[(size_t)tr * 3 + 2] = full.z - sh.z;
    }
    if (hit) color = full;  // a missed training pixel keeps the environment (Le)
  }
  if (short_color) short_color[i] = make_float4(color.x, color.y, color.z, 0.f);
}

// RENDER pass composite and the three accumulations (cuda/hair_msnn.cu:314-356).
__global__ void k_msnn_composite(int n, const uint8_t* hit_flags, const float4* short_color, const float* nn_out,
                                 float4* acc_s, float4* acc_n, float4* acc_f, float4* avg_s, float4* avg_n,
                                 float4* avg_f, int accum_id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 s4 = short_color[i];
  V3 sc = v3(s4.x, s4.y, s4.z);
  V3 nn = v3(nn_out[3 * (size_t)i], nn_out[3 * (size_t)i + 1], nn_out[3 * (size_t)i + 2]);
  const uint8_t f = hit_flags[i];
  V3 col;
  if (!(f & 1u) || (f & 2u)) {  // escaped or surface: short path only
    col = sc;
    nn = col;
  } else {
    col = sc + nn;
  }
  // NaN scrub as in writePixel (utils.cuh:20-21); the reference's composite
  // has none, which lets one NaN sample poison a pixel's running sum.
  if (any_nan(sc)) sc = v3(0.f);
  if (any_nan(nn)) nn = v3(0.f);
  if (any_nan(col)) col = v3(0.f);
  if (accum_id > 0) {
    const float4 a = acc_s[i], b = acc_n[i], c = acc_f[i];
    sc = sc + v3(a.x, a.y, a.z);
    nn = nn + v3(b.x, b.y, b.z);
    col = col + v3(c.x, c.y, c.z);
  }
  acc_s[i] = make_float4(sc.x, sc.y, sc.z, 1.f);
  acc_n[i] = make_float4(nn.x, nn.y, nn.z, 1.f);
  acc_f[i] = make_float4(col.x, col.y, col.z, 1.f);
  const float inv = 1.f / (float)(accum_id + 1);
  avg_s[i] = make_float4(sc.x * inv, sc.y * inv, sc.z * inv, 1.f);
  avg_n[i] = make_float4(nn.x * inv, nn.y * inv, nn.z * inv, 1.f);
  avg_f[i] = make_float4(col.x * inv, col.y * inv, col.z * inv, 1.f);
}

// A training pixel's G-buffer colour is its full path (the reference keeps
// the training path's colour there, cuda/hair_msnn.cu:270-279).
__global__ void k_msnn_merge(int n, const int* pix, const float4* full, float4* short_color) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) short_color[pix[j]] = full[j];
}

// Learning-rate schedule of the reference's optimizer config
// (tcnn_hairmsnn.json:5-18): ExponentialDecay from step 4000, x0.33 every 4000.
float lr_at(float base, int step) {
  if (step < 4000) return base;
  return base * std::pow(0.33f, (float)((step - 4000) / 4000 + 1));
}

// Scene samples on the hair (fetchSceneSamples, render_hair_msnn.cu:34-97):
// about one million points, per segment in proportion to its length and at
// least one per segment, uniform along the segment. Made once per scene and
// kept on the device.
int ensure_scene_points(Context& c, uint64_t seed, std::string& err) {
  if (c.scene_points) return kOk;
  const int nseg = (int)(c.hair.seg.size() / 8);
  double total = 0.0;
  std::vector<double> len(nseg);
  for (int i = 0; i < nseg; ++i) {
    const float* q = &c.hair.seg[8 * (size_t)i];
    len[i] = std::sqrt((double)(q[4] - q[0]) * (q[4] - q[0]) + (double)(q[5] - q[1]) * (q[5] - q[1]) +
                       (double)(q[6] - q[2]) * (q[6] - q[2]));
    total += len[i];
  }
  const double target_samples = 1000000.0;
  std::vector<int> count(nseg);
  size_t np = 0;
  for (int i = 0; i < nseg; ++i) {
    count[i] = std::max(1, (int)(len[i] / total * target_samples));
    np += (size_t)count[i];
  }
  std::vector<float> pts(3 * np);
  uint64_t st = seed * 0x2545F4914F6CDD1Dull + 3;
  size_t k = 0;
  const bool curves = !c.hair.crv.empty();
  for (int i = 0; i < nseg; ++i) {
    const float* q = &c.hair.seg[8 * (size_t)i];
    CurveSeg cs{};
    if (curves) {
      // on the round curve itself (interpolator.position3(eps1),
      // render_hair_msnn.cu:51-64); its chord length sets the count
      const float* w = &c.hair.crv[(size_t)kCurveFloats * i];
      cs = make_curve(make_float4(w[0], w[1], w[2], w[3]), make_float4(w[4], w[5], w[6], w[7]),
                      make_float4(w[8], w[9], w[10], w[11]), make_float4(w[12], w[13], w[14], w[15]));
    }
    for (int j = 0; j < count[i]; ++j, ++k) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      const float u = (float)((st >> 40) * (1.0 / 16777216.0));
      if (curves) {
        const V3 p = curve_pos(cs, u);
        pts[3 * k] = p.x;
        pts[3 * k + 1] = p.y;
        pts[3 * k + 2] = p.z;
        continue;
      }
      pts[3 * k] = q[0] + u * (q[4] - q[0]);
      pts[3 * k + 1] = q[1] + u * (q[5] - q[1]);
      pts[3 * k + 2] = q[2] + u * (q[6] - q[2]);
    }
  }
  HM_TRY(dev_alloc((void**)&c.scene_points, sizeof(float) * pts.size(), err));
  HM_HIP(hipMemcpy(c.scene_points, pts.data(), sizeof(float) * pts.size(), hipMemcpyHostToDevice), err);
  c.num_scene_points = (int)np;
  return kOk;
}

int msnn_setup(Context& c, int BETA, std::string& err) {
  if (!c.scene.has_hair) {
    err = "HairMSNN needs hair geometry";
    return kEInval;
  }
  msnn_release(c.msnn);
  c.msnn = nullptr;
  MsnnState* m = new MsnnState();
  c.msnn = m;
  const Scene& s = c.scene;
  m->BETA = BETA;
  m->kbeta = BETA - 1;  // render_hair_msnn.cu:1170
  m->n_train = std::max(128, s.msnn_train_records / 128 * 128);
  const int npx = s.width * s.height;
  m->every_nth = std::max(1, npx / m->n_train);  // render_hair_msnn.cu:106 (integer division)
  m->grid_w = 1;
  while (m->grid_w * m->grid_w < m->n_train) m->grid_w <<= 1;
  m->lr = s.msnn_lr;
  m->loss = s.msnn_loss == "RelativeL2" ? HMSNN_LOSS_RELATIVE_L2
            : s.msnn_loss == "L2"        ? HMSNN_LOSS_L2
                                         : HMSNN_LOSS_RELATIVE_L2_LUMINANCE;
  m->rng_state = s.msnn_seed * 0x2545F4914F6CDD1Dull + 1;
  const int prec = s.msnn_precision == "bf16" ? HMSNN_BF16 : s.msnn_precision == "fp8" ? HMSNN_FP8 : HMSNN_FP16;
  HM_TRY(mlp_create(prec, s.msnn_seed, &m->mlp, err));
  // tcnn.init_weights (scene.cpp:330-331 -> render_hair_msnn.cu:109-110)
  if (!s.nrc_weights.empty()) HM_TRY(mlp_load(*m->mlp, scene_path(s, s.nrc_weights), err));
  m->ctx_stream = c.stream;
  const int spill_levels = c.stack_levels();
  // a warm-up pass holds warm_batch training batches; the slots' path buffers
  // hold a frame or one such pass
  m->warm_batch = std::max(1, std::min(MsnnState::kWarmBatch, std::max(npx, m->n_train) / m->n_train));
  const size_t rec = (size_t)m->warm_batch * m->n_train;
  m->slot_paths = (int)std::max((size_t)npx, rec);
  HM_HIP(hipHostMalloc((void**)&m->job_status, sizeof(long long) * 4), err);
  for (hipEvent_t& e : m->job_ev) HM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), err);
  for (MsnnSlot& sl : m->slots) {
    HM_HIP(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking), err);
    for (hipEvent_t* e : {&sl.ev_gb, &sl.ev_done, &sl.ev_copy})
      HM_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming), err);
    HM_TRY(alloc_path_buffers(sl.pb, sl.allocs, std::max((size_t)npx, rec), true, spill_levels,
                              c.trace_blocks * kTraceBlock, sl.st, err));
    auto A = [&](void** p, size_t bytes) -> int {
      HM_TRY(dev_alloc(p, bytes, err));
      sl.allocs.push_back(*p);
      return kOk;
    };
    HM_TRY(A((void**)&sl.nn_input, sizeof(float) * 12 * (size_t)npx));
    HM_TRY(A((void**)&sl.hit_flags, (size_t)npx));
    HM_TRY(A((void**)&sl.short_color, sizeof(float4) * (size_t)npx));
    HM_TRY(A((void**)&sl.train_input, sizeof(float) * 12 * rec));
    HM_TRY(A((void**)&sl.train_target, sizeof(float) * 3 * rec));
    HM_TRY(A((void**)&sl.train_hit, rec));
    HM_TRY(A((void**)&sl.train_idx, sizeof(int) * (size_t)m->n_train));
    HM_TRY(A((void**)&sl.gen_points, sizeof(float) * 3 * rec));
    HM_HIP(hipHostMalloc((void**)&sl.idx_host, sizeof(int) * (size_t)m->n_train), err);
  }
  for (TrainBatch& b : m->tb) {
    for (hipEvent_t* e : {&b.ev_ready, &b.ev_free, &b.ev_copy})
      HM_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming), err);
    HM_TRY(alloc_path_buffers(b.pb, b.allocs, rec, true, spill_levels, c.trace_blocks * kTraceBlock, c.stream, err));
    auto A = [&](void** p, size_t bytes) -> int {
      HM_TRY(dev_alloc(p, bytes, err));
      b.allocs.push_back(*p);
      return kOk;
    };
    HM_TRY(A((void**)&b.idx_dev, sizeof(int) * rec));
    HM_TRY(A((void**)&b.pix, sizeof(int) * rec));
    HM_TRY(A((void**)&b.nn_input, sizeof(float) * 12 * rec));
    HM_TRY(A((void**)&b.hit, rec));
    HM_TRY(A((void**)&b.train_input, sizeof(float) * 12 * rec));
    HM_TRY(A((void**)&b.train_target, sizeof(float) * 3 * rec));
    HM_TRY(A((void**)&b.full, sizeof(float4) * rec));
    HM_HIP(hipHostMalloc((void**)&b.idx_host, sizeof(int) * rec), err);
  }
  HM_TRY(dev_alloc((void**)&m->nn_output, sizeof(float) * 3 * (size_t)npx, err));
  for (int k = 0; k < 3; ++k) {
    HM_TRY(dev_alloc((void**)&m->acc[k], sizeof(float4) * (size_t)npx, err));
    HM_TRY(dev_alloc((void**)&m->avg[k], sizeof(float4) * (size_t)npx, err));
    HM_HIP(hipMemsetAsync(m->avg[k], 0, sizeof(float4) * (size_t)npx, c.stream), err);
  }
  // the slots' first passes start after the set-up work on the context stream
  HM_HIP(hipStreamSynchronize(c.stream), err);
  for (MsnnSlot& sl : m->slots) HM_HIP(hipStreamSynchronize(sl.st), err);
  HM_TRY(ensure_scene_points(c, s.msnn_seed, err));
  return kOk;
}

// Training-data targets for one warm-up pass: n uniform draws (with
// replacement) from the scene samples, hashed from (seed, i). The reference
// shuffles its sample indices and takes the first records
// (render_hair_msnn.cu:640-660); over ~3M samples the two differ only in the
// rare repeated draw.
__global__ void k_sample_points(const float* pts, int num_points, uint64_t seed, int n, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const int p = (int)(z % (uint64_t)num_points);
  out[3 * (size_t)i] = pts[3 * (size_t)p];
  out[3 * (size_t)i + 1] = pts[3 * (size_t)p + 1];
  out[3 * (size_t)i + 2] = pts[3 * (size_t)p + 2];
}

// Draws for `batches` consecutive training-data passes, batch b at
// gen_points + b * n_train (one seed per pass, in pass order).
int sample_gen_points(Context& c, MsnnState& m, hipStream_t st, float* dst, int batches, std::string& err) {
  for (int b = 0; b < batches; ++b)
    hipLaunchKernelGGL(k_sample_points, dim3((m.n_train + 255) / 256), dim3(256), 0, st, c.scene_points,
                       c.num_scene_points, m.next_u64(), m.n_train, dst + 3 * (size_t)b * m.n_train);
  return check_launch("msnn sample points", err);
}

// The slot's stream waits until the context stream has consumed the slot's
// previous contents.
int acquire_slot(MsnnSlot& sl, std::string& err) {
  if (sl.used) HM_HIP(hipStreamWaitEvent(sl.st, sl.ev_done, 0), err);
  sl.used = true;
  return kOk;
}

PassConfig msnn_pass_config(const Context& c, const MsnnState& m) {
  PassConfig cfg = c.pt_pass_config();
  cfg.msnn = 1;
  cfg.dl_zero = 1;  // msnnPathTrace lights every vertex (hair_msnn.cu:142, 166)
  cfg.dl_from = 1;
  cfg.max_vertex = c.scene.path_v2 - 1;
  cfg.trunc_beta = m.kbeta;
  return cfg;
}

int train_step(Context& c, MsnnState& m, const float* input, const float* target, std::string& err,
               float* loss_out) {
  const float lr = lr_at(m.lr, m.train_steps);
  HM_TRY(mlp_loss_grad_step(*m.mlp, input, target, m.n_train, m.loss, nullptr, loss_out, true, lr, c.stream, err));
  m.train_steps++;
  return kOk;
}

// Warm-up: training-data passes toward sampled hair points, each followed by
// one training step (render_hair_msnn.cu:633-665). The passes do not depend
// on the network, so consecutive ones run on consecutive slots and overlap;
// the training steps run in order on the context stream. This is the
// per-pass schedule; msnn_warmup_jobs traces the same records.
int msnn_warmup_passes(Context& c, MsnnState& m, int steps, std::string& err) {
  for (int k = 0; k < steps;) {
    const int nb = std::min(m.warm_batch, steps - k);
    const int n = nb * m.n_train;
    MsnnSlot& sl = m.take_slot();
    HM_TRY(acquire_slot(sl, err));
    HM_TRY(sample_gen_points(c, m, sl.st, sl.gen_points, nb, err));
    PassConfig cfg = msnn_pass_config(c, m);
    cfg.camera_mode = 1;
    cfg.n_slots = n;
    cfg.chunk = m.n_train;
    cfg.grid_w = m.grid_w;
    cfg.points = sl.gen_points;
    cfg.accum_id = m.gen_id;
    m.gen_id += nb;
    cfg.gb.enabled = 1;
    cfg.gb.nn_input = sl.train_input;
    cfg.gb.hit_flags = sl.train_hit;
    cfg.gb.scene_scale = c.scene_scale;
    HM_TRY(pt_run_pass(c.view_no_caps(), sl.pb, cfg, sl.st, err));
    hipLaunchKernelGGL(k_msnn_finalize, dim3((n + 255) / 256), dim3(256), 0, sl.st, sl.pb.L, sl.pb.L_short,
                       sl.pb.flags, n, 1, 1, n, sl.train_input, sl.train_hit, sl.train_input, sl.train_target,
                       (float4*)nullptr);
    HM_TRY(check_launch("msnn finalize (warm-up)", err));
    HM_HIP(hipEventRecord(sl.ev_gb, sl.st), err);
    HM_HIP(hipStreamWaitEvent(c.stream, sl.ev_gb, 0), err);
    for (int b = 0; b < nb; ++b)
      HM_TRY(train_step(c, m, sl.train_input + 12 * (size_t)b * m.n_train, sl.train_target + 3 * (size_t)b * m.n_train,
                        err, nullptr));
    HM_HIP(hipEventRecord(sl.ev_done, c.stream), err);
    k += nb;
  }
  return kOk;
}

// Warm-up steps per job run: their training-data records are traced as one
// job-schedule run (the slot's path buffers refilled with new records as
// paths finish), then trained on in step order while the next run traces.
constexpr int kWarmJobSteps = 32;

// The records of every warm-up step, as msnn_warmup_passes makes them (same
// points, seeds and paths), on the job schedule: runs of up to
// kWarmJobSteps steps alternate between two record buffers, so a run traces
// while the previous one's steps train.
int msnn_warmup_jobs(Context& c, MsnnState& m, int steps, std::string& err) {
  if (steps <= 0) return kOk;
  const int per = std::min(steps, kWarmJobSteps);
  const size_t rec = (size_t)per * m.n_train;
  struct Buf {
    float* pts = nullptr;
    float* input = nullptr;
    float* target = nullptr;
    uint8_t* hit = nullptr;
    float4* L = nullptr;
    float4* Ls = nullptr;
    hipEvent_t ev_free = nullptr;
    bool used = false;
  } buf[2];
  std::vector<void*> allocs;
  auto release = [&]() {
    (void)hipStreamSynchronize(c.stream);
    for (void* p : allocs) dev_free(p);
    for (Buf& b : buf)
      if (b.ev_free) (void)hipEventDestroy(b.ev_free);
  };
  auto body = [&]() -> int {
    const int nbuf = steps > per ? 2 : 1;
    for (int k = 0; k < nbuf; ++k) {
      Buf& b = buf[k];
      auto A = [&](void** p, size_t bytes) -> int {
        HM_TRY(dev_alloc(p, bytes, err));
        allocs.push_back(*p);
        return kOk;
      };
      HM_TRY(A((void**)&b.pts, sizeof(float) * 3 * rec));
      HM_TRY(A((void**)&b.input, sizeof(float) * 12 * rec));
      HM_TRY(A((void**)&b.target, sizeof(float) * 3 * rec));
      HM_TRY(A((void**)&b.hit, rec));
      HM_TRY(A((void**)&b.L, sizeof(float4) * rec));
      HM_TRY(A((void**)&b.Ls, sizeof(float4) * rec));
      HM_HIP(hipEventCreateWithFlags(&b.ev_free, hipEventDisableTiming), err);
    }
    const SceneView sc = c.view_no_caps();
    for (int k = 0, r = 0; k < steps; ++r) {
      const int nb = std::min(per, steps - k);
      const int n = nb * m.n_train;
      Buf& b = buf[r % 2];
      MsnnSlot& sl = m.take_slot();
      HM_TRY(acquire_slot(sl, err));
      if (b.used) HM_HIP(hipStreamWaitEvent(sl.st, b.ev_free, 0), err);
      b.used = true;
      HM_TRY(sample_gen_points(c, m, sl.st, b.pts, nb, err));
      PassConfig cfg = msnn_pass_config(c, m);
      cfg.camera_mode = 1;
      cfg.n_slots = n;
      cfg.chunk = m.n_train;
      cfg.grid_w = m.grid_w;
      cfg.points = b.pts;
      cfg.accum_id = m.gen_id;
      m.gen_id += nb;
      cfg.gb.enabled = 1;
      cfg.gb.nn_input = b.input;
      cfg.gb.hit_flags = b.hit;
      cfg.gb.scene_scale = c.scene_scale;
      JobRun run;
      HM_TRY(pt_jobs_start(sc, sl.pb, cfg, std::min(n, m.slot_paths), run, sl.st, err));
      HM_TRY(pt_jobs_drain(sc, sl.pb, cfg, run, b.L, b.Ls, m.job_status, m.job_ev, sl.st, err));
      hipLaunchKernelGGL(k_msnn_finalize, dim3((n + 255) / 256), dim3(256), 0, sl.st, b.L, b.Ls,
                         (const uint32_t*)nullptr, n, 1, 1, n, b.input, b.hit, b.input, b.target, (float4*)nullptr);
      HM_TRY(check_launch("msnn finalize (warm-up)", err));
      HM_HIP(hipEventRecord(sl.ev_gb, sl.st), err);
      HM_HIP(hipStreamWaitEvent(c.stream, sl.ev_gb, 0), err);
      for (int j = 0; j < nb; ++j)
        HM_TRY(train_step(c, m, b.input + 12 * (size_t)j * m.n_train, b.target + 3 * (size_t)j * m.n_train, err,
                          nullptr));
      HM_HIP(hipEventRecord(b.ev_free, c.stream), err);
      HM_HIP(hipEventRecord(sl.ev_done, c.stream), err);
      k += nb;
    }
    return kOk;
  };
  const int rc = body();
  release();
  return rc;
}

int msnn_warmup(Context& c, MsnnState& m, int steps, std::string& err) {
  return c.pt_schedule == 0 ? msnn_warmup_jobs(c, m, steps, err) : msnn_warmup_passes(c, m, steps, err);
}

struct EventTimer {
  hipEvent_t e[8];
  EventTimer() {
    for (auto& x : e) (void)hipEventCreate(&x);
  }
  EventTimer(const EventTimer&) = delete;
  EventTimer& operator=(const EventTimer&) = delete;
  ~EventTimer() {
    for (auto& x : e) (void)hipEventDestroy(x);
  }
  static double ms(hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

// Training paths of frames f0 .. f0 + warm_batch - 1 as one pass on stream
// st, into batch buffer (f0 / warm_batch) % 2. The frames' index shuffles
// (render_hair_msnn.cu:712-714) are drawn here, in frame order. The batch
// runs on the per-pass schedule on either setting of the path-tracing
// schedule: traced on the job schedule, the frame loop measured slower
// (64 C1 frames: 3.97 ms per frame with all records in flight, 4.15 with
// half, 4.45 with a quarter, against 3.75), because the batch's dense
// kernels overlap the frames' G_BUFFER passes less well.
int gen_train_batch(Context& c, MsnnState& m, int f0, hipStream_t st, std::string& err) {
  const int B = m.warm_batch;
  TrainBatch& b = m.tb[(f0 / B) % 2];
  if (b.used) HM_HIP(hipStreamWaitEvent(st, b.ev_free, 0), err);
  b.used = true;
  HM_HIP(hipEventSynchronize(b.ev_copy), err);
  for (int f = 0; f < B; ++f) {
    int* idx = b.idx_host + (size_t)f * m.n_train;
    for (int i = 0; i < m.n_train; ++i) idx[i] = i;
    for (int i = m.n_train - 1; i > 0; --i) std::swap(idx[i], idx[(int)(m.next_u64() % (uint64_t)(i + 1))]);
  }
  const int n = B * m.n_train;
  HM_HIP(hipMemcpyAsync(b.idx_dev, b.idx_host, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st), err);
  HM_HIP(hipEventRecord(b.ev_copy, st), err);
  PassConfig cfg = msnn_pass_config(c, m);
  cfg.camera_mode = 2;
  cfg.n_slots = n;
  cfg.chunk = m.n_train;
  cfg.every_nth = m.every_nth;
  cfg.train_idx = b.idx_dev;
  cfg.pix = b.pix;
  cfg.accum_id = f0;
  cfg.gb.enabled = 1;
  cfg.gb.nn_input = b.nn_input;
  cfg.gb.hit_flags = b.hit;
  cfg.gb.scene_scale = c.scene_scale;
  HM_TRY(pt_run_pass(c.view_no_caps(), b.pb, cfg, st, err));
  hipLaunchKernelGGL(k_msnn_finalize, dim3((n + 255) / 256), dim3(256), 0, st, b.pb.L, b.pb.L_short, b.pb.flags, n,
                     1, 1, n, b.nn_input, b.hit, b.train_input, b.train_target, b.full);
  HM_TRY(check_launch("msnn finalize (training batch)", err));
  HM_HIP(hipEventRecord(b.ev_ready, st), err);
  return kOk;
}

// One frame (render_hair_msnn.cu:699-771). On the frame's slot: the G_BUFFER
// pass over every pixel, truncated at the kernel beta, and its finalize; the
// frame's training pixels, traced in full by their batch, then overwrite
// their G-buffer colour. In frame order on the context stream: training on
// the frame's records, inference, and the RENDER composite. tm, when given,
// receives events: [0], [1] around the slot's work; [2]..[5] on the context
// stream.
int msnn_frame(Context& c, MsnnState& m, EventTimer* tm, std::string& err) {
  const int npx = c.scene.width * c.scene.height;
  const int B = m.warm_batch;
  const int f = m.frame, chunk = f % B;
  MsnnSlot& sl = m.take_slot();
  HM_TRY(acquire_slot(sl, err));
  if (tm) (void)hipEventRecord(tm->e[0], sl.st);
  if (chunk == 0) {
    // The frame's batch (unless it was traced ahead), then the next batch on
    // the batch stream, so that its long tail of late-depth launches overlaps
    // this batch's frames instead of delaying the next batch's first frame.
    // Training paths do not depend on the network, and the batches draw their
    // shuffles in the same order either way. A batch is traced ahead only
    // when its first frame belongs to the current call, so every batch a call
    // starts is consumed (and waited for) by that call's frames.
    if (m.batches_made <= f / B) {
      HM_TRY(gen_train_batch(c, m, f, sl.st, err));
      m.batches_made = f / B + 1;
    }
    if (f + B < m.call_end) {
      if (!m.batch_stream) HM_HIP(hipStreamCreateWithFlags(&m.batch_stream, hipStreamNonBlocking), err);
      HM_TRY(gen_train_batch(c, m, f + B, m.batch_stream, err));
      m.batches_made = f / B + 2;
    }
  }
  TrainBatch& b = m.tb[(f / B) % 2];
  // ---- G_BUFFER (every pixel truncated; no training flags) ----
  PassConfig cfg = msnn_pass_config(c, m);
  cfg.accum_id = f;
  cfg.every_nth = m.every_nth;
  cfg.n_slots = 0;
  // the last vertex a truncated path lights: the break test `bounces > beta`
  // runs from the second vertex on, so kernel beta -1 (BETA 0) truncates like
  // beta 0 (hair_msnn.cu:166-179)
  cfg.max_vertex = std::min(cfg.max_vertex, std::max(m.kbeta, 0) + 1);
  cfg.gb.enabled = 1;
  cfg.gb.nn_input = sl.nn_input;
  cfg.gb.hit_flags = sl.hit_flags;
  cfg.gb.scene_scale = c.scene_scale;
  HM_TRY(pt_run_pass(c.view_no_caps(), sl.pb, cfg, sl.st, err));
  hipLaunchKernelGGL(k_msnn_finalize, dim3((npx + 255) / 256), dim3(256), 0, sl.st, sl.pb.L, sl.pb.L_short,
                     sl.pb.flags, npx, 0, m.every_nth, 0,
                     sl.nn_input, sl.hit_flags, (float*)nullptr, (float*)nullptr, sl.short_color);
  HM_TRY(check_launch("msnn finalize", err));
  HM_HIP(hipStreamWaitEvent(sl.st, b.ev_ready, 0), err);
  const size_t r0 = (size_t)chunk * m.n_train;
  hipLaunchKernelGGL(k_msnn_merge, dim3((m.n_train + 255) / 256), dim3(256), 0, sl.st, m.n_train, b.pix + r0,
                     b.full + r0, sl.short_color);
  HM_TRY(check_launch("msnn merge", err));
  if (tm) (void)hipEventRecord(tm->e[1], sl.st);
  HM_HIP(hipEventRecord(sl.ev_gb, sl.st), err);
  HM_HIP(hipStreamWaitEvent(c.stream, sl.ev_gb, 0), err);
  if (tm) (void)hipEventRecord(tm->e[2], c.stream);
  // ---- train ----
  if (c.scene.nrc_train) HM_TRY(train_step(c, m, b.train_input + 12 * r0, b.train_target + 3 * r0, err, nullptr));
  if (chunk == B - 1) HM_HIP(hipEventRecord(b.ev_free, c.stream), err);
  if (tm) (void)hipEventRecord(tm->e[3], c.stream);
  // ---- infer ----
  HM_TRY(mlp_forward(*m.mlp, sl.nn_input, npx, m.nn_output, c.stream, err));
  if (tm) (void)hipEventRecord(tm->e[4], c.stream);
  // ---- RENDER ----
  hipLaunchKernelGGL(k_msnn_composite, dim3((npx + 255) / 256), dim3(256), 0, c.stream, npx, sl.hit_flags,
                     sl.short_color, m.nn_output, m.acc[0], m.acc[1], m.acc[2], m.avg[0], m.avg[1], m.avg[2], m.frame);
  HM_TRY(check_launch("msnn composite", err));
  if (tm) (void)hipEventRecord(tm->e[5], c.stream);
  HM_HIP(hipEventRecord(sl.ev_done, c.stream), err);
  m.frame++;
  return kOk;
}

}  // namespace

}  // namespace hm

using namespace hm;

struct hmsnn_ctx {
  Context c;
};

extern "C" {

int hmsnn_msnn_begin(hmsnn_ctx* ctx, int BETA) {
  if (!ctx) return fail(HMSNN_EINVAL, "NULL context");
  std::string err;
  Context& c = ctx->c;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = msnn_setup(c, BETA, err);
  if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) rc = HMSNN_EDEVICE, err = "setup failed";
  const auto t1 = std::chrono::steady_clock::now();
  if (!rc) rc = msnn_warmup(c, *c.msnn, c.scene.msnn_warmup_steps, err);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
    const auto t2 = std::chrono::steady_clock::now();
    c.msnn->setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    c.msnn->warmup_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  }
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_msnn_frames(hmsnn_ctx* ctx, int frames, hmsnn_msnn_stats* stats) {
  if (!ctx || frames < 0) return fail(HMSNN_EINVAL, "hmsnn_msnn_frames: bad arguments");
  Context& c = ctx->c;
  if (!c.msnn) return fail(HMSNN_EINVAL, "call hmsnn_msnn_begin first");
  std::string err;
  std::vector<EventTimer> timers(stats ? (size_t)frames : 0);
  c.msnn->call_end = c.msnn->frame + frames;
  for (int k = 0; k < frames; ++k) {
    int rc = msnn_frame(c, *c.msnn, stats ? &timers[k] : nullptr, err);
    if (rc) return fail(rc, err);
  }
  hipError_t e = hipStreamSynchronize(c.stream);
  // every batch traced ahead was consumed by a frame of this call, so the
  // batch stream is idle here; waiting on it keeps that true if the rule
  // above ever changes
  if (e == hipSuccess && c.msnn->batch_stream) e = hipStreamSynchronize(c.msnn->batch_stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  if (stats) {
    hmsnn_msnn_stats local{};
    for (EventTimer& tm : timers) {
      local.gbuffer_ms += EventTimer::ms(tm.e[0], tm.e[1]);
      local.train_ms += EventTimer::ms(tm.e[2], tm.e[3]);
      local.infer_ms += EventTimer::ms(tm.e[3], tm.e[4]);
      local.render_ms += EventTimer::ms(tm.e[4], tm.e[5]);
    }
    if (frames > 0) local.total_ms = EventTimer::ms(timers[0].e[0], timers[frames - 1].e[5]);
    local.frames = frames;
    local.loss = c.msnn->loss_last;
    local.train_steps = c.msnn->train_steps;
    local.setup_ms = c.msnn->setup_ms;
    local.warmup_ms = c.msnn->warmup_ms;
    *stats = local;
  }
  return HMSNN_OK;
}

int hmsnn_msnn_save_weights(hmsnn_ctx* ctx, const char* path) {
  if (!ctx || !path) return fail(HMSNN_EINVAL, "hmsnn_msnn_save_weights: bad arguments");
  Context& c = ctx->c;
  if (!c.msnn) return fail(HMSNN_EINVAL, "call hmsnn_msnn_begin first");
  std::string err;
  int rc = hipStreamSynchronize(c.stream) == hipSuccess ? kOk : kEDevice;
  if (rc) err = "hmsnn_msnn_save_weights: device error";
  if (!rc) rc = mlp_save(*c.msnn->mlp, path, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_msnn_get_image(hmsnn_ctx* ctx, int which, float* rgba) {
  if (!ctx || !rgba || which < 0 || which > 2) return fail(HMSNN_EINVAL, "hmsnn_msnn_get_image: bad arguments");
  Context& c = ctx->c;
  if (!c.msnn) return fail(HMSNN_EINVAL, "call hmsnn_msnn_begin first");
  // which: 0 final, 1 short path, 2 network
  const int slot = which == 0 ? 2 : (which == 1 ? 0 : 1);
  const size_t bytes = sizeof(float4) * (size_t)c.scene.width * c.scene.height;
  hipError_t e = hipMemcpyAsync(rgba, c.msnn->avg[slot], bytes, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, hipGetErrorString(e));
}

int hmsnn_msnn_get_sizes(hmsnn_ctx* ctx, int sizes[4]) {
  if (!ctx || !sizes) return fail(HMSNN_EINVAL, "hmsnn_msnn_get_sizes: bad arguments");
  const MsnnState* m = ctx->c.msnn;
  if (!m) return fail(HMSNN_EINVAL, "call hmsnn_msnn_begin first");
  sizes[0] = m->n_train;
  sizes[1] = m->every_nth;
  sizes[2] = m->frame;
  sizes[3] = m->train_steps;
  return HMSNN_OK;
}

int hmsnn_msnn_get_buffer(hmsnn_ctx* ctx, int which, void* host, size_t bytes) {
  if (!ctx || !host) return fail(HMSNN_EINVAL, "hmsnn_msnn_get_buffer: bad arguments");
  Context& c = ctx->c;
  MsnnState* m = c.msnn;
  if (!m) return fail(HMSNN_EINVAL, "call hmsnn_msnn_begin first");
  if (m->frame == 0) return fail(HMSNN_EINVAL, "no frame rendered yet");
  const size_t npx = (size_t)c.scene.width * c.scene.height, nt = (size_t)m->n_train;
  const int f = m->frame - 1, B = m->warm_batch;
  const MsnnSlot& sl = m->slots[(m->next_slot + MsnnState::kSlots - 1) % MsnnState::kSlots];
  const TrainBatch& b = m->tb[(f / B) % 2];
  const size_t r0 = (size_t)(f % B) * nt;
  const void* src = nullptr;
  size_t want = 0;
  switch (which) {
    case HMSNN_MSNN_TRAIN_INPUT: src = b.train_input + 12 * r0, want = sizeof(float) * 12 * nt; break;
    case HMSNN_MSNN_TRAIN_TARGET: src = b.train_target + 3 * r0, want = sizeof(float) * 3 * nt; break;
    case HMSNN_MSNN_TRAIN_PIXEL: src = b.pix + r0, want = sizeof(int) * nt; break;
    case HMSNN_MSNN_GB_INPUT: src = sl.nn_input, want = sizeof(float) * 12 * npx; break;
    case HMSNN_MSNN_GB_HIT: src = sl.hit_flags, want = npx; break;
    case HMSNN_MSNN_GB_COLOR: src = sl.short_color, want = sizeof(float4) * npx; break;
    case HMSNN_MSNN_NN_OUTPUT: src = m->nn_output, want = sizeof(float) * 3 * npx; break;
    default: return fail(HMSNN_EINVAL, "hmsnn_msnn_get_buffer: unknown buffer");
  }
  if (bytes != want) return fail(HMSNN_EINVAL, "hmsnn_msnn_get_buffer: size mismatch");
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, hipGetErrorString(e));
}

int hmsnn_render_hair_msnn(hmsnn_ctx* ctx, int BETA, int spp, float* rgba, hmsnn_msnn_stats* stats) {
  if (!ctx || spp <= 0) return fail(HMSNN_EINVAL, "hmsnn_render_hair_msnn: bad arguments");
  int rc = hmsnn_msnn_begin(ctx, BETA);
  if (rc) return rc;
  rc = hmsnn_msnn_frames(ctx, spp, stats);
  if (rc) return rc;
  return rgba ? hmsnn_msnn_get_image(ctx, 0, rgba) : HMSNN_OK;
}

}  // extern "C"
