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
  uint64_t rng_state = 0;
  double setup_ms = 0.0, warmup_ms = 0.0;  // host wall clock of msnn_begin's two parts
  // device buffers
  float* nn_input = nullptr;     // W*H*12
  uint8_t* hit_flags = nullptr;  // W*H
  float4* short_color = nullptr; // W*H
  float* nn_output = nullptr;    // W*H*3
  float* train_input = nullptr;  // n_train*12
  float* train_target = nullptr; // n_train*3
  uint8_t* train_hit = nullptr;  // n_train
  int* train_idx = nullptr;      // n_train
  float* gen_points = nullptr;   // points of one training-data pass, n_train*3
  float4* acc[3] = {nullptr, nullptr, nullptr};  // short, network, final
  float4* avg[3] = {nullptr, nullptr, nullptr};
  double loss_last = 0.0;
  ~MsnnState() {
    if (mlp) mlp_destroy(mlp);
    dev_free(nn_input);
    dev_free(hit_flags);
    dev_free(short_color);
    dev_free(nn_output);
    dev_free(train_input);
    dev_free(train_target);
    dev_free(train_hit);
    dev_free(train_idx);
    dev_free(gen_points);
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
};

void msnn_release(MsnnState* m) { delete m; }

namespace {

// Training records and the G-buffer colour after a pass
// (cuda/hair_msnn.cu:243-311). train_slot < 0: pixel pass, records of the
// flagged pixels go to index pixel / every_nth; otherwise every slot is a
// record.
__global__ void k_msnn_finalize(PathBuffers pb, int n, int all_slots, int every_nth, int n_train,
                                const float* rec_in, const uint8_t* hit_in, float* train_input, float* train_target,
                                float4* short_color) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 L = pb.L[i];
  V3 color = v3(L.x, L.y, L.z);
  const bool training = all_slots || (pb.flags[i] & 1u);
  if (training) {
    const int tr = all_slots ? i : i / every_nth;
    const bool hit = hit_in[i] & 1u;
    V3 full = v3(0.f), sh = v3(0.f);
    if (hit) {
      full = color;
      const float4 S = pb.L_short[i];
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
  for (int i = 0; i < nseg; ++i) {
    const float* q = &c.hair.seg[8 * (size_t)i];
    for (int j = 0; j < count[i]; ++j, ++k) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      const float u = (float)((st >> 40) * (1.0 / 16777216.0));
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
  HM_TRY(dev_alloc((void**)&m->nn_input, sizeof(float) * 12 * (size_t)npx, err));
  HM_TRY(dev_alloc((void**)&m->hit_flags, (size_t)npx, err));
  HM_TRY(dev_alloc((void**)&m->short_color, sizeof(float4) * (size_t)npx, err));
  HM_TRY(dev_alloc((void**)&m->nn_output, sizeof(float) * 3 * (size_t)npx, err));
  HM_TRY(dev_alloc((void**)&m->train_input, sizeof(float) * 12 * (size_t)m->n_train, err));
  HM_TRY(dev_alloc((void**)&m->train_target, sizeof(float) * 3 * (size_t)m->n_train, err));
  HM_TRY(dev_alloc((void**)&m->train_hit, (size_t)m->n_train, err));
  HM_TRY(dev_alloc((void**)&m->train_idx, sizeof(int) * (size_t)m->n_train, err));
  HM_TRY(dev_alloc((void**)&m->gen_points, sizeof(float) * 3 * (size_t)m->n_train, err));
  for (int k = 0; k < 3; ++k) {
    HM_TRY(dev_alloc((void**)&m->acc[k], sizeof(float4) * (size_t)npx, err));
    HM_TRY(dev_alloc((void**)&m->avg[k], sizeof(float4) * (size_t)npx, err));
    HM_HIP(hipMemsetAsync(m->avg[k], 0, sizeof(float4) * (size_t)npx, c.stream), err);
  }
  HM_TRY(ensure_scene_points(c, s.msnn_seed, err));
  HM_TRY(c.ensure_path_buffers(true, err, m->n_train));
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

int sample_gen_points(Context& c, MsnnState& m, std::string& err) {
  hipLaunchKernelGGL(k_sample_points, dim3((m.n_train + 255) / 256), dim3(256), 0, c.stream, c.scene_points,
                     c.num_scene_points, m.next_u64(), m.n_train, m.gen_points);
  return check_launch("msnn sample points", err);
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

int train_step(Context& c, MsnnState& m, std::string& err, float* loss_out) {
  const float lr = lr_at(m.lr, m.train_steps);
  HM_TRY(mlp_loss_grad_step(*m.mlp, m.train_input, m.train_target, m.n_train, m.loss, nullptr, loss_out, true, lr,
                            c.stream, err));
  m.train_steps++;
  return kOk;
}

// Warm-up: training-data passes toward sampled hair points, each followed by
// one training step (render_hair_msnn.cu:633-665).
int msnn_warmup(Context& c, MsnnState& m, int steps, std::string& err) {
  for (int k = 0; k < steps; ++k) {
    HM_TRY(sample_gen_points(c, m, err));
    PassConfig cfg = msnn_pass_config(c, m);
    cfg.camera_mode = 1;
    cfg.n_slots = m.n_train;
    cfg.grid_w = m.grid_w;
    cfg.points = m.gen_points;
    cfg.accum_id = m.gen_id++;
    cfg.gb.enabled = 1;
    cfg.gb.nn_input = m.train_input;
    cfg.gb.hit_flags = m.train_hit;
    cfg.gb.scene_scale = c.scene_scale;
    HM_TRY(pt_run_pass(c.view, c.pb, cfg, c.stream, err));
    hipLaunchKernelGGL(k_msnn_finalize, dim3((m.n_train + 255) / 256), dim3(256), 0, c.stream, c.pb, m.n_train, 1, 1,
                       m.n_train, m.train_input, m.train_hit, m.train_input, m.train_target, (float4*)nullptr);
    HM_TRY(check_launch("msnn finalize (warm-up)", err));
    float l = 0.f;
    HM_TRY(train_step(c, m, err, &l));
    m.loss_last = l;
  }
  return kOk;
}

struct EventTimer {
  hipEvent_t e[8];
  EventTimer() {
    for (auto& x : e) (void)hipEventCreate(&x);
  }
  ~EventTimer() {
    for (auto& x : e) (void)hipEventDestroy(x);
  }
  static double ms(hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

int msnn_frame(Context& c, MsnnState& m, hmsnn_msnn_stats* st, std::string& err) {
  const int npx = c.scene.width * c.scene.height;
  // shuffle of the training-pixel indices (render_hair_msnn.cu:712-714)
  std::vector<int> idx(m.n_train);
  std::iota(idx.begin(), idx.end(), 0);
  for (int i = m.n_train - 1; i > 0; --i) std::swap(idx[i], idx[(int)(m.next_u64() % (uint64_t)(i + 1))]);
  HM_HIP(hipMemcpyAsync(m.train_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c.stream), err);
  EventTimer tm;
  (void)hipEventRecord(tm.e[0], c.stream);
  // ---- G_BUFFER ----
  PassConfig cfg = msnn_pass_config(c, m);
  cfg.accum_id = m.frame;
  cfg.train_idx = m.train_idx;
  cfg.every_nth = m.every_nth;
  cfg.n_slots = m.n_train;
  cfg.gb.enabled = 1;
  cfg.gb.nn_input = m.nn_input;
  cfg.gb.hit_flags = m.hit_flags;
  cfg.gb.scene_scale = c.scene_scale;
  HM_TRY(pt_run_pass(c.view, c.pb, cfg, c.stream, err));
  hipLaunchKernelGGL(k_msnn_finalize, dim3((npx + 255) / 256), dim3(256), 0, c.stream, c.pb, npx, 0, m.every_nth,
                     m.n_train, m.nn_input, m.hit_flags, m.train_input, m.train_target, m.short_color);
  HM_TRY(check_launch("msnn finalize", err));
  (void)hipEventRecord(tm.e[1], c.stream);
  // ---- train ----
  float l = 0.f;
  if (c.scene.nrc_train) HM_TRY(train_step(c, m, err, nullptr));
  (void)hipEventRecord(tm.e[2], c.stream);
  // ---- infer ----
  HM_TRY(mlp_forward(*m.mlp, m.nn_input, npx, m.nn_output, c.stream, err));
  (void)hipEventRecord(tm.e[3], c.stream);
  // ---- RENDER ----
  hipLaunchKernelGGL(k_msnn_composite, dim3((npx + 255) / 256), dim3(256), 0, c.stream, npx, m.hit_flags,
                     m.short_color, m.nn_output, m.acc[0], m.acc[1], m.acc[2], m.avg[0], m.avg[1], m.avg[2], m.frame);
  HM_TRY(check_launch("msnn composite", err));
  (void)hipEventRecord(tm.e[4], c.stream);
  m.frame++;
  if (st) {
    HM_HIP(hipEventSynchronize(tm.e[4]), err);
    st->gbuffer_ms += EventTimer::ms(tm.e[0], tm.e[1]);
    st->train_ms += EventTimer::ms(tm.e[1], tm.e[2]);
    st->infer_ms += EventTimer::ms(tm.e[2], tm.e[3]);
    st->render_ms += EventTimer::ms(tm.e[3], tm.e[4]);
    st->total_ms += EventTimer::ms(tm.e[0], tm.e[4]);
    st->frames += 1;
  }
  (void)l;
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
  hmsnn_msnn_stats local{};
  for (int k = 0; k < frames; ++k) {
    int rc = msnn_frame(c, *c.msnn, stats ? &local : nullptr, err);
    if (rc) return fail(rc, err);
  }
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  if (stats) {
    local.loss = c.msnn->loss_last;
    local.train_steps = c.msnn->train_steps;
    local.setup_ms = c.msnn->setup_ms;
    local.warmup_ms = c.msnn->warmup_ms;
    *stats = local;
  }
  return HMSNN_OK;
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

int hmsnn_render_hair_msnn(hmsnn_ctx* ctx, int BETA, int spp, float* rgba, hmsnn_msnn_stats* stats) {
  if (!ctx || spp <= 0) return fail(HMSNN_EINVAL, "hmsnn_render_hair_msnn: bad arguments");
  int rc = hmsnn_msnn_begin(ctx, BETA);
  if (rc) return rc;
  rc = hmsnn_msnn_frames(ctx, spp, stats);
  if (rc) return rc;
  return rgba ? hmsnn_msnn_get_image(ctx, 0, rgba) : HMSNN_OK;
}

}  // extern "C"
