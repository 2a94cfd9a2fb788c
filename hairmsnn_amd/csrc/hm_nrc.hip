// Neural radiance caching, the reference's comparator renderer
// (render_nrc.cu:116-701, cuda/nrc.cu:18-398). Per frame:
//   shuffle  : the training-pixel indices (one pixel in every_nth)
//   G_BUFFER : every pixel's path until its spread exceeds the camera
//              footprint, where it queries the cache; training pixels go on
//              with a suffix that ends in a second query (or, for every 16th,
//              runs to the end unbiased) and record every vertex
//   infer    : the cache at all queries (frame + training suffixes)
//   RENDER   : pixel = radiance before the query + throughput x cache;
//              training targets = suffix radiance sums + throughput x cache
//   train    : one Adam step on the training records
// The cache network is the fused MLP of hm_mlp.hip (12-float records; NRC
// fills position / scene scale, wo and the shading normal, channels 9-11 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"
#include "hm_context.h"
#include "hm_mlp.h"

namespace hm {

// Per-frame state of one pipeline slot: path buffers, the G_BUFFER arrays
// and training records of one frame, and the stream its G_BUFFER pass runs
// on. The pass does not depend on the network, so consecutive frames' passes
// run on consecutive slots and overlap; inference, RENDER and training run in
// frame order on the context stream.
struct NrcSlot {
  PathBuffers pb{};
  std::vector<void*> allocs;
  hipStream_t st = nullptr;
  int* train_idx = nullptr;
  int* idx_host = nullptr;  // pinned staging of train_idx
  float* query = nullptr;   // n_query x 12
  float4* gb_beta = nullptr;
  uint8_t* gb_hit = nullptr;
  float4* state = nullptr;
  float4* prev = nullptr;
  float4* tb[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // pos, wo, n, rad, beta
  int* tb_bounces = nullptr;
  uint8_t* tb_hit = nullptr;
  hipEvent_t ev_gb = nullptr;    // the slot's G_BUFFER pass is done
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

struct NrcState {
  MlpState* mlp = nullptr;
  int n_rec = 0;       // training records per step (numTrainingRecords)
  int n_tpx = 0;       // training pixels (numTrainingPixels)
  int every_nth = 1;
  int frame_size = 0;  // pixels
  int n_query = 0;     // nnFrameSize: frame + training queries, a multiple of 128
  float c = 0.01f;
  int all_unbiased = 0;
  int loss = HMSNN_LOSS_RELATIVE_L2_LUMINANCE;
  float lr = 1e-2f;
  int train_steps = 0;
  int frame = 0;
  uint64_t rng_state = 0;
  std::vector<int> idx;  // current permutation of the training indices
  // frames in flight: 4 took 10.3 ms per C1 frame against 12.3 ms with 3
  // (0.54 s against 0.57 s for 64 frames in the benchmark process)
  static constexpr int kSlots = 4;
  NrcSlot slots[kSlots];
  int next_slot = 0, last_slot = -1;
  // shared, used in frame order on the context stream
  std::vector<void*> allocs;
  float* cache = nullptr;    // n_query x 3
  float4* gb_rad = nullptr;  // radiance before the query (RENDER pass input), per pixel
  float* train_in = nullptr;   // n_rec x 12
  float* train_tgt = nullptr;  // n_rec x 3
  float4* acc = nullptr;
  float4* avg = nullptr;
  int* counts_dev = nullptr;   // [0] frame cache queries, [1] training records with data
  hipStream_t ctx_stream = nullptr;
  ~NrcState() {
    if (ctx_stream) (void)hipStreamSynchronize(ctx_stream);
    for (NrcSlot& sl : slots) sl.release();
    for (void* p : allocs) dev_free(p);
    if (mlp) mlp_destroy(mlp);
  }
  uint64_t next_u64() {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

void nrc_release(NrcState* s) { delete s; }

namespace {

// Adds v, summed over a 256-thread block, to *counter with one atomic per
// block (every wave of a frame-sized launch adding to the one counter
// serialised there). Every thread of the block calls it.
__device__ __forceinline__ void block_count(int* counter, int v) {
  __shared__ int part[4];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int sum = part[0] + part[1] + part[2] + part[3];
    if (sum) atomicAdd(counter, sum);
  }
}

// RENDER pass, pixel side (cuda/nrc.cu:378-392): radiance before the query
// plus throughput x max(cache, 0), a NaN cache counting as 0; accumulated
// like writePixel (utils.cuh:13-35, NaN scrub). Also stores the G-buffer
// radiance for inspection and counts the frame's cache queries.
__global__ void k_nrc_render(PathBuffers pb, NrcPassArgs nr, const float* cache, float4* gb_rad, float4* acc,
                             float4* avg, int accum_id, int* counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < nr.frame_size;
  bool queried = false;
  if (in) {
    const float4 l4 = (pb.flags[i] & kNrcSuffix) ? pb.L_short[i] : pb.L[i];
    const float4 b4 = nr.gb_beta[i];
    V3 cr = v3(cache[3 * (size_t)i], cache[3 * (size_t)i + 1], cache[3 * (size_t)i + 2]);
    if (any_nan(cr)) cr = v3(0.f);
    cr = v3(fmaxf(cr.x, 0.f), fmaxf(cr.y, 0.f), fmaxf(cr.z, 0.f));
    V3 col = v3(l4.x, l4.y, l4.z) + v3(b4.x, b4.y, b4.z) * cr;
    gb_rad[i] = make_float4(l4.x, l4.y, l4.z, 0.f);
    queried = b4.x != 0.f || b4.y != 0.f || b4.z != 0.f;
    if (any_nan(col)) col = v3(0.f);
    if (accum_id > 0) {
      const float4 a = acc[i];
      col = col + v3(a.x, a.y, a.z);
    }
    acc[i] = make_float4(col.x, col.y, col.z, 1.f);
    const float inv = 1.f / (float)(accum_id + 1);
    avg[i] = make_float4(col.x * inv, col.y * inv, col.z * inv, 1.f);
  }
  block_count(&counts[0], queried ? 1 : 0);
}

// RENDER pass, training side (nrcGenerateTrainingData, cuda/nrc.cu:69-133):
// for training pixel tr (whose camera ray hit) and each recorded vertex b <
// bounces, target = sum over the later vertices of their direct light times
// the throughput from b, plus that throughput times the suffix's cache value
// (0 for unbiased suffixes or suffixes without a query); NaN targets become
// 0. The sums run in the reference's order. Records beyond `bounces` stay 0.
__global__ void k_nrc_train_data(NrcPassArgs nr, const int* train_idx, const float* cache, float* train_in,
                                 float* train_tgt, int n_rec, int* counts) {
  const int tr = blockIdx.x * blockDim.x + threadIdx.x;
  int used = 0;
  if (tr < nr.n_tpx) {
    const int px = tr * nr.every_nth + train_idx[tr] % nr.every_nth;
    if (px < nr.frame_size && nr.gb_hit[px]) {
      const int B = nr.tb_bounces[tr];
      V3 cr = v3(0.f);
      if (nr.tb_hit[tr] && tr % 16 != 0 && !nr.all_unbiased) {
        const size_t qi = (size_t)nr.frame_size + tr;
        cr = v3(cache[3 * qi], cache[3 * qi + 1], cache[3 * qi + 2]);
      }
      const size_t base = (size_t)tr * kNrcMaxBounces;
      for (int b = 0; b < B && b < kNrcMaxBounces; ++b) {
        V3 color = v3(0.f), beta = v3(1.f);
        for (int sub = b; sub < B && sub < kNrcMaxBounces; ++sub) {
          const float4 r = nr.tb_rad[base + sub], vb = nr.tb_beta[base + sub];
          color = color + beta * v3(r.x, r.y, r.z);
          beta = beta * v3(vb.x, vb.y, vb.z);
        }
        const size_t rec = base + b;
        if ((int)rec >= n_rec) break;
        const float4 p = nr.tb_pos[rec], wo = nr.tb_wo[rec], n = nr.tb_n[rec];
        float* in = train_in + rec * kNnInputCh;
        in[0] = p.x; in[1] = p.y; in[2] = p.z;
        in[3] = wo.x; in[4] = wo.y; in[5] = wo.z;
        in[6] = n.x; in[7] = n.y; in[8] = n.z;
        in[9] = 0.f; in[10] = 0.f; in[11] = 0.f;
        V3 t = color + beta * cr;
        if (any_nan(t)) t = v3(0.f);
        train_tgt[3 * rec] = t.x;
        train_tgt[3 * rec + 1] = t.y;
        train_tgt[3 * rec + 2] = t.z;
        ++used;
      }
    }
  }
  block_count(&counts[1], used);  // the records written
}

struct NrcTimer {
  hipEvent_t e[6];
  NrcTimer() {
    for (auto& x : e) (void)hipEventCreate(&x);
  }
  NrcTimer(const NrcTimer&) = delete;
  NrcTimer& operator=(const NrcTimer&) = delete;
  ~NrcTimer() {
    for (auto& x : e) (void)hipEventDestroy(x);
  }
  static double ms(hipEvent_t a, hipEvent_t b) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

float nrc_lr_at(float base, int step) {
  if (step < 4000) return base;
  return base * std::pow(0.33f, (float)((step - 4000) / 4000 + 1));
}

int nrc_setup(Context& c, std::string& err) {
  nrc_release(c.nrc);
  c.nrc = nullptr;
  NrcState* m = new NrcState();
  c.nrc = m;
  const Scene& s = c.scene;
  const int npx = s.width * s.height;
  m->frame_size = npx;
  m->n_rec = std::max(128, s.nrc_train_records / 128 * 128);
  // render_nrc.cu:130-146, integer divisions as in the reference
  m->n_tpx = std::max(1, m->n_rec / kNrcMaxBounces);
  m->every_nth = std::max(1, npx / m->n_tpx);
  m->n_query = npx + m->n_tpx - m->n_tpx % 128 + 128;
  m->c = s.nrc_c;
  m->all_unbiased = s.nrc_all_unbiased ? 1 : 0;
  m->lr = s.msnn_lr;
  m->loss = s.msnn_loss == "RelativeL2" ? HMSNN_LOSS_RELATIVE_L2
            : s.msnn_loss == "L2"        ? HMSNN_LOSS_L2
                                         : HMSNN_LOSS_RELATIVE_L2_LUMINANCE;
  m->rng_state = s.msnn_seed * 0x2545F4914F6CDD1Dull + 5;
  const int prec = s.msnn_precision == "bf16" ? HMSNN_BF16 : s.msnn_precision == "fp8" ? HMSNN_FP8 : HMSNN_FP16;
  HM_TRY(mlp_create(prec, s.msnn_seed, &m->mlp, err));
  // tcnn.init_weights (scene.cpp:330-331 -> render_nrc.cu:149-150)
  if (!s.nrc_weights.empty()) HM_TRY(mlp_load(*m->mlp, scene_path(s, s.nrc_weights), err));
  m->ctx_stream = c.stream;
  auto A = [&](void** p, size_t bytes) -> int {
    HM_TRY(dev_alloc(p, bytes, err));
    m->allocs.push_back(*p);
    return kOk;
  };
  const size_t nv = (size_t)m->n_tpx * kNrcMaxBounces;
  for (NrcSlot& sl : m->slots) {
    HM_HIP(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking), err);
    for (hipEvent_t* e : {&sl.ev_gb, &sl.ev_done, &sl.ev_copy})
      HM_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming), err);
    HM_TRY(alloc_path_buffers(sl.pb, sl.allocs, (size_t)npx, true, c.stack_levels(), c.trace_blocks * kTraceBlock,
                              sl.st, err));
    auto S = [&](void** p, size_t bytes) -> int {
      HM_TRY(dev_alloc(p, bytes, err));
      sl.allocs.push_back(*p);
      return kOk;
    };
    HM_TRY(S((void**)&sl.train_idx, sizeof(int) * (size_t)m->n_tpx));
    HM_TRY(S((void**)&sl.query, sizeof(float) * kNnInputCh * (size_t)m->n_query));
    HM_TRY(S((void**)&sl.gb_beta, sizeof(float4) * (size_t)npx));
    HM_TRY(S((void**)&sl.gb_hit, (size_t)npx));
    HM_TRY(S((void**)&sl.state, sizeof(float4) * (size_t)npx));
    HM_TRY(S((void**)&sl.prev, sizeof(float4) * (size_t)npx));
    for (float4*& t : sl.tb) HM_TRY(S((void**)&t, sizeof(float4) * nv));
    HM_TRY(S((void**)&sl.tb_bounces, sizeof(int) * (size_t)m->n_tpx));
    HM_TRY(S((void**)&sl.tb_hit, (size_t)m->n_tpx));
    HM_HIP(hipHostMalloc((void**)&sl.idx_host, sizeof(int) * (size_t)m->n_tpx), err);
  }
  HM_TRY(A((void**)&m->cache, sizeof(float) * 3 * (size_t)m->n_query));
  HM_TRY(A((void**)&m->gb_rad, sizeof(float4) * (size_t)npx));
  HM_TRY(A((void**)&m->train_in, sizeof(float) * kNnInputCh * (size_t)m->n_rec));
  HM_TRY(A((void**)&m->train_tgt, sizeof(float) * 3 * (size_t)m->n_rec));
  HM_TRY(A((void**)&m->acc, sizeof(float4) * (size_t)npx));
  HM_TRY(A((void**)&m->avg, sizeof(float4) * (size_t)npx));
  HM_TRY(A((void**)&m->counts_dev, sizeof(int) * 4));
  HM_HIP(hipMemsetAsync(m->avg, 0, sizeof(float4) * (size_t)npx, c.stream), err);
  m->idx.resize(m->n_tpx);
  for (int i = 0; i < m->n_tpx; ++i) m->idx[i] = i;  // thrust::sequence (render_nrc.cu:170-173)
  // the slots' first passes start after the set-up work
  HM_HIP(hipStreamSynchronize(c.stream), err);
  for (NrcSlot& sl : m->slots) HM_HIP(hipStreamSynchronize(sl.st), err);
  return kOk;
}

NrcPassArgs nrc_args(const Context& c, const NrcState& m, const NrcSlot& sl) {
  NrcPassArgs a{};
  a.enabled = 1;
  a.all_unbiased = m.all_unbiased;
  a.c = m.c;
  a.scene_scale = c.scene_scale;
  a.frame_size = m.frame_size;
  a.n_tpx = m.n_tpx;
  a.every_nth = m.every_nth;
  a.query = sl.query;
  a.gb_beta = sl.gb_beta;
  a.gb_hit = sl.gb_hit;
  a.state = sl.state;
  a.prev = sl.prev;
  a.tb_pos = sl.tb[0];
  a.tb_wo = sl.tb[1];
  a.tb_n = sl.tb[2];
  a.tb_rad = sl.tb[3];
  a.tb_beta = sl.tb[4];
  a.tb_bounces = sl.tb_bounces;
  a.tb_hit = sl.tb_hit;
  return a;
}

// One frame (render_nrc.cu:640-701): the shuffle and G_BUFFER pass on the
// frame's slot, then inference, RENDER and training in frame order on the
// context stream. tm, when given, receives events: [0], [1] around the
// G_BUFFER pass on the slot's stream; [2] inference start, [3] RENDER start,
// [4] training start and [5] end on the context stream.
int nrc_frame(Context& c, NrcState& m, NrcTimer* tm, std::string& err) {
  const int npx = m.frame_size;
  NrcSlot& sl = m.slots[m.next_slot];
  m.last_slot = m.next_slot;
  m.next_slot = (m.next_slot + 1) % NrcState::kSlots;
  hipStream_t st = sl.st, cs = c.stream;
  if (sl.used) HM_HIP(hipStreamWaitEvent(st, sl.ev_done, 0), err);
  sl.used = true;
  if (tm) (void)hipEventRecord(tm->e[0], st);
  // ---- shuffle (render_nrc.cu:652-655): seeded, applied to the previous order ----
  HM_HIP(hipEventSynchronize(sl.ev_copy), err);  // the slot's staging buffer is free again
  for (int i = m.n_tpx - 1; i > 0; --i) std::swap(m.idx[i], m.idx[(int)(m.next_u64() % (uint64_t)(i + 1))]);
  std::copy(m.idx.begin(), m.idx.end(), sl.idx_host);
  HM_HIP(hipMemcpyAsync(sl.train_idx, sl.idx_host, sizeof(int) * (size_t)m.n_tpx, hipMemcpyHostToDevice, st), err);
  HM_HIP(hipEventRecord(sl.ev_copy, st), err);
  // ---- clear the frame's buffers (owlBufferClear + RESET, render_nrc.cu:684-692) ----
  const size_t nv = (size_t)m.n_tpx * kNrcMaxBounces;
  HM_HIP(hipMemsetAsync(sl.query, 0, sizeof(float) * kNnInputCh * (size_t)m.n_query, st), err);
  HM_HIP(hipMemsetAsync(sl.gb_beta, 0, sizeof(float4) * (size_t)npx, st), err);
  HM_HIP(hipMemsetAsync(sl.gb_hit, 0, (size_t)npx, st), err);
  for (float4* t : sl.tb) HM_HIP(hipMemsetAsync(t, 0, sizeof(float4) * nv, st), err);
  HM_HIP(hipMemsetAsync(sl.tb_bounces, 0, sizeof(int) * (size_t)m.n_tpx, st), err);
  HM_HIP(hipMemsetAsync(sl.tb_hit, 0, (size_t)m.n_tpx, st), err);
  // ---- G_BUFFER ----
  PassConfig cfg = c.pt_pass_config();
  cfg.accum_id = m.frame;
  cfg.dl_zero = 1;  // nrcTracePaths lights every vertex (nrc.cu:166-169)
  cfg.dl_from = 1;
  cfg.max_vertex = kNrcMaxBounces - 1;
  cfg.train_idx = sl.train_idx;
  cfg.every_nth = m.every_nth;
  cfg.nrc = nrc_args(c, m, sl);
  HM_TRY(pt_run_pass(c.view_no_caps(), sl.pb, cfg, st, err));
  if (tm) (void)hipEventRecord(tm->e[1], st);
  HM_HIP(hipEventRecord(sl.ev_gb, st), err);
  HM_HIP(hipStreamWaitEvent(cs, sl.ev_gb, 0), err);
  // ---- infer (context stream, frame order) ----
  if (tm) (void)hipEventRecord(tm->e[2], cs);
  HM_HIP(hipMemsetAsync(m.train_in, 0, sizeof(float) * kNnInputCh * (size_t)m.n_rec, cs), err);
  HM_HIP(hipMemsetAsync(m.train_tgt, 0, sizeof(float) * 3 * (size_t)m.n_rec, cs), err);
  HM_HIP(hipMemsetAsync(m.counts_dev, 0, sizeof(int) * 4, cs), err);
  HM_TRY(mlp_forward(*m.mlp, sl.query, m.n_query, m.cache, cs, err));
  if (tm) (void)hipEventRecord(tm->e[3], cs);
  // ---- RENDER ----
  hipLaunchKernelGGL(k_nrc_render, dim3((npx + 255) / 256), dim3(256), 0, cs, sl.pb, cfg.nrc, m.cache, m.gb_rad,
                     m.acc, m.avg, m.frame, m.counts_dev);
  hipLaunchKernelGGL(k_nrc_train_data, dim3((m.n_tpx + 255) / 256), dim3(256), 0, cs, cfg.nrc, sl.train_idx,
                     m.cache, m.train_in, m.train_tgt, m.n_rec, m.counts_dev);
  HM_TRY(check_launch("nrc render", err));
  if (tm) (void)hipEventRecord(tm->e[4], cs);
  // ---- train ----
  if (c.scene.nrc_train) {
    const float lr = nrc_lr_at(m.lr, m.train_steps);
    HM_TRY(mlp_loss_grad_step(*m.mlp, m.train_in, m.train_tgt, m.n_rec, m.loss, nullptr, nullptr, true, lr, cs,
                              err));
    m.train_steps++;
  }
  if (tm) (void)hipEventRecord(tm->e[5], cs);
  HM_HIP(hipEventRecord(sl.ev_done, cs), err);
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

int hmsnn_nrc_begin(hmsnn_ctx* ctx) {
  if (!ctx) return fail(HMSNN_EINVAL, "NULL context");
  std::string err;
  Context& c = ctx->c;
  int rc = nrc_setup(c, err);
  if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) rc = HMSNN_EDEVICE, err = "NRC setup failed";
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_nrc_set_params(hmsnn_ctx* ctx, float c_spread, int all_unbiased) {
  if (!ctx || !(c_spread >= 0.f)) return fail(HMSNN_EINVAL, "hmsnn_nrc_set_params: bad arguments");
  if (!ctx->c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  ctx->c.nrc->c = c_spread;
  ctx->c.nrc->all_unbiased = all_unbiased ? 1 : 0;
  return HMSNN_OK;
}

int hmsnn_nrc_frames(hmsnn_ctx* ctx, int frames, hmsnn_nrc_stats* stats) {
  if (!ctx || frames < 0) return fail(HMSNN_EINVAL, "hmsnn_nrc_frames: bad arguments");
  Context& c = ctx->c;
  if (!c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  std::string err;
  std::vector<NrcTimer> timers(stats ? (size_t)frames : 0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < frames; ++k) {
    int rc = nrc_frame(c, *c.nrc, stats ? &timers[k] : nullptr, err);
    if (rc) return fail(rc, err);
  }
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  const auto t1 = std::chrono::steady_clock::now();
  if (stats) {
    hmsnn_nrc_stats local{};
    for (NrcTimer& tm : timers) {
      local.gbuffer_ms += NrcTimer::ms(tm.e[0], tm.e[1]);
      local.infer_ms += NrcTimer::ms(tm.e[2], tm.e[3]);
      local.render_ms += NrcTimer::ms(tm.e[3], tm.e[4]);
      local.train_ms += NrcTimer::ms(tm.e[4], tm.e[5]);
    }
    local.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    local.frames = frames;
    local.train_steps = c.nrc->train_steps;
    int counts[4] = {0, 0, 0, 0};
    e = hipMemcpy(counts, c.nrc->counts_dev, sizeof(counts), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(HMSNN_EDEVICE, hipGetErrorString(e));
    local.cache_queries = counts[0];
    local.train_records_used = counts[1];
    *stats = local;
  }
  return HMSNN_OK;
}

int hmsnn_nrc_get_image(hmsnn_ctx* ctx, float* rgba) {
  if (!ctx || !rgba) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_image: bad arguments");
  Context& c = ctx->c;
  if (!c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  const size_t bytes = sizeof(float4) * (size_t)c.scene.width * c.scene.height;
  hipError_t e = hipMemcpyAsync(rgba, c.nrc->avg, bytes, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, hipGetErrorString(e));
}

int hmsnn_nrc_get_sizes(hmsnn_ctx* ctx, int sizes[4]) {
  if (!ctx || !sizes) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_sizes: bad arguments");
  const NrcState* m = ctx->c.nrc;
  if (!m) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  sizes[0] = m->n_tpx;
  sizes[1] = m->every_nth;
  sizes[2] = m->n_query;
  sizes[3] = m->n_rec;
  return HMSNN_OK;
}

int hmsnn_nrc_get_buffer(hmsnn_ctx* ctx, int which, void* host, size_t bytes) {
  if (!ctx || !host) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_buffer: bad arguments");
  const NrcState* m = ctx->c.nrc;
  if (!m) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  const size_t npx = (size_t)m->frame_size;
  if (m->last_slot < 0) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_buffer: no frame rendered yet");
  const NrcSlot& sl = m->slots[m->last_slot];
  const void* src = nullptr;
  size_t need = 0;
  switch (which) {
    case HMSNN_NRC_QUERY: src = sl.query, need = sizeof(float) * kNnInputCh * (size_t)m->n_query; break;
    case HMSNN_NRC_CACHE: src = m->cache, need = sizeof(float) * 3 * (size_t)m->n_query; break;
    case HMSNN_NRC_GB_RADIANCE: src = m->gb_rad, need = sizeof(float4) * npx; break;
    case HMSNN_NRC_GB_BETA: src = sl.gb_beta, need = sizeof(float4) * npx; break;
    case HMSNN_NRC_GB_HIT: src = sl.gb_hit, need = npx; break;
    case HMSNN_NRC_TRAIN_INPUT: src = m->train_in, need = sizeof(float) * kNnInputCh * (size_t)m->n_rec; break;
    case HMSNN_NRC_TRAIN_TARGET: src = m->train_tgt, need = sizeof(float) * 3 * (size_t)m->n_rec; break;
    case HMSNN_NRC_TRAIN_IDX: src = sl.train_idx, need = sizeof(int) * (size_t)m->n_tpx; break;
    case HMSNN_NRC_TRAIN_BOUNCES: src = sl.tb_bounces, need = sizeof(int) * (size_t)m->n_tpx; break;
    case HMSNN_NRC_TRAIN_HIT: src = sl.tb_hit, need = (size_t)m->n_tpx; break;
    default: return fail(HMSNN_EINVAL, "hmsnn_nrc_get_buffer: unknown buffer");
  }
  if (bytes != need) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_buffer: size mismatch");
  hipError_t e = hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->c.stream);
  return e == hipSuccess ? HMSNN_OK : fail(HMSNN_EDEVICE, hipGetErrorString(e));
}

int hmsnn_nrc_get_mlp_weights(hmsnn_ctx* ctx, float* w_host, int n) {
  if (!ctx || !w_host || n != kMlpWeightCount) return fail(HMSNN_EINVAL, "hmsnn_nrc_get_mlp_weights: bad arguments");
  if (!ctx->c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  std::string err;
  if (hipStreamSynchronize(ctx->c.stream) != hipSuccess) return fail(HMSNN_EDEVICE, "stream failed");
  int rc = mlp_get_weights(*ctx->c.nrc->mlp, w_host, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_nrc_save_weights(hmsnn_ctx* ctx, const char* path) {
  if (!ctx || !path) return fail(HMSNN_EINVAL, "hmsnn_nrc_save_weights: bad arguments");
  if (!ctx->c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  std::string err;
  if (hipStreamSynchronize(ctx->c.stream) != hipSuccess) return fail(HMSNN_EDEVICE, "stream failed");
  int rc = mlp_save(*ctx->c.nrc->mlp, path, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_nrc_set_mlp_weights(hmsnn_ctx* ctx, const float* w_host, int n) {
  if (!ctx || !w_host || n != kMlpWeightCount) return fail(HMSNN_EINVAL, "hmsnn_nrc_set_mlp_weights: bad arguments");
  if (!ctx->c.nrc) return fail(HMSNN_EINVAL, "call hmsnn_nrc_begin first");
  std::string err;
  if (hipStreamSynchronize(ctx->c.stream) != hipSuccess) return fail(HMSNN_EDEVICE, "stream failed");
  int rc = mlp_set_weights(*ctx->c.nrc->mlp, w_host, err);
  return rc ? fail(rc, err) : HMSNN_OK;
}

int hmsnn_render_nrc(hmsnn_ctx* ctx, int spp, float* rgba, hmsnn_nrc_stats* stats) {
  if (!ctx || spp <= 0) return fail(HMSNN_EINVAL, "hmsnn_render_nrc: bad arguments");
  int rc = hmsnn_nrc_begin(ctx);
  if (rc) return rc;
  rc = hmsnn_nrc_frames(ctx, spp, stats);
  if (rc) return rc;
  return rgba ? hmsnn_nrc_get_image(ctx, rgba) : HMSNN_OK;
}

}  // extern "C"
