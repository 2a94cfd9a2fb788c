// Wavefront path-tracing kernels.
//
// One sample pass over the image runs:
//   camera   : seed each pixel's stream, jitter, primary ray   (path_tracing.cu:19-50)
//   per vertex depth = 0 .. max_vertex:
//     extend : closest hit for every live path (BVH traversal)
//     shade  : hit setup, direct-lighting samples, Russian roulette and the
//              continuation sample; survivors are compacted into the next queue
//     shadow : any-hit tests for the pending light samples, then accumulate
//   accumulate: running sum and mean per pixel                   (utils.cuh:13-35)
// Queues are compacted with one atomic per wave (ballot + popcount). Kernels
// read the live count from device memory, so a pass needs no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hm_pt.h"
#include "hm_pt_kernels.h"

namespace hm {

namespace {

__device__ __forceinline__ int wave_append(int* counter, bool pred) {
  const unsigned long long mask = __ballot(pred);
  const int lane = __lane_id();
  const int total = __popcll(mask);
  int base = 0;
  const int leader = __ffsll((long long)mask) - 1;
  if (total > 0 && lane == leader) base = atomicAdd(counter, total);
  base = __shfl(base, leader < 0 ? 0 : leader);
  const int rank = __popcll(mask & ((1ull << lane) - 1ull));
  return pred ? base + rank : -1;
}

__device__ __forceinline__ TraceStack make_stack(int* lds_base, const PathBuffers& pb) {
  TraceStack st;
  st.lds = lds_base + threadIdx.x;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  st.spill = pb.spill + gtid;
  st.spill_stride = pb.spill_stride;
  st.spill_levels = pb.spill_levels;
  st.sp = 0;
  return st;
}

// Wave-reduces the traversal counters and adds them to nodes/segs/rays.
__device__ __forceinline__ void flush_counters(const TraceCounters& c, unsigned long long* nodes,
                                               unsigned long long* segs, unsigned long long* rays) {
  unsigned long long n = c.nodes, s = c.segs, r = c.rays;
  for (int off = 32; off > 0; off >>= 1) {
    n += __shfl_xor(n, off);
    s += __shfl_xor(s, off);
    r += __shfl_xor(r, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(nodes, n);
    atomicAdd(segs, s);
    atomicAdd(rays, r);
  }
}

}  // namespace

__global__ void k_reset_counts(int* counts, int a, int b, int c) {
  if (threadIdx.x == 0) {
    counts[a] = 0;
    counts[b] = 0;
    counts[c] = 0;
  }
}

// Primary rays for every pixel of pass `accum_id` (path_tracing.cu:29-48).
// HairMSNN training pixels (hair_msnn.cu:206-209) are flagged here.
__global__ void k_camera(SceneView sc, PathBuffers pb, int accum_id, int msnn, const int* train_idx,
                         int every_nth, int n_train) {
  const int n = sc.width * sc.height;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[0] = n;
  if (i >= n) return;
  const int x = i % sc.width, y = i / sc.width;
  uint32_t rng = path_seed((uint32_t)(accum_id + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)sc.width);
  const float jx = draw(rng), jy = draw(rng);
  const float sx = ((float)x + jx) / (float)sc.width;
  const float sy = ((float)y + jy) / (float)sc.height;
  const V3 d = normalize(sc.cam_d00 + sx * sc.cam_du + sy * sc.cam_dv);
  pb.ray_o[i] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[i] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[i] = rng;
  pb.queue[0][i] = i;
  uint32_t flags = 0;
  if (msnn) {
    const int tr = i / every_nth;
    if (tr < n_train && i % every_nth == train_idx[tr] % every_nth) flags |= 1u;
    pb.beta_short[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    pb.L_short[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  pb.flags[i] = flags;
}

// HairMSNN training pixels of several frames as one pass: slot i holds
// record j = i % chunk of frame accum_id + i / chunk, whose pixel is
// j * every_nth + train_idx[i] % every_nth (the pixel k_camera flags,
// hair_msnn.cu:206-209), traced from that pixel's seed and jitter, so the
// paths are those the frame's own pass would trace. pix receives the pixels.
__global__ void k_camera_pixels(SceneView sc, PathBuffers pb, int accum_id, int n, int chunk, int every_nth,
                                const int* train_idx, int* pix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[0] = n;
  if (i >= n) return;
  const int c = i / chunk, j = i - c * chunk;
  const int p = j * every_nth + train_idx[i] % every_nth;
  pix[i] = p;
  const int x = p % sc.width, y = p / sc.width;
  uint32_t rng = path_seed((uint32_t)(accum_id + c + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)sc.width);
  const float jx = draw(rng), jy = draw(rng);
  const float sx = ((float)x + jx) / (float)sc.width;
  const float sy = ((float)y + jy) / (float)sc.height;
  const V3 d = normalize(sc.cam_d00 + sx * sc.cam_du + sy * sc.cam_dv);
  pb.ray_o[i] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[i] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.beta_short[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L_short[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[i] = rng;
  pb.queue[0][i] = i;
  pb.flags[i] = 1u;
}

// Training-data rays: camera position toward sampled hair points; every slot
// is a training path (cuda/hair_msnn.cu:222-233). The slots hold consecutive
// chunks of `chunk` paths, chunk c being the pass of index accum_id + c; a
// path's seed uses its launch index within its chunk's grid_w-wide 2D launch,
// so a batch of chunks reproduces the separate passes exactly.
__global__ void k_camera_points(SceneView sc, PathBuffers pb, int accum_id, int n, int chunk, int grid_w,
                                const float* pts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[0] = n;
  if (i >= n) return;
  const int c = i / chunk, j = i - c * chunk;
  const int x = j % grid_w, y = j / grid_w;
  const uint32_t rng = path_seed((uint32_t)(accum_id + c + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)grid_w);
  const V3 p = v3(pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]);
  const V3 d = normalize(p - sc.cam_pos);
  pb.ray_o[i] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[i] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.beta_short[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L_short[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[i] = rng;
  pb.queue[0][i] = i;
  pb.flags[i] = 1u;
}

template <bool kCount>
__global__ void HM_TRACE_BOUNDS k_extend(SceneView sc, PathBuffers pb, int cur,
                                                        unsigned long long* stats) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st = make_stack(lds_stack, pb);
  TraceCounters cnt{0, 0, 0};
  const int n = pb.counts[cur];
  const int* q = pb.queue[cur];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int slot = q[i];
    const float4 o = pb.ray_o[slot], d = pb.ray_d[slot];
    float t, u;
    int prim;
    if (kCount) cnt.rays++;
    closest_hit<kCount>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, t, prim, u, cnt);
    pb.hit_t[slot] = t;
    pb.hit_prim[slot] = prim;
    pb.hit_u[slot] = u;
  }
  if (kCount) flush_counters(cnt, &stats[0], &stats[1], &stats[4]);
}

__global__ void __launch_bounds__(256) k_shade(SceneView sc, PathBuffers pb, ShadeParams sp,
                                               GBufferOut gb) {
  const int nxt = sp.cur ^ 1;
  const int n = pb.counts[sp.cur];
  const int* q = pb.queue[sp.cur];
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (n + stride - 1) / stride * stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    bool cont = false, shadow = false, ray1 = false, ray2 = false;
    int slot = -1;
    if (i < n) {
      slot = q[i];
      const int prim = pb.hit_prim[slot];
      const float4 o4 = pb.ray_o[slot], d4 = pb.ray_d[slot];
      const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
      const bool training = sp.msnn && (pb.flags[slot] & 1u);
      if (prim == kPrimMiss) {
        if (sp.depth == 0) {
          // camera ray escaped: the miss program's radiance (optix_common.cuh:604-616)
          V3 le = v3(0.f);
          if (sc.env.enabled) le = env_radiance(sc.env, d);
          pb.L[slot] = make_float4(le.x, le.y, le.z, 0.f);
          if (gb.enabled) gbuffer_write_miss(gb, slot, d);
        }
      } else {
        Hit h;
        setup_hit(sc, o, d, pb.hit_t[slot], prim, pb.hit_u[slot], h);
        if (sp.depth == 0 && gb.enabled) gbuffer_write_hit(gb, sc, slot, h, prim);
        uint32_t rng = pb.rng[slot];
        float4 b4 = pb.beta[slot];
        V3 beta = v3(b4.x, b4.y, b4.z);
        V3 beta_s = v3(0.f);
        if (training) {
          const float4 s4 = pb.beta_short[slot];
          beta_s = v3(s4.x, s4.y, s4.z);
        }
        // ---- direct lighting at this vertex ----
        if (sp.depth == 0 ? sp.dl_zero != 0 : sp.depth >= sp.dl_from) {
          DirectOut dl;
          direct_lighting(sc, h, rng, dl);
          if (dl.has1 || dl.has2) {
            shadow = true;
            ray1 = dl.has1 != 0;
            ray2 = dl.has2 != 0;
            pb.sh_o1[slot] = make_float4(dl.o1.x, dl.o1.y, dl.o1.z, dl.has1 ? 1.f : 0.f);
            pb.sh_d1[slot] = make_float4(dl.d1.x, dl.d1.y, dl.d1.z, 0.f);
            pb.sh_c1[slot] = make_float4(dl.c1.x, dl.c1.y, dl.c1.z, 0.f);
            pb.sh_o2[slot] = make_float4(dl.o2.x, dl.o2.y, dl.o2.z, dl.has2 ? 1.f : 0.f);
            pb.sh_d2[slot] = make_float4(dl.d2.x, dl.d2.y, dl.d2.z, 0.f);
            pb.sh_c2[slot] = make_float4(dl.c2.x, dl.c2.y, dl.c2.z, 0.f);
            pb.sh_beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
            if (training) pb.sh_beta_short[slot] = make_float4(beta_s.x, beta_s.y, beta_s.z, 0.f);
          }
        }
        // ---- Russian roulette (optix_common.cuh:398-405; hair_msnn.cu:110-128, 171-181) ----
        bool alive = true;
        if (sp.depth >= 1) {
          const float qq = fmaxf(0.05f, 1.f - luminance(beta));
          const float eps = draw(rng);
          if (training) {
            const float qs = fmaxf(0.05f, 1.f - luminance(beta_s));
            if (eps < qs || sp.depth > sp.trunc_beta) beta_s = v3(0.f);
            if (eps < qq) {
              alive = false;
            } else {
              beta = beta / (1.f - qq);
              if (!is_zero(beta_s)) beta_s = beta_s / (1.f - qs);
            }
          } else {
            if (eps < qq || (sp.msnn && sp.depth > sp.trunc_beta)) alive = false;
            else beta = beta / (1.f - qq);
          }
        }
        // ---- continuation (nextPathVertex, optix_common.cuh:311-360) ----
        if (alive && sp.depth + 1 <= sp.max_vertex) {
          const BsdfSample bs = sample_bsdf(sc, h, rng);
          const V3 wl = bs.wl, wi = bs.wi, f = bs.f;
          const float pdf = bs.pdf;
          V3 mul = (pdf == 0.f) ? f : f / pdf;
          if (any_nan(mul)) mul = v3(1.f);
          beta = beta * mul;
          if (training) beta_s = beta_s * mul;
          V3 p = h.p, nd = h.n;
          if (dot(wi, h.n) < 0.f && !h.surface) {
            nd = -1.f * h.n;
            p = p + (2.f * h.radius) * nd;
          }
          const V3 no = p + 1e-3f * nd;
          pb.ray_o[slot] = make_float4(no.x, no.y, no.z, 0.f);
          pb.ray_d[slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
          pb.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
          if (training) pb.beta_short[slot] = make_float4(beta_s.x, beta_s.y, beta_s.z, 0.f);
          cont = true;
        }
        pb.rng[slot] = rng;
      }
    }
    const int qpos = wave_append(&pb.counts[nxt], cont);
    if (cont) pb.queue[nxt][qpos] = slot;
    const int spos = wave_append(&pb.counts[2], shadow);
    if (shadow) pb.shadow_queue[spos] = slot;
    // one entry per shadow ray (slot * 2 + which), so k_shadow traces one ray per lane
    const int r1 = wave_append(&pb.counts[3], ray1);
    if (ray1) pb.shadow_rays[r1] = 2 * slot;
    const int r2 = wave_append(&pb.counts[3], ray2);
    if (ray2) pb.shadow_rays[r2] = 2 * slot + 1;
  }
}

// Any-hit test of each pending shadow ray, one ray per lane; writes the
// visibility byte of (slot, which).
template <bool kCount>
__global__ void HM_TRACE_BOUNDS k_shadow(SceneView sc, PathBuffers pb, unsigned long long* stats) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st = make_stack(lds_stack, pb);
  TraceCounters cnt{0, 0, 0};
  const int n = pb.counts[3];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int e = pb.shadow_rays[i];
    const int slot = e >> 1;
    const float4 o = (e & 1) ? pb.sh_o2[slot] : pb.sh_o1[slot];
    const float4 d = (e & 1) ? pb.sh_d2[slot] : pb.sh_d1[slot];
    if (kCount) cnt.rays++;
    pb.sh_vis[e] = occluded<kCount>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, cnt) ? 0 : 1;
  }
  if (kCount) flush_counters(cnt, &stats[2], &stats[3], &stats[5]);
}

// Adds each path's visible direct light (directLighting's sum with its NaN
// scrub, optix_common.cuh:279-309) times the throughput at the vertex.
__global__ void k_shadow_resolve(PathBuffers pb, int msnn) {
  const int n = pb.counts[2];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int slot = pb.shadow_queue[i];
    V3 sum = v3(0.f);
    if (pb.sh_o1[slot].w != 0.f && pb.sh_vis[2 * slot]) {
      const float4 c = pb.sh_c1[slot];
      sum = sum + v3(c.x, c.y, c.z);
    }
    if (pb.sh_o2[slot].w != 0.f && pb.sh_vis[2 * slot + 1]) {
      const float4 c = pb.sh_c2[slot];
      sum = sum + v3(c.x, c.y, c.z);
    }
    if (any_nan(sum)) sum = v3(0.f);  // optix_common.cuh:305-306
    const float4 b = pb.sh_beta[slot];
    float4 L = pb.L[slot];
    L.x += b.x * sum.x;
    L.y += b.y * sum.y;
    L.z += b.z * sum.z;
    pb.L[slot] = L;
    if (msnn && (pb.flags[slot] & 1u)) {
      const float4 bs = pb.sh_beta_short[slot];
      float4 Ls = pb.L_short[slot];
      Ls.x += bs.x * sum.x;
      Ls.y += bs.y * sum.y;
      Ls.z += bs.z * sum.z;
      pb.L_short[slot] = Ls;
    }
  }
}

// Running sum and mean per pixel, NaN scrubbed (utils.cuh:13-35).
__global__ void k_accumulate(const float4* L, float4* accum, float4* average, int n, int accum_id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 l = L[i];
  V3 c = v3(l.x, l.y, l.z);
  if (any_nan(c)) c = v3(0.f);
  if (accum_id > 0) {
    const float4 a = accum[i];
    c = c + v3(a.x, a.y, a.z);
  }
  accum[i] = make_float4(c.x, c.y, c.z, 1.f);
  const V3 m = (1.f / (float)(accum_id + 1)) * c;
  average[i] = make_float4(m.x, m.y, m.z, 1.f);
}

// Closest or any hit for an explicit ray list (8 floats per ray).
template <bool kCount>
__global__ void HM_TRACE_BOUNDS k_trace_rays(SceneView sc, const float4* rays, int n, int any_hit,
                                                            int* prim_out, float* t_out, float* u_out, int* spill,
                                                            int spill_stride, int spill_levels,
                                                            unsigned long long* stats) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st;
  st.lds = lds_stack + threadIdx.x;
  st.spill = spill + blockIdx.x * blockDim.x + threadIdx.x;
  st.spill_stride = spill_stride;
  st.spill_levels = spill_levels;
  st.sp = 0;
  TraceCounters cnt{0, 0, 0};
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 o = rays[2 * (size_t)i], d = rays[2 * (size_t)i + 1];
    if (kCount) cnt.rays++;
    if (any_hit) {
      const bool occ = occluded<kCount>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, cnt);
      prim_out[i] = occ ? 1 : 0;
    } else {
      float t, u;
      int prim;
      closest_hit<kCount>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, t, prim, u, cnt);
      prim_out[i] = prim;
      if (t_out) t_out[i] = t;
      if (u_out) u_out[i] = u;
    }
  }
  if (kCount) flush_counters(cnt, &stats[0], &stats[1], &stats[4]);
}

// ---------------------------------------------------------------------------
// Host launchers

int pt_run_pass(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, hipStream_t stream,
                std::string& err) {
  if (cfg.camera_mode == 2) {
    const int n = cfg.n_slots;
    hipLaunchKernelGGL(k_camera_pixels, dim3((n + 255) / 256), dim3(256), 0, stream, sc, pb,
                       cfg.accum_id + cfg.seed_offset, n, cfg.chunk > 0 ? cfg.chunk : n, cfg.every_nth, cfg.train_idx,
                       cfg.pix);
  } else if (cfg.camera_mode == 1) {
    const int n = cfg.n_slots;
    hipLaunchKernelGGL(k_camera_points, dim3((n + 255) / 256), dim3(256), 0, stream, sc, pb,
                       cfg.accum_id + cfg.seed_offset, n, cfg.chunk > 0 ? cfg.chunk : n, cfg.grid_w, cfg.points);
  } else {
    const int n = sc.width * sc.height;
    hipLaunchKernelGGL(k_camera, dim3((n + 255) / 256), dim3(256), 0, stream, sc, pb, cfg.accum_id + cfg.seed_offset,
                       cfg.msnn, cfg.train_idx, cfg.every_nth, cfg.n_slots);
  }
  HM_TRY(check_launch("camera", err));
  const bool count = cfg.stats != nullptr;
  PassTimer* tm = cfg.timer;
  auto tic = [&](int kind) {
    if (tm && tm->used < tm->capacity) {
      tm->kinds[tm->used] = kind;
      (void)hipEventRecord(tm->events[2 * tm->used], stream);
    }
  };
  auto toc = [&]() {
    if (tm && tm->used < tm->capacity) {
      (void)hipEventRecord(tm->events[2 * tm->used + 1], stream);
      tm->used++;
    }
  };
  // Grids sized to the pass's paths (queues only shrink with depth), so a
  // small training-data pass does not dispatch the full-frame grid per depth.
  const int n0 = cfg.camera_mode != 0 ? cfg.n_slots : sc.width * sc.height;
  const int ext_blocks = std::min(cfg.trace_blocks, std::max(1, (n0 + kTraceBlock - 1) / kTraceBlock));
  const int sh_blocks = std::min(cfg.trace_blocks, std::max(1, (2 * n0 + kTraceBlock - 1) / kTraceBlock));
  const int shade_blocks = std::min(cfg.shade_blocks, std::max(1, (n0 + 255) / 256));
  int cur = 0;
  for (int depth = 0; depth <= cfg.max_vertex; ++depth) {
    const int nxt = cur ^ 1;
    hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, stream, pb.counts, nxt, 2, 3);
    tic(0);
    if (count)
      hipLaunchKernelGGL(k_extend<true>, dim3(ext_blocks), dim3(kTraceBlock), 0, stream, sc, pb, cur,
                         cfg.stats);
    else
      hipLaunchKernelGGL(k_extend<false>, dim3(ext_blocks), dim3(kTraceBlock), 0, stream, sc, pb, cur,
                         cfg.stats);
    toc();
    ShadeParams sp;
    sp.depth = depth;
    sp.dl_zero = cfg.dl_zero;
    sp.dl_from = cfg.dl_from;
    sp.max_vertex = cfg.max_vertex;
    sp.trunc_beta = cfg.trunc_beta;
    sp.msnn = cfg.msnn;
    sp.cur = cur;
    tic(1);
    hipLaunchKernelGGL(k_shade, dim3(shade_blocks), dim3(256), 0, stream, sc, pb, sp, cfg.gb);
    toc();
    tic(2);
    if (count)
      hipLaunchKernelGGL(k_shadow<true>, dim3(sh_blocks), dim3(kTraceBlock), 0, stream, sc, pb, cfg.stats);
    else
      hipLaunchKernelGGL(k_shadow<false>, dim3(sh_blocks), dim3(kTraceBlock), 0, stream, sc, pb, cfg.stats);
    hipLaunchKernelGGL(k_shadow_resolve, dim3(shade_blocks), dim3(256), 0, stream, pb, cfg.msnn);
    toc();
    HM_TRY(check_launch("wavefront vertex", err));
    cur = nxt;
  }
  return kOk;
}

int pt_accumulate(const float4* L, float4* accum, float4* average, int n, int accum_id, hipStream_t stream,
                  std::string& err) {
  hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, stream, L, accum, average, n, accum_id);
  return check_launch("accumulate", err);
}

int trace_rays(const SceneView& sc, const float* rays, int n, int any_hit, int* prim, float* t, float* u,
               int* spill, int spill_stride, int spill_levels, int blocks, unsigned long long* stats,
               hipStream_t stream, std::string& err) {
  if (n <= 0) return kOk;
  if (stats)
    hipLaunchKernelGGL(k_trace_rays<true>, dim3(blocks), dim3(kTraceBlock), 0, stream, sc,
                       reinterpret_cast<const float4*>(rays), n, any_hit, prim, t, u, spill, spill_stride, spill_levels,
                       stats);
  else
    hipLaunchKernelGGL(k_trace_rays<false>, dim3(blocks), dim3(kTraceBlock), 0, stream, sc,
                       reinterpret_cast<const float4*>(rays), n, any_hit, prim, t, u, spill, spill_stride, spill_levels,
                       stats);
  return check_launch("trace rays", err);
}

}  // namespace hm
