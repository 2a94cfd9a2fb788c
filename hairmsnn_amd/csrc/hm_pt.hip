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
// Queues are compacted with ballots and popcounts and one atomic per block and
// queue (shading) or per wave (the rest), each queue length in a cache line of
// its own. Kernels read the live count from device memory, so a pass needs no
// host round trip.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hm_pt.h"
#include "hm_pt_kernels.h"

namespace hm {

namespace {

// Appends to a queue from a 256-thread block: each lane with pred set gets a
// position. The four waves' counts meet in LDS and one thread makes the
// block's single atomic (one atomic per wave, every wave of a launch on the
// same counter, serialised there). Every thread of the block calls it the
// same number of times.
__device__ __forceinline__ int block_append1(int* counter, bool pred) {
  __shared__ int tot[4], base[4];
  const unsigned long long mask = __ballot(pred);
  const int w = threadIdx.x >> 6, lane = __lane_id();
  if (lane == 0) tot[w] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int sum = tot[0] + tot[1] + tot[2] + tot[3];
    int b = sum > 0 ? atomicAdd(counter, sum) : 0;
    for (int k = 0; k < 4; ++k) {
      base[k] = b;
      b += tot[k];
    }
  }
  __syncthreads();
  const int q = pred ? base[w] + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
  __syncthreads();  // base[] is read before the next call overwrites it
  return q;
}

// Queue positions of one vertex after a shading step.
struct Appended {
  int next, done, shadow, ray1, ray2;
};

// The appends of a 256-thread block (four waves) to the queues a shading
// step fills, with one atomic per queue and block: each wave's counts go to
// LDS, thread q adds queue q's block total to its counter and hands each wave
// its base. The shadow-ray queue takes a wave's ray1 entries, then its ray2
// entries. Every thread of the block calls this (the grid-stride loops here
// run the same number of iterations in every thread of a block); a null
// counter must receive no entries.
__device__ __forceinline__ Appended block_append(int* next_q, int* done_q, int* shadow_q, int* ray_q, bool next,
                                                 bool done, bool shadow, bool ray1, bool ray2) {
  __shared__ int tot[4][4], base[4][4];
  const unsigned long long mn = __ballot(next), md = __ballot(done), ms = __ballot(shadow), m1 = __ballot(ray1),
                           m2 = __ballot(ray2);
  const int w = threadIdx.x >> 6, lane = __lane_id();
  if (lane == 0) {
    tot[w][0] = __popcll(mn);
    tot[w][1] = __popcll(md);
    tot[w][2] = __popcll(ms);
    tot[w][3] = __popcll(m1) + __popcll(m2);
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    const int sum = tot[0][q] + tot[1][q] + tot[2][q] + tot[3][q];
    int* c = q == 0 ? next_q : q == 1 ? done_q : q == 2 ? shadow_q : ray_q;
    int b = sum > 0 ? atomicAdd(c, sum) : 0;
    for (int k = 0; k < 4; ++k) {
      base[k][q] = b;
      b += tot[k][q];
    }
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  Appended a;
  a.next = next ? base[w][0] + __popcll(mn & below) : -1;
  a.done = done ? base[w][1] + __popcll(md & below) : -1;
  a.shadow = shadow ? base[w][2] + __popcll(ms & below) : -1;
  a.ray1 = ray1 ? base[w][3] + __popcll(m1 & below) : -1;
  a.ray2 = ray2 ? base[w][3] + __popcll(m1) + __popcll(m2 & below) : -1;
  return a;
}

// Launches a kernel; with a timer that has room, as record `kind` whose
// events are taken from the dispatch itself (the kernel's start and end, no
// marker packets around it).
template <typename K, typename... Args>
void launch_timed(PassTimer* tm, int kind, K kernel, dim3 grid, dim3 block, hipStream_t st, Args... args) {
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (tm && tm->used < tm->capacity) {
    tm->kinds[tm->used] = kind;
    t0 = tm->events[2 * tm->used];
    t1 = tm->events[2 * tm->used + 1];
    tm->used++;
  }
  hipExtLaunchKernelGGL(kernel, grid, block, 0, st, t0, t1, 0, args...);
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

// Camera ray of pixel px in the pass whose seed index is seed_pass
// (path_tracing.cu:29-48), into path slot i: two jitter draws, then the path
// state reset.
__device__ __forceinline__ void camera_path(const SceneView& sc, const PathBuffers& pb, int i, int px,
                                            int seed_pass) {
  const int x = px % sc.width, y = px / sc.width;
  uint32_t rng = path_seed((uint32_t)(seed_pass + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)sc.width);
  const float jx = draw(rng), jy = draw(rng);
  const float sx = ((float)x + jx) / (float)sc.width;
  const float sy = ((float)y + jy) / (float)sc.height;
  const V3 d = normalize(sc.cam_d00 + sx * sc.cam_du + sy * sc.cam_dv);
  pb.ray_o[i] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[i] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[i] = rng;
}

__global__ void k_reset_counts(int* counts, int a, int b, int c) {
  if (threadIdx.x == 0) {
    counts[kCountStride * a] = 0;
    counts[kCountStride * b] = 0;
    counts[kCountStride * c] = 0;
    counts[kCountStride * kDoneCount] = 0;
  }
}

// Primary rays for every pixel of pass `accum_id` (path_tracing.cu:29-48).
// HairMSNN training pixels (hair_msnn.cu:206-209) are flagged here.
// NRC (nrc != 0): training pixels by the same rule (cuda/nrc.cu:371-377),
// every 16th of them, or all with all_unbiased, tracing an unbiased suffix.
__global__ void k_camera(SceneView sc, PathBuffers pb, int accum_id, int msnn, const int* train_idx,
                         int every_nth, int n_train, int nrc, int all_unbiased) {
  const int n = sc.width * sc.height;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[kCountStride * 0] = n;
  if (i >= n) return;
  camera_path(sc, pb, i, i, accum_id);
  pb.queue[0][i] = i;
  uint32_t flags = 0;
  if (msnn) {
    const int tr = i / every_nth;
    if (tr < n_train && i % every_nth == train_idx[tr] % every_nth) flags |= 1u;
    pb.beta_short[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    pb.L_short[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (nrc) {
    const int tr = i / every_nth;
    if (tr < n_train && i % every_nth == train_idx[tr] % every_nth) {
      flags |= kNrcTrain;
      if (tr % 16 == 0 || all_unbiased) flags |= kNrcUnbiased;
    }
  }
  pb.flags[i] = flags;
}

// HairMSNN training pixels of several frames as one pass: record i is
// record j = i % chunk of frame accum_id + i / chunk, whose pixel is
// j * every_nth + train_idx[i] % every_nth (the pixel k_camera flags,
// hair_msnn.cu:206-209), traced from that pixel's seed and jitter, so the
// paths are those the frame's own pass would trace. pix receives the pixels.
// The path goes to position q of pb.
__device__ __forceinline__ void camera_train_pixel(const SceneView& sc, const PathBuffers& pb, int q, int i,
                                                   int accum_id, int chunk, int every_nth, const int* train_idx,
                                                   int* pix) {
  const int c = i / chunk, j = i - c * chunk;
  const int p = j * every_nth + train_idx[i] % every_nth;
  pix[i] = p;
  const int x = p % sc.width, y = p / sc.width;
  uint32_t rng = path_seed((uint32_t)(accum_id + c + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)sc.width);
  const float jx = draw(rng), jy = draw(rng);
  const float sx = ((float)x + jx) / (float)sc.width;
  const float sy = ((float)y + jy) / (float)sc.height;
  const V3 d = normalize(sc.cam_d00 + sx * sc.cam_du + sy * sc.cam_dv);
  pb.ray_o[q] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[q] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[q] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.beta_short[q] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L_short[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[q] = rng;
}

// Training-data ray i: camera position toward sampled hair point i, a
// training path (cuda/hair_msnn.cu:222-233). Records come in consecutive
// chunks of `chunk`, chunk c being the pass of index accum_id + c; a path's
// seed uses its launch index within its chunk's grid_w-wide 2D launch, so a
// batch of chunks reproduces the separate passes exactly. The path goes to
// position q of pb.
__device__ __forceinline__ void camera_train_point(const SceneView& sc, const PathBuffers& pb, int q, int i,
                                                   int accum_id, int chunk, int grid_w, const float* pts) {
  const int c = i / chunk, j = i - c * chunk;
  const int x = j % grid_w, y = j / grid_w;
  const uint32_t rng = path_seed((uint32_t)(accum_id + c + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)grid_w);
  const V3 p = v3(pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]);
  const V3 d = normalize(p - sc.cam_pos);
  pb.ray_o[q] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.f);
  pb.ray_d[q] = make_float4(d.x, d.y, d.z, 0.f);
  pb.beta[q] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.beta_short[q] = make_float4(1.f, 1.f, 1.f, 0.f);
  pb.L_short[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  pb.rng[q] = rng;
}

// Slot i = training record i (camera_train_pixel), every slot a training path.
__global__ void k_camera_pixels(SceneView sc, PathBuffers pb, int accum_id, int n, int chunk, int every_nth,
                                const int* train_idx, int* pix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[kCountStride * 0] = n;
  if (i >= n) return;
  camera_train_pixel(sc, pb, i, i, accum_id, chunk, every_nth, train_idx, pix);
  pb.queue[0][i] = i;
  pb.flags[i] = 1u;
}

// Slot i = training-data ray i (camera_train_point), every slot a training path.
__global__ void k_camera_points(SceneView sc, PathBuffers pb, int accum_id, int n, int chunk, int grid_w,
                                const float* pts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) pb.counts[kCountStride * 0] = n;
  if (i >= n) return;
  camera_train_point(sc, pb, i, i, accum_id, chunk, grid_w, pts);
  pb.queue[0][i] = i;
  pb.flags[i] = 1u;
}

// NRC G_BUFFER vertex (nrcTracePaths, cuda/nrc.cu:135-343) in wavefront
// form. The reference's loop iteration b lights vertex b, samples and traces
// the continuation, and then decides on the new hit; here shading depth d
// first finishes iteration d - 1 with the hit just traced (spread, cache
// query, termination) and then runs the head of iteration d (training
// record, direct light, continuation). No Russian roulette. The radiance of
// a training pixel is frozen in L_short when its primary path queries the
// cache and its suffix goes on.
__global__ void __launch_bounds__(256) k_shade_nrc(SceneView sc, PathBuffers pb, int depth, int cur,
                                                   NrcPassArgs nr) {
  const int nxt = cur ^ 1;
  const int n = pb.counts[kCountStride * cur];
  const int* q = pb.queue[cur];
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (n + stride - 1) / stride * stride;
  const float kPiRef = 3.1415926f;  // Pi (utils.cuh:10)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    bool cont = false, shadow = false, ray1 = false, ray2 = false;
    int slot = -1;
    if (i < n) {
      slot = q[i];
      const int prim = pb.hit_prim[slot];
      const float4 o4 = pb.ray_o[slot], d4 = pb.ray_d[slot];
      const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
      uint32_t fl = pb.flags[slot];
      const bool training = fl & kNrcTrain;
      const int tr = slot / nr.every_nth;
      const bool miss = prim == kPrimMiss;
      Hit h;
      if (!miss) setup_hit(sc, o, d, pb.hit_t[slot], prim, pb.hit_u[slot], h);
      bool alive = true;
      float4 ns = make_float4(0.f, 0.f, nr.c, 0.f);  // a0, spread, c, pdf of the last continuation
      if (depth == 0) {
        if (miss || (h.surface && h.wo_local.z < 0.f)) {
          // G-buffer miss: the escaped ray's Le, or Le = 0 on a back face (nrc.cu:314-322)
          V3 le = v3(0.f);
          if (miss && sc.env.enabled) le = env_radiance(sc.env, d);
          pb.L[slot] = make_float4(le.x, le.y, le.z, 0.f);
          alive = false;
        } else {
          const float l = length(h.p - sc.cam_pos);
          ns.x = l * l / (4.f * kPiRef) / fabsf(h.wo_local.z);  // camera footprint a0 (nrc.cu:153)
          nr.gb_hit[slot] = 1;
        }
      } else {
        ns = nr.state[slot];
        const float4 pv = nr.prev[slot];
        const V3 prev = v3(pv.x, pv.y, pv.z);
        if (miss) {
          // the path left the scene (nrc.cu:209-220)
          if (training) {
            nr.tb_bounces[tr] = depth - 1;
            nr.tb_hit[tr] = 0;
          }
          alive = false;
        } else {
          const float l = length(h.p - prev);
          if (training && (fl & kNrcRecalcA0)) {
            ns.x = l * l / (4.f * kPiRef) / fabsf(h.wo_local.z);  // suffix footprint (nrc.cu:189-192)
            fl &= ~kNrcRecalcA0;
          }
          ns.y += sqrtf(l * l / ns.w / fabsf(h.wo_local.z));  // nrcSpread (utils.cuh:86-90)
          const bool spread_cond = ns.y * ns.y > ns.z * ns.x;
          if (spread_cond && !(fl & kNrcSuffix)) {
            // primary path ends in a cache query at this vertex (nrc.cu:222-271)
            const float4 b4 = pb.beta[slot];
            nr.gb_beta[slot] = b4;
            float* r = nr.query + (size_t)slot * kNnInputCh;
            const V3 p = h.p / nr.scene_scale;
            r[0] = p.x; r[1] = p.y; r[2] = p.z;
            r[3] = h.wo.x; r[4] = h.wo.y; r[5] = h.wo.z;
            r[6] = h.n.x; r[7] = h.n.y; r[8] = h.n.z;
            r[9] = 0.f; r[10] = 0.f; r[11] = 0.f;
            if (training) {
              fl |= kNrcSuffix | kNrcRecalcA0;
              ns.y = 0.f;
              pb.L_short[slot] = pb.L[slot];
            } else {
              alive = false;
            }
          } else if (spread_cond) {
            // training suffix ends in a cache query (nrc.cu:273-309)
            float* r = nr.query + ((size_t)nr.frame_size + tr) * kNnInputCh;
            const V3 p = h.p / nr.scene_scale;
            r[0] = p.x; r[1] = p.y; r[2] = p.z;
            r[3] = h.wo.x; r[4] = h.wo.y; r[5] = h.wo.z;
            r[6] = h.n.x; r[7] = h.n.y; r[8] = h.n.z;
            r[9] = 0.f; r[10] = 0.f; r[11] = 0.f;
            nr.tb_bounces[tr] = depth - 1;
            nr.tb_hit[tr] = 1;
            if (fl & kNrcUnbiased) ns.z = 1e30f;  // the unbiased suffix goes on to the end
            else alive = false;
          }
        }
      }
      if (alive) {
        uint32_t rng = pb.rng[slot];
        const float4 b4 = pb.beta[slot];
        V3 beta = v3(b4.x, b4.y, b4.z);
        const int rec = tr * kNrcMaxBounces + depth;
        if (training) {
          const V3 p = h.p / nr.scene_scale;
          nr.tb_pos[rec] = make_float4(p.x, p.y, p.z, 0.f);
          nr.tb_wo[rec] = make_float4(h.wo.x, h.wo.y, h.wo.z, 0.f);
          nr.tb_n[rec] = make_float4(h.n.x, h.n.y, h.n.z, 0.f);
        }
        ShadowSink sink(pb, slot);
        const bool last = depth == kNrcMaxBounces - 1;
        // direct light, then nextPathVertexNRC (nrc.cu:16-67) unless this is
        // the last bounce; NRC has no Russian roulette
        const Continuation c = shade_vertex(sc, h, rng, true, false, !last, [](float) { return true; },
                                            [&](int w, V3 lo, V3 ld, V3 lc) { sink(w, lo, ld, lc); });
        if (sink.finish(beta)) {
          shadow = true;
          ray1 = sink.has1;
          ray2 = sink.has2;
        }
        if (last) {
          // last bounce (nrc.cu:209-220)
          if (training) {
            nr.tb_bounces[tr] = depth;
            nr.tb_hit[tr] = 0;
          }
        } else {
          const V3 mul = c.mul;
          beta = beta * mul;
          if (training) nr.tb_beta[rec] = make_float4(mul.x, mul.y, mul.z, 0.f);
          ns.w = c.pdf;
          nr.prev[slot] = make_float4(h.p.x, h.p.y, h.p.z, 0.f);
          const V3 no = c.o;
          const V3 bs_wi = c.wi;
          pb.ray_o[slot] = make_float4(no.x, no.y, no.z, 0.f);
          pb.ray_d[slot] = make_float4(bs_wi.x, bs_wi.y, bs_wi.z, 0.f);
          pb.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
          cont = true;
        }
        pb.rng[slot] = rng;
        nr.state[slot] = ns;
      }
      pb.flags[slot] = fl;
    }
    const Appended a = block_append(&pb.counts[kCountStride * nxt], nullptr, &pb.counts[kCountStride * 2],
                                    &pb.counts[kCountStride * 3], cont, false, shadow, ray1, ray2);
    if (cont) pb.queue[nxt][a.next] = slot;
    if (shadow) pb.shadow_queue[a.shadow] = slot;
    if (ray1) pb.shadow_rays[a.ray1] = 2 * slot;
    if (ray2) pb.shadow_rays[a.ray2] = 2 * slot + 1;
  }
}

// kDense: the wavefront is compacted (path i is at index i), no queue.
template <bool kCount, bool kDense = false, bool kCurves = false>
__global__ void HM_TRACE_BOUNDS_T(kCurves) k_extend(SceneView sc, PathBuffers pb, int cur,
                                                        unsigned long long* stats) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st = make_stack(lds_stack, pb);
  TraceCounters cnt{0, 0, 0};
  const int n = pb.counts[kCountStride * cur];
  const int* q = pb.queue[cur];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int slot = kDense ? i : q[i];
    const float4 o = pb.ray_o[slot], d = pb.ray_d[slot];
    float t, u;
    int prim;
    if (kCount) cnt.rays++;
    closest_hit<kCount, kCurves>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, t, prim, u, cnt);
    pb.hit_t[slot] = t;
    pb.hit_prim[slot] = prim;
    pb.hit_u[slot] = u;
  }
  if (kCount) flush_counters(cnt, &stats[0], &stats[1], &stats[4]);
}

#ifndef HM_SHADE_MIN_WAVES
#define HM_SHADE_MIN_WAVES 3
#endif
__global__ void __launch_bounds__(256, HM_SHADE_MIN_WAVES) k_shade(SceneView sc, PathBuffers pb, ShadeParams sp,
                                                                   GBufferOut gb) {
  const int nxt = sp.cur ^ 1;
  const int n = pb.counts[kCountStride * (sp.cur)];
  const int* q = pb.queue[sp.cur];
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (n + stride - 1) / stride * stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    bool cont = false, shadow = false, ray1 = false, ray2 = false;
    int slot = -1;
    if (i < n) {
      slot = q[i];
      const int depth = sp.depth;
      const int prim = pb.hit_prim[slot];
      const float4 o4 = pb.ray_o[slot], d4 = pb.ray_d[slot];
      const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
      const bool training = sp.msnn && (pb.flags[slot] & 1u);
      if (prim == kPrimMiss) {
        if (depth == 0) {
          // camera ray escaped: the miss program's radiance (optix_common.cuh:604-616)
          V3 le = v3(0.f);
          if (sc.env.enabled) le = env_radiance(sc.env, d);
          pb.L[slot] = make_float4(le.x, le.y, le.z, 0.f);
          if (gb.enabled) gbuffer_write_miss(gb, slot, d);
        }
      } else {
        Hit h;
        setup_hit(sc, o, d, pb.hit_t[slot], prim, pb.hit_u[slot], h);
        if (depth == 0 && gb.enabled) gbuffer_write_hit(gb, sc, slot, h, prim);
        uint32_t rng = pb.rng[slot];
        float4 b4 = pb.beta[slot];
        V3 beta = v3(b4.x, b4.y, b4.z);
        V3 beta_s = v3(0.f);
        if (training) {
          const float4 s4 = pb.beta_short[slot];
          beta_s = v3(s4.x, s4.y, s4.z);
        }
        const V3 beta_dl = beta, beta_s_dl = beta_s;
        ShadowSink sink(pb, slot);
        // Russian roulette (optix_common.cuh:398-405; hair_msnn.cu:110-128, 171-181)
        auto roulette = [&](float eps) {
          const float qq = fmaxf(0.05f, 1.f - luminance(beta));
          if (training) {
            const float qs = fmaxf(0.05f, 1.f - luminance(beta_s));
            if (eps < qs || depth > sp.trunc_beta) beta_s = v3(0.f);
            if (eps < qq) return false;
            beta = beta / (1.f - qq);
            if (!is_zero(beta_s)) beta_s = beta_s / (1.f - qs);
            return true;
          }
          if (eps < qq || (sp.msnn && depth > sp.trunc_beta)) return false;
          beta = beta / (1.f - qq);
          return true;
        };
        const Continuation c = shade_vertex(sc, h, rng, depth == 0 ? sp.dl_zero != 0 : depth >= sp.dl_from,
                                            depth >= 1, depth + 1 <= sp.max_vertex, roulette,
                                            [&](int w, V3 lo, V3 ld, V3 lc) { sink(w, lo, ld, lc); });
        if (sink.finish(beta_dl)) {
          shadow = true;
          ray1 = sink.has1;
          ray2 = sink.has2;
          if (training) pb.sh_beta_short[slot] = make_float4(beta_s_dl.x, beta_s_dl.y, beta_s_dl.z, 0.f);
        }
        // continuation (nextPathVertex, optix_common.cuh:311-360)
        if (c.cont) {
          beta = beta * c.mul;
          if (training) beta_s = beta_s * c.mul;
          pb.ray_o[slot] = make_float4(c.o.x, c.o.y, c.o.z, 0.f);
          pb.ray_d[slot] = make_float4(c.wi.x, c.wi.y, c.wi.z, 0.f);
          pb.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
          if (training) pb.beta_short[slot] = make_float4(beta_s.x, beta_s.y, beta_s.z, 0.f);
          cont = true;
        }
        pb.rng[slot] = rng;
      }
    }
    const Appended a = block_append(&pb.counts[kCountStride * nxt], nullptr, &pb.counts[kCountStride * 2],
                                    &pb.counts[kCountStride * 3], cont, false, shadow, ray1, ray2);
    if (cont) pb.queue[nxt][a.next] = slot;
    if (shadow) pb.shadow_queue[a.shadow] = slot;
    // one entry per shadow ray (slot * 2 + which), so k_shadow traces one ray per lane
    if (ray1) pb.shadow_rays[a.ray1] = 2 * slot;
    if (ray2) pb.shadow_rays[a.ray2] = 2 * slot + 1;
  }
}

// Compacted wavefront vertex: path i of wavefront pb is shaded as k_shade
// shades a slot (the same operations in the same order), then a continuing
// path is written densely into wavefront nb and a finished one into the
// finished list; pb.dest[i] records where it went, for
// k_shadow_resolve_dense. Shadow records stay at index i. kMsnn: every path
// is a HairMSNN training path (full and short throughput, the short one cut
// after vertex trunc_beta); the primary hit's network input goes to gb at
// the path's job (slot_px), and the short radiance travels with the path.
template <bool kMsnn>
__device__ __forceinline__ void shade_dense(const SceneView& sc, const PathBuffers& pb, const PathBuffers& nb,
                                            const ShadeParams& sp, const GBufferOut& gb) {
  const int nxt = sp.cur ^ 1;
  const int n = pb.counts[kCountStride * (sp.cur)];
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (n + stride - 1) / stride * stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    bool cont = false, shadow = false, ray1 = false, ray2 = false;
    const bool live = i < n;
    float4 L = make_float4(0.f, 0.f, 0.f, 0.f), no4 = L, wi4 = L, beta4 = L;
    float4 Ls = L, bs4 = L;
    uint32_t rng = 0;
    int depth = 0;
    if (live) {
      depth = pb.slot_depth[i];
      const int prim = pb.hit_prim[i];
      const float4 o4 = pb.ray_o[i], d4 = pb.ray_d[i];
      const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
      L = pb.L[i];
      if constexpr (kMsnn) Ls = pb.L_short[i];
      if (prim == kPrimMiss) {
        if (depth == 0) {
          // camera ray escaped: the miss program's radiance (optix_common.cuh:604-616)
          V3 le = v3(0.f);
          if (sc.env.enabled) le = env_radiance(sc.env, d);
          L = make_float4(le.x, le.y, le.z, 0.f);
          if constexpr (kMsnn) gbuffer_write_miss(gb, pb.slot_px[i], d);
        }
      } else {
        Hit h;
        setup_hit(sc, o, d, pb.hit_t[i], prim, pb.hit_u[i], h);
        if constexpr (kMsnn)
          if (depth == 0) gbuffer_write_hit(gb, sc, pb.slot_px[i], h, prim);
        rng = pb.rng[i];
        const float4 b4 = pb.beta[i];
        V3 beta = v3(b4.x, b4.y, b4.z);
        V3 beta_s = v3(0.f);
        if constexpr (kMsnn) {
          const float4 s4 = pb.beta_short[i];
          beta_s = v3(s4.x, s4.y, s4.z);
        }
        const V3 beta_dl = beta, beta_s_dl = beta_s;
        ShadowSink sink(pb, i);
        // Russian roulette (optix_common.cuh:398-405; hair_msnn.cu:110-128,
        // 171-181) from the second vertex on
        auto roulette = [&](float eps) {
          const float qq = fmaxf(0.05f, 1.f - luminance(beta));
          if constexpr (kMsnn) {
            const float qs = fmaxf(0.05f, 1.f - luminance(beta_s));
            if (eps < qs || depth > sp.trunc_beta) beta_s = v3(0.f);
            if (eps < qq) return false;
            beta = beta / (1.f - qq);
            if (!is_zero(beta_s)) beta_s = beta_s / (1.f - qs);
            return true;
          } else {
            if (eps < qq) return false;
            beta = beta / (1.f - qq);
            return true;
          }
        };
        const Continuation c = shade_vertex(sc, h, rng, depth == 0 ? sp.dl_zero != 0 : depth >= sp.dl_from,
                                            depth >= 1, depth + 1 <= sp.max_vertex, roulette,
                                            [&](int w, V3 lo, V3 ld, V3 lc) { sink(w, lo, ld, lc); });
        if (sink.finish(beta_dl)) {
          shadow = true;
          ray1 = sink.has1;
          ray2 = sink.has2;
          if constexpr (kMsnn) pb.sh_beta_short[i] = make_float4(beta_s_dl.x, beta_s_dl.y, beta_s_dl.z, 0.f);
        }
        if (c.cont) {
          beta = beta * c.mul;
          no4 = make_float4(c.o.x, c.o.y, c.o.z, 0.f);
          wi4 = make_float4(c.wi.x, c.wi.y, c.wi.z, 0.f);
          beta4 = make_float4(beta.x, beta.y, beta.z, 0.f);
          if constexpr (kMsnn) {
            beta_s = beta_s * c.mul;
            bs4 = make_float4(beta_s.x, beta_s.y, beta_s.z, 0.f);
          }
          cont = true;
        }
      }
    }
    const bool done = live && !cont;
    const Appended a = block_append(&pb.counts[kCountStride * nxt], &pb.counts[kCountStride * kDoneCount],
                                    &pb.counts[kCountStride * 2], &pb.counts[kCountStride * 3], cont, done, shadow,
                                    ray1, ray2);
    const int p = a.next, dp = a.done;
    if (cont) {
      nb.ray_o[p] = no4;
      nb.ray_d[p] = wi4;
      nb.beta[p] = beta4;
      nb.L[p] = L;
      nb.rng[p] = rng;
      nb.slot_pass[p] = pb.slot_pass[i];
      nb.slot_px[p] = pb.slot_px[i];
      nb.slot_depth[p] = depth + 1;
      if constexpr (kMsnn) {
        nb.beta_short[p] = bs4;
        nb.L_short[p] = Ls;
      }
      pb.dest[i] = p;
    }
    if (done) {
      pb.done_L[dp] = L;
      pb.done_px[dp] = pb.slot_px[i];
      pb.done_pass[dp] = pb.slot_pass[i];
      if constexpr (kMsnn) pb.done_Ls[dp] = Ls;
      pb.dest[i] = ~dp;
    }
    if (shadow) pb.shadow_queue[a.shadow] = i;
    if (ray1) pb.shadow_rays[a.ray1] = 2 * i;
    if (ray2) pb.shadow_rays[a.ray2] = 2 * i + 1;
  }
}

__global__ void __launch_bounds__(256, HM_SHADE_MIN_WAVES) k_shade_dense(SceneView sc, PathBuffers pb,
                                                                          PathBuffers nb, ShadeParams sp) {
  shade_dense<false>(sc, pb, nb, sp, GBufferOut{});
}

__global__ void __launch_bounds__(256, HM_SHADE_MIN_WAVES) k_shade_dense_msnn(SceneView sc, PathBuffers pb,
                                                                               PathBuffers nb, ShadeParams sp,
                                                                               GBufferOut gb) {
  shade_dense<true>(sc, pb, nb, sp, gb);
}

// The resolve of k_shadow_resolve for a compacted wavefront: the visible
// contributions go to the path where k_shade_dense moved it (kMsnn: the
// short radiance too).
template <bool kMsnn>
__device__ __forceinline__ void shadow_resolve_dense(const PathBuffers& pb, const PathBuffers& nb) {
  const int n = pb.counts[kCountStride * 2];
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const int slot = pb.shadow_queue[j];
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
    const int dst = pb.dest[slot];
    float4* Lp = dst >= 0 ? nb.L + dst : pb.done_L + ~dst;
    float4 L = *Lp;
    L.x += b.x * sum.x;
    L.y += b.y * sum.y;
    L.z += b.z * sum.z;
    *Lp = L;
    if constexpr (kMsnn) {
      const float4 bs = pb.sh_beta_short[slot];
      float4* Lsp = dst >= 0 ? nb.L_short + dst : pb.done_Ls + ~dst;
      float4 Ls = *Lsp;
      Ls.x += bs.x * sum.x;
      Ls.y += bs.y * sum.y;
      Ls.z += bs.z * sum.z;
      *Lsp = Ls;
    }
  }
}

__global__ void k_shadow_resolve_dense(PathBuffers pb, PathBuffers nb) { shadow_resolve_dense<false>(pb, nb); }

__global__ void k_shadow_resolve_dense_msnn(PathBuffers pb, PathBuffers nb) { shadow_resolve_dense<true>(pb, nb); }

// Any-hit test of each pending shadow ray, one ray per lane; writes the
// visibility byte of (slot, which).
template <bool kCount, bool kCurves = false>
__global__ void HM_TRACE_BOUNDS_T(kCurves) k_shadow(SceneView sc, PathBuffers pb, unsigned long long* stats) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st = make_stack(lds_stack, pb);
  TraceCounters cnt{0, 0, 0};
  const int n = pb.counts[kCountStride * 3];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int e = pb.shadow_rays[i];
    const int slot = e >> 1;
    const float4 o = (e & 1) ? pb.sh_o2[slot] : pb.sh_o1[slot];
    const float4 d = (e & 1) ? pb.sh_d2[slot] : pb.sh_d1[slot];
    if (kCount) cnt.rays++;
    pb.sh_vis[e] = occluded<kCount, kCurves>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, cnt) ? 0 : 1;
  }
  if (kCount) flush_counters(cnt, &stats[2], &stats[3], &stats[5]);
}

// NRC training pixels also record the vertex's direct light (vertRadiance,
// cuda/nrc.cu:200).
__global__ void k_shadow_resolve(PathBuffers pb, int msnn, int depth, NrcPassArgs nr) {
  const int n = pb.counts[kCountStride * 2];
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
    if (nr.enabled && (pb.flags[slot] & kNrcTrain)) {
      const int rec = slot / nr.every_nth * kNrcMaxBounces + depth;
      nr.tb_rad[rec] = make_float4(sum.x, sum.y, sum.z, 0.f);
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
template <bool kCount, bool kCurves = false>
__global__ void HM_TRACE_BOUNDS_T(kCurves) k_trace_rays(SceneView sc, const float4* rays, int n, int any_hit,
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
      const bool occ = occluded<kCount, kCurves>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, cnt);
      prim_out[i] = occ ? 1 : 0;
    } else {
      float t, u;
      int prim;
      closest_hit<kCount, kCurves>(sc, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), st, t, prim, u, cnt);
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
                       cfg.msnn, cfg.train_idx, cfg.every_nth, cfg.nrc.enabled ? cfg.nrc.n_tpx : cfg.n_slots,
                       cfg.nrc.enabled, cfg.nrc.all_unbiased);
  }
  HM_TRY(check_launch("camera", err));
  const bool count = cfg.stats != nullptr;
  const bool curves = sc.curves != nullptr;
  PassTimer* tm = cfg.timer;

  // Grids sized to the pass's paths (queues only shrink with depth), so a
  // small training-data pass does not dispatch the full-frame grid per depth.
  const int n0 = cfg.camera_mode != 0 ? cfg.n_slots : sc.width * sc.height;
  const int ext_blocks = std::min(cfg.trace_blocks, std::max(1, (n0 + kTraceBlock - 1) / kTraceBlock));
  const int sh_blocks = std::min(cfg.trace_blocks, std::max(1, (2 * n0 + kTraceBlock - 1) / kTraceBlock));
  const int shade_blocks = std::min(cfg.shade_blocks, std::max(1, (n0 + 255) / 256));
  int cur = 0;
  // depth 0 always runs: a camera miss shows the environment even when
  // pathTrace itself returns black (path_tracing.cu:50-57)
  for (int depth = 0; depth <= std::max(0, cfg.max_vertex); ++depth) {
    const int nxt = cur ^ 1;
    hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, stream, pb.counts, nxt, 2, 3);
    const dim3 eg(ext_blocks), tb(kTraceBlock);
    if (curves && count)
      launch_timed(tm, 0, k_extend<true, false, true>, eg, tb, stream, sc, pb, cur, cfg.stats);
    else if (curves)
      launch_timed(tm, 0, k_extend<false, false, true>, eg, tb, stream, sc, pb, cur, cfg.stats);
    else if (count)
      launch_timed(tm, 0, k_extend<true>, eg, tb, stream, sc, pb, cur, cfg.stats);
    else
      launch_timed(tm, 0, k_extend<false>, eg, tb, stream, sc, pb, cur, cfg.stats);
    ShadeParams sp;
    sp.depth = depth;
    sp.dl_zero = cfg.dl_zero;
    sp.dl_from = cfg.dl_from;
    sp.max_vertex = cfg.max_vertex;
    sp.trunc_beta = cfg.trunc_beta;
    sp.msnn = cfg.msnn;
    sp.cur = cur;
    if (cfg.nrc.enabled)
      launch_timed(tm, 1, k_shade_nrc, dim3(shade_blocks), dim3(256), stream, sc, pb, depth, cur, cfg.nrc);
    else
      launch_timed(tm, 1, k_shade, dim3(shade_blocks), dim3(256), stream, sc, pb, sp, cfg.gb);
    const dim3 sg(sh_blocks);
    if (curves && count)
      launch_timed(tm, 2, k_shadow<true, true>, sg, tb, stream, sc, pb, cfg.stats);
    else if (curves)
      launch_timed(tm, 2, k_shadow<false, true>, sg, tb, stream, sc, pb, cfg.stats);
    else if (count)
      launch_timed(tm, 2, k_shadow<true>, sg, tb, stream, sc, pb, cfg.stats);
    else
      launch_timed(tm, 2, k_shadow<false>, sg, tb, stream, sc, pb, cfg.stats);
    launch_timed(tm, 3, k_shadow_resolve, dim3(shade_blocks), dim3(256), stream, pb, cfg.msnn, depth, cfg.nrc);
    HM_TRY(check_launch("wavefront vertex", err));
    cur = nxt;
  }
  return kOk;
}

// ---- streaming schedule ------------------------------------------------------
// A job queue over a compacted wavefront. A job is (pass, pixel): job j of a
// call is pass j / cnt of the group's pixel j % cnt, so jobs are handed out in
// pass order, and a pixel's passes can run at the same time on different
// paths. Each iteration: extend the current wavefront (path i at index i),
// shade it (continuing paths written densely into the other copy of the
// per-path state, finished ones into a finished list), trace and resolve the
// shadow rays, move the finished radiance into a ring of kStreamRing passes
// (indexed by the global pixel), plan, then add the passes whose jobs are all
// done to the image (in pass order, so the accumulation and the image are the
// per-pass schedule's) and append new jobs' camera rays after the continuing
// paths, up to the group's path capacity. Jobs of pass p start only while
// p < (accumulated passes) + kStreamRing, so the ring never overwrites a pass
// that is not accumulated yet. The pixels are split into K groups, each with
// its own path buffers and stream, so that one group's shading overlaps
// another's traversal: group g owns the 64-pixel chunks g, g + K, g + 2K, ...
// (balanced work, coherent camera rays within a wave).
// ctl (64-bit, per group): two copies of the scheduler state, each [0] next
// job, [1] passes accumulated, [2] free path capacity (set by the finishing
// kernel), [8 + r] finished jobs of ring entry r.
__global__ void k_stream_init(long long* ctl, int* counts, int slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kStreamCtl) ctl[i] = i == 2 ? slots : 0;
  if (i <= kDoneCount) counts[kCountStride * i] = 0;
}

// Pixel px's planned passes p0 .. p0 + k - 1 from the ring into the running
// sum and mean, in pass order (writePixel, utils.cuh:13-35).
__device__ __forceinline__ void accumulate_planned(const float4* ring, float4* accum, float4* average, long long p0,
                                                   int k, int first_pass, int npx, int px) {
  V3 c = v3(0.f);
  int pass = 0;
  for (int j = 0; j < k; ++j) {
    const long long p = p0 + j;
    pass = first_pass + (int)p;
    const float4 l = ring[(size_t)(p % kStreamRing) * npx + px];
    V3 v = v3(l.x, l.y, l.z);
    if (any_nan(v)) v = v3(0.f);
    if (pass > 0) {
      const float4 a = j == 0 ? accum[px] : make_float4(c.x, c.y, c.z, 1.f);
      v = v + v3(a.x, a.y, a.z);
    }
    c = v;
  }
  accum[px] = make_float4(c.x, c.y, c.z, 1.f);
  const float inv = 1.f / (float)(pass + 1);
  average[px] = make_float4(c.x * inv, c.y * inv, c.z * inv, 1.f);
}

// Finished paths from the finished list into the ring, counted per ring
// entry; the free capacity for the plan; the new jobs' camera rays appended
// to wavefront nb after the continuing paths.
__global__ void k_stream_finish_dense(PathBuffers pb, long long* ctl, float4* ring, int npx, int nxt, int slots) {
  // finished jobs per ring entry, counted in LDS; one global atomic per
  // block and entry (every block's waves adding to the same few counters
  // serialised on them)
  __shared__ int per_entry[kStreamRing];
  if (threadIdx.x < kStreamRing) per_entry[threadIdx.x] = 0;
  __syncthreads();
  const int n = pb.counts[kCountStride * kDoneCount];
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[2] = slots - pb.counts[kCountStride * nxt];  // free path capacity
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int r = pb.done_pass[i] % kStreamRing;
    ring[(size_t)r * npx + pb.done_px[i]] = pb.done_L[i];
    atomicAdd(&per_entry[r], 1);
  }
  __syncthreads();
  if (threadIdx.x < kStreamRing && per_entry[threadIdx.x] > 0)
    atomicAdd((unsigned long long*)&ctl[8 + threadIdx.x], (unsigned long long)per_entry[threadIdx.x]);
}

// Plans and runs the next step of a group from scheduler state S (written
// by the previous call and updated by the finishing kernel in between): every
// thread derives the same plan from S (passes whose jobs are all done, at
// most a ring's worth, are accumulated; free capacity x jobs inside the
// window are issued), one thread writes the state after it to T and clears
// the iteration's counters. Then the completed passes are added to the image
// and the new jobs' camera rays appended to wavefront nb.
__global__ void k_stream_next_dense(SceneView sc, PathBuffers nb, const long long* S, long long* T, const float4* ring,
                                    float4* accum, float4* average, int cur, int nxt, int first_pass, int seed_offset,
                                    int npx, int g, int K, int cnt, int passes) {
  const long long acc = S[1], job0 = S[0];
  int k = 0;
  while (acc + k < passes && k < kStreamRing && S[8 + (int)((acc + k) % kStreamRing)] == cnt) ++k;
  const long long limit = (long long)min((long long)passes, acc + k + kStreamRing) * cnt;
  const int n = (int)max(0ll, min(S[2], limit - job0));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    T[0] = job0 + n;
    T[1] = acc + k;
    for (int r = 0; r < kStreamRing; ++r) {
      const long long d = (r - acc % kStreamRing + kStreamRing) % kStreamRing;  // ring entry r holds pass acc + d
      T[8 + r] = d < k ? 0 : S[8 + r];
    }
    nb.counts[kCountStride * cur] = 0;
    nb.counts[kCountStride * 2] = 0;
    nb.counts[kCountStride * 3] = 0;
    nb.counts[kCountStride * kDoneCount] = 0;
  }
  const int work = max(k > 0 ? cnt : 0, n);
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (work + stride - 1) / stride * stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    if (k > 0 && i < cnt) accumulate_planned(ring, accum, average, acc, k, first_pass, npx, stream_group_pixel(i, g, K));
    const bool go = i < n;
    const int q = block_append1(&nb.counts[kCountStride * nxt], go);
    if (go) {
      const long long job = job0 + i;
      const int pass = (int)(job / cnt);
      const int px = stream_group_pixel((int)(job % cnt), g, K);
      camera_path(sc, nb, q, px, first_pass + pass + seed_offset);
      nb.slot_pass[q] = pass;
      nb.slot_px[q] = px;
      nb.slot_depth[q] = 0;
    }
  }
}

// The same buffers with the two per-path state copies exchanged.
static PathBuffers other_state(const PathBuffers& pb) {
  PathBuffers v = pb;
  v.ray_o = pb.alt.ray_o;
  v.ray_d = pb.alt.ray_d;
  v.beta = pb.alt.beta;
  v.L = pb.alt.L;
  v.rng = pb.alt.rng;
  v.slot_pass = pb.alt.pass;
  v.slot_px = pb.alt.px;
  v.slot_depth = pb.alt.depth;
  v.beta_short = pb.alt.beta_short;
  v.L_short = pb.alt.L_short;
  v.alt = PathState{pb.ray_o,     pb.ray_d,   pb.beta,       pb.L,         pb.rng,
                    pb.slot_pass, pb.slot_px, pb.slot_depth, pb.beta_short, pb.L_short};
  return v;
}

int stream_group_slots(int npx, int g, int K) {
  // chunks of 64 pixels dealt round-robin; a partial last chunk (scenes hold
  // a multiple of 128 pixels, so it does not occur from a scene file) is the
  // last chunk of its group, so the group's pixel indices stay contiguous
  const int chunks = (npx + 63) / 64;
  if (g >= chunks) return 0;
  int n = 64 * ((chunks - g + K - 1) / K);
  if ((chunks - 1) % K == g) n -= 64 * chunks - npx;
  return n;
}

int pt_run_stream(const SceneView& sc, PathBuffers* const* pbs, const hipStream_t* streams, int K,
                  const PassConfig& cfg, int passes, float4* accum, float4* average, float4* ring, long long* ctl,
                  int* live_host, hipEvent_t* live_ev, std::string& err) {
  const int npx = sc.width * sc.height;
  if (passes <= 0 || npx <= 0 || K < 1) return passes <= 0 ? kOk : (err = "bad stream groups", kEInval);
  PassTimer* tm = cfg.timer;

// Two paths in flight per pixel of a group (four groups: 2M paths at 1024²)
// measured 3% faster than one; four were 4% slower (a longer final drain).
#ifndef HM_STREAM_SLOT_FACTOR
#define HM_STREAM_SLOT_FACTOR 2
#endif
  constexpr int kBatch = 8;  // even: each group's live queue is back at index 0 after a batch
  std::vector<int> cur(K, 0), live(K, 0), calls(K, 0);
  std::vector<long long> iters(K, 0);
  // paths in flight per group: HM_STREAM_SLOT_FACTOR per pixel of the group,
  // at most the npx slots its path buffers hold (jobs are still the group's
  // pixels x passes)
  auto group_slots = [&](int cnt) { return std::min(npx, HM_STREAM_SLOT_FACTOR * cnt); };
  // first jobs: every slot free, then one issue round into queue 0
  for (int g = 0; g < K; ++g) {
    const int cnt = stream_group_slots(npx, g, K);
    live[g] = cnt > 0;
    if (!cnt) continue;
    long long* c = ctl + (size_t)g * kStreamCtl;
    const int slots = group_slots(cnt);
    const int blocks = (std::max(cnt, slots) + 255) / 256;
    hipLaunchKernelGGL(k_stream_init, dim3(1), dim3(kStreamCtl), 0, streams[g], c, pbs[g]->counts, slots);
    hipLaunchKernelGGL(k_stream_next_dense, dim3(blocks), dim3(256), 0, streams[g], sc, *pbs[g], c, c + kStreamHalf,
                       ring, accum, average, 1, 0, cfg.accum_id, cfg.seed_offset, npx, g, K, cnt, passes);
    calls[g] = 1;
  }
  HM_TRY(check_launch("stream start", err));
  // A path lives at most max_vertex + 1 iterations and every group holds at
  // least as many slots as pixels (slots >= cnt), so a group drains within
  // passes x (max_vertex + 1) iterations; the live count is read one batch
  // late, which adds up to two batches. Past twice that bound the schedule is
  // reported as stuck.
  for (int g = 0; g < K; ++g) {
    const int cnt = stream_group_slots(npx, g, K);
    if (cnt > 0 && group_slots(cnt) < cnt) {
      err = "stream group holds fewer slots than pixels";
      return kEInval;
    }
  }
  const long long max_iters = 2 * ((long long)passes * std::max(1, cfg.max_vertex + 1) + 2 * kBatch);
  for (int batch = 0;; ++batch) {
    bool any = false;
    for (int g = 0; g < K; ++g) {
      if (!live[g]) continue;
      any = true;
      PathBuffers& pb = *pbs[g];
      const PathBuffers views[2] = {pb, other_state(pb)};  // state copy 0 / 1 as the current wavefront
      hipStream_t st = streams[g];
      const int cnt = stream_group_slots(npx, g, K);
      const int slots = group_slots(cnt);
      const int ext_blocks = std::min(cfg.trace_blocks, std::max(1, (slots + kTraceBlock - 1) / kTraceBlock));
      const int sh_blocks = std::min(cfg.trace_blocks, std::max(1, (2 * slots + kTraceBlock - 1) / kTraceBlock));
      const int shade_blocks = std::min(cfg.shade_blocks, std::max(1, (std::max(cnt, slots) + 255) / 256));
      const bool timed = g == 0 && tm;
      for (int k = 0; k < kBatch; ++k) {
        const int nxt = cur[g] ^ 1;
        const PathBuffers& A = views[cur[g]];
        const PathBuffers& B = views[nxt];
        PassTimer* t = timed ? tm : nullptr;
        if (sc.curves)
          launch_timed(t, 0, k_extend<false, true, true>, dim3(ext_blocks), dim3(kTraceBlock), st, sc, A, cur[g],
                       (unsigned long long*)nullptr);
        else
          launch_timed(t, 0, k_extend<false, true>, dim3(ext_blocks), dim3(kTraceBlock), st, sc, A, cur[g],
                       (unsigned long long*)nullptr);
        ShadeParams sp;
        sp.depth = 0;
        sp.dl_zero = cfg.dl_zero;
        sp.dl_from = cfg.dl_from;
        sp.max_vertex = cfg.max_vertex;
        sp.trunc_beta = cfg.trunc_beta;
        sp.msnn = 0;
        sp.cur = cur[g];
        launch_timed(t, 1, k_shade_dense, dim3(shade_blocks), dim3(256), st, sc, A, B, sp);
        if (sc.curves)
          launch_timed(t, 2, k_shadow<false, true>, dim3(sh_blocks), dim3(kTraceBlock), st, sc, A,
                       (unsigned long long*)nullptr);
        else
          launch_timed(t, 2, k_shadow<false>, dim3(sh_blocks), dim3(kTraceBlock), st, sc, A,
                       (unsigned long long*)nullptr);
        launch_timed(t, 3, k_shadow_resolve_dense, dim3(shade_blocks), dim3(256), st, A, B);
        // scheduler state: next call m reads half m % 2 and writes the other
        long long* c = ctl + (size_t)g * kStreamCtl;
        long long* S = c + kStreamHalf * (calls[g] & 1);
        long long* T = c + kStreamHalf * ((calls[g] + 1) & 1);
        launch_timed(t, 3, k_stream_finish_dense, dim3(shade_blocks), dim3(256), st, A, S, ring, npx, nxt, slots);
        launch_timed(t, 3, k_stream_next_dense, dim3(shade_blocks), dim3(256), st, sc, B, S, T, ring, accum,
                     average, cur[g], nxt, cfg.accum_id, cfg.seed_offset, npx, g, K, cnt, passes);
        ++calls[g];
        cur[g] = nxt;
      }
      iters[g] += kBatch;
      HM_TRY(check_launch("stream iteration", err));
      const int s = 2 * g + (batch & 1);
      HM_HIP(hipMemcpyAsync(&live_host[s], pb.counts + kCountStride * cur[g], sizeof(int), hipMemcpyDeviceToHost, st),
             err);
      HM_HIP(hipEventRecord(live_ev[s], st), err);
    }
    if (!any) break;
    if (batch > 0) {
      // the previous batch's counts, read while this batch runs
      for (int g = 0; g < K; ++g) {
        if (!live[g]) continue;
        const int s = 2 * g + ((batch - 1) & 1);
        HM_HIP(hipEventSynchronize(live_ev[s]), err);
        if (live_host[s] == 0) {
          live[g] = 0;  // drained: never an error, however late
          continue;
        }
        if (iters[g] > max_iters) {
          err = "streaming schedule did not drain";
          return kEDevice;
        }
      }
    }
  }
  return kOk;
}

// ---- job schedule -----------------------------------------------------------
// Records of a camera_mode 1 / 2 pass on a compacted wavefront (the path
// kernels are those of the streaming schedule with HairMSNN training state):
// each iteration extends, shades, traces and resolves the current
// wavefront, moves the finished paths' radiance to their records' outputs
// (k_jobs_finish) and appends new records' camera rays after the continuing
// paths, up to the capacity (k_jobs_next). ctl holds two copies of the
// scheduler words; step m reads copy m % 2 and writes the other.

struct JobCamera {
  int mode;  // 1: rays toward sampled hair points, 2: training pixels
  int n_jobs, accum_id, chunk, grid_w, every_nth;
  const float* points;
  const int* train_idx;
  int* pix;
};

__global__ void k_jobs_init(long long* ctl, int* counts, int capacity) {
  const int i = threadIdx.x;
  if (i < kJobCtl) ctl[i] = i == 2 ? capacity : 0;
  if (i <= kDoneCount) counts[kCountStride * i] = 0;
}

// Finished paths' radiance to their records; the free capacity for the
// next issue; one more live iteration if the wavefront just shaded (cur)
// held paths.
__global__ void k_jobs_finish(PathBuffers pb, long long* S, long long* ctl, int cur, int nxt, int capacity,
                              float4* out_L, float4* out_Ls) {
  const int n = pb.counts[kCountStride * kDoneCount];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S[2] = capacity - pb.counts[kCountStride * nxt];
    if (pb.counts[kCountStride * cur] > 0) ctl[kJobCtlLive] += 1;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int job = pb.done_px[i];
    out_L[job] = pb.done_L[i];
    out_Ls[job] = pb.done_Ls[i];
  }
}

// Issues min(free capacity, records left) new records into wavefront nb
// after its continuing paths, and clears the counters the next iteration
// appends to. Every thread derives the same count from S.
__global__ void k_jobs_next(SceneView sc, PathBuffers nb, const long long* S, long long* T, int cur, int nxt,
                            JobCamera jc) {
  const long long job0 = S[0];
  const int n = (int)max(0ll, min(S[2], (long long)jc.n_jobs - job0));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    T[0] = job0 + n;
    nb.counts[kCountStride * cur] = 0;
    nb.counts[kCountStride * 2] = 0;
    nb.counts[kCountStride * 3] = 0;
    nb.counts[kCountStride * kDoneCount] = 0;
  }
  const int stride = gridDim.x * blockDim.x;
  const int nloop = (n + stride - 1) / stride * stride;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nloop; i += stride) {
    const bool go = i < n;
    const int q = block_append1(&nb.counts[kCountStride * nxt], go);
    if (go) {
      const int job = (int)(job0 + i);
      if (jc.mode == 1)
        camera_train_point(sc, nb, q, job, jc.accum_id, jc.chunk, jc.grid_w, jc.points);
      else
        camera_train_pixel(sc, nb, q, job, jc.accum_id, jc.chunk, jc.every_nth, jc.train_idx, jc.pix);
      nb.slot_pass[q] = 0;
      nb.slot_px[q] = job;
      nb.slot_depth[q] = 0;
    }
  }
}

static JobCamera job_camera(const PassConfig& cfg) {
  JobCamera jc{};
  jc.mode = cfg.camera_mode;
  jc.n_jobs = cfg.n_slots;
  jc.accum_id = cfg.accum_id + cfg.seed_offset;
  jc.chunk = cfg.chunk > 0 ? cfg.chunk : cfg.n_slots;
  jc.grid_w = cfg.grid_w;
  jc.every_nth = cfg.every_nth;
  jc.points = cfg.points;
  jc.train_idx = cfg.train_idx;
  jc.pix = cfg.pix;
  return jc;
}

static int job_blocks(int paths, int cap_blocks, int block) {
  return std::min(cap_blocks, std::max(1, (paths + block - 1) / block));
}

int pt_jobs_start(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, int capacity, JobRun& run,
                  hipStream_t stream, std::string& err) {
  if ((cfg.camera_mode != 1 && cfg.camera_mode != 2) || capacity < 1 || !pb.done_Ls || !pb.alt.L_short) {
    err = "job schedule: training-path records and HairMSNN path buffers required";
    return kEInval;
  }
  run = JobRun{};
  run.capacity = capacity;
  hipLaunchKernelGGL(k_jobs_init, dim3(1), dim3(64), 0, stream, pb.job_ctl, pb.counts, capacity);
  // first issue into wavefront 0 (state copy 0)
  hipLaunchKernelGGL(k_jobs_next, dim3(job_blocks(capacity, cfg.shade_blocks, 256)), dim3(256), 0, stream, sc, pb,
                     pb.job_ctl, pb.job_ctl + kJobCtlHalf, 1, 0, job_camera(cfg));
  run.calls = 1;
  run.cur = 0;
  return check_launch("job schedule start", err);
}

int pt_jobs_iterate(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, JobRun& run, int iters,
                    float4* out_L, float4* out_Ls, hipStream_t stream, std::string& err) {
  const PathBuffers views[2] = {pb, other_state(pb)};
  const int C = run.capacity;
  const int ext_blocks = job_blocks(C, cfg.trace_blocks, kTraceBlock);
  const int sh_blocks = job_blocks(2 * C, cfg.trace_blocks, kTraceBlock);
  const int shade_blocks = job_blocks(C, cfg.shade_blocks, 256);
  const JobCamera jc = job_camera(cfg);
  ShadeParams sp;
  sp.depth = 0;
  sp.dl_zero = cfg.dl_zero;
  sp.dl_from = cfg.dl_from;
  sp.max_vertex = cfg.max_vertex;
  sp.trunc_beta = cfg.trunc_beta;
  sp.msnn = 1;
  long long* c = pb.job_ctl;
  for (int k = 0; k < iters; ++k) {
    const int cur = run.cur, nxt = cur ^ 1;
    const PathBuffers& A = views[cur];
    const PathBuffers& B = views[nxt];
    sp.cur = cur;
    if (sc.curves)
      hipLaunchKernelGGL((k_extend<false, true, true>), dim3(ext_blocks), dim3(kTraceBlock), 0, stream, sc, A, cur,
                         (unsigned long long*)nullptr);
    else
      hipLaunchKernelGGL((k_extend<false, true>), dim3(ext_blocks), dim3(kTraceBlock), 0, stream, sc, A, cur,
                         (unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_shade_dense_msnn, dim3(shade_blocks), dim3(256), 0, stream, sc, A, B, sp, cfg.gb);
    if (sc.curves)
      hipLaunchKernelGGL((k_shadow<false, true>), dim3(sh_blocks), dim3(kTraceBlock), 0, stream, sc, A,
                         (unsigned long long*)nullptr);
    else
      hipLaunchKernelGGL((k_shadow<false>), dim3(sh_blocks), dim3(kTraceBlock), 0, stream, sc, A,
                         (unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_shadow_resolve_dense_msnn, dim3(shade_blocks), dim3(256), 0, stream, A, B);
    long long* S = c + kJobCtlHalf * (run.calls & 1);
    long long* T = c + kJobCtlHalf * ((run.calls + 1) & 1);
    hipLaunchKernelGGL(k_jobs_finish, dim3(shade_blocks), dim3(256), 0, stream, A, S, c, cur, nxt, C, out_L, out_Ls);
    hipLaunchKernelGGL(k_jobs_next, dim3(shade_blocks), dim3(256), 0, stream, sc, B, S, T, cur, nxt, jc);
    ++run.calls;
    ++run.iters;
    run.cur = nxt;
  }
  return check_launch("job schedule iteration", err);
}

int pt_jobs_status(const PathBuffers& pb, const JobRun& run, long long* host, hipEvent_t ev, hipStream_t stream,
                   std::string& err) {
  // the wavefront count is an int: it lands in the low half of host[0]
  // (little-endian), whose high half is cleared here
  host[0] = 0;
  HM_HIP(hipMemcpyAsync(host, pb.counts + kCountStride * run.cur, sizeof(int), hipMemcpyDeviceToHost, stream), err);
  HM_HIP(hipMemcpyAsync(host + 1, pb.job_ctl + kJobCtlLive, sizeof(long long), hipMemcpyDeviceToHost, stream), err);
  HM_HIP(hipEventRecord(ev, stream), err);
  return kOk;
}

int pt_jobs_drain(const SceneView& sc, PathBuffers& pb, const PassConfig& cfg, JobRun& run, float4* out_L,
                  float4* out_Ls, long long* host, const hipEvent_t* ev, hipStream_t stream, std::string& err,
                  int* live_iters) {
  constexpr int kBatch = 8;
  const int bound = pt_jobs_bound(cfg, run.capacity) + 2 * kBatch;
  for (int b = 0;; ++b) {
    HM_TRY(pt_jobs_iterate(sc, pb, cfg, run, kBatch, out_L, out_Ls, stream, err));
    HM_TRY(pt_jobs_status(pb, run, host + 2 * (b & 1), ev[b & 1], stream, err));
    if (b > 0) {
      const long long* h = host + 2 * ((b - 1) & 1);
      HM_HIP(hipEventSynchronize(ev[(b - 1) & 1]), err);
      if ((int)h[0] == 0) {  // drained one batch ago; this batch found nothing
        if (live_iters) *live_iters = (int)h[1];
        return kOk;
      }
    }
    if (run.iters > bound) {
      err = "job schedule did not drain";
      return kEDevice;
    }
  }
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
  const float4* r4 = reinterpret_cast<const float4*>(rays);
  const dim3 g(blocks), b(kTraceBlock);
  if (sc.curves && stats)
    launch_timed(nullptr, 0, k_trace_rays<true, true>, g, b, stream, sc, r4, n, any_hit, prim, t, u, spill,
                 spill_stride, spill_levels, stats);
  else if (sc.curves)
    launch_timed(nullptr, 0, k_trace_rays<false, true>, g, b, stream, sc, r4, n, any_hit, prim, t, u, spill,
                 spill_stride, spill_levels, stats);
  else if (stats)
    launch_timed(nullptr, 0, k_trace_rays<true>, g, b, stream, sc, r4, n, any_hit, prim, t, u, spill, spill_stride,
                 spill_levels, stats);
  else
    launch_timed(nullptr, 0, k_trace_rays<false>, g, b, stream, sc, r4, n, any_hit, prim, t, u, spill, spill_stride,
                 spill_levels, stats);
  return check_launch("trace rays", err);
}

}  // namespace hm
