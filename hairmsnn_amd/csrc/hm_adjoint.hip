// Path-replay adjoint of the path tracer for differentiable hair appearance:
// gradients of L(mean image) w.r.t. per-strand sigma_a and global beta_m,
// beta_n, alpha, given dL/d(image).
//
// Each (pixel, pass) path is traced twice from the same seed (the reference's
// per-pixel streams, cuda/path_tracing.cu:29): a primal pass that sums the
// path's radiance, then a replay that differentiates every factor that
// depends on the parameters at each hair vertex, with adj = dL/dI / spp:
//   direct-lighting sample k:  c_k = f_k * w_k * (parameter-free factor), so
//     dc_k = (c_k / f_k) * df_k + c_k * dw_k / w_k, where w_k is the
//     power-heuristic MIS weight through the BSDF pdf (1 for delta lights);
//   continuation: the radiance still to come after the vertex is
//     proportional to f, so it contributes rem * df / f.
// Sampled directions, the sampling pdfs in the estimators' denominators and
// Russian roulette are held fixed (detached): the estimator stays unbiased
// for the derivative of the expected image without storing a tape.
// Per-strand gradients are reduced with FP64 atomics.
#include <hip/hip_runtime.h>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"
#include "hm_context.h"
#include "hm_hair_ad.h"

namespace hm {

namespace {

struct AdjArgs {
  HairConstsAD hcd;
  const float4* dl_dimg;  // w*h
  float inv_spp;
  int pass0, npass;
  int dl_zero, dl_from, max_vertex;
  const float4* ring;  // primal radiance of pass pass0 + p at ring[p * npx + pixel]
  double* g_sigma;  // num_strands * 3
  double* g_glob;   // 3
  float* g_px;      // optional: w*h*3 per-pixel d/d(beta_m, beta_n, alpha)
  int* spill;
  int spill_stride, spill_levels;
};

HM_D int strand_of(const SceneView& sc, int prim) { return (int)sc.seg_by_id[8 * (size_t)prim + 7]; }

HM_D V3 strand_sigma_of(const SceneView& sc, int prim) {
  if (!sc.strand_sigma) return sc.sigma_a;
  const int s = strand_of(sc, prim);
  return v3(sc.strand_sigma[3 * s], sc.strand_sigma[3 * s + 1], sc.strand_sigma[3 * s + 2]);
}

// g[k] = sum_c wk_c * f_c'[k] + wc * r[k]; the global part goes to gg, the
// sigma part to the strand's FP64 slots.
HM_D void add_grad(const SceneView& sc, const AdjArgs& a, int prim, const Dual f[3], V3 wk, float wc, const float r[kAdParams],
                   float* gg) {
  float g[kAdParams];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < kAdParams; ++k) {
    g[k] = wk.x * f[0].d[k] + wk.y * f[1].d[k] + wk.z * f[2].d[k] + wc * r[k];
    ok = ok && isfinite(g[k]);
  }
  if (!ok) return;
  gg[0] += g[0];
  gg[1] += g[1];
  gg[2] += g[2];
  const int strand = strand_of(sc, prim);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if (g[3 + c] != 0.f) atomicAdd(&a.g_sigma[3 * (size_t)strand + c], (double)g[3 + c]);
}

HM_D V3 safe_ratio(V3 c, V3 f) {
  return v3(f.x != 0.f ? c.x / f.x : 0.f, f.y != 0.f ? c.y / f.y : 0.f, f.z != 0.f ? c.z / f.z : 0.f);
}

// A replayed path between vertices: the ray to trace next and what the
// vertices so far left (the operations and random draws are those of
// k_camera / k_shade / k_shadow for the path's seed).
struct AdjPath {
  uint32_t rng;
  V3 o, d;      // the ray to trace next
  V3 beta, L;   // throughput and radiance gathered so far
  V3 Lt, adj;   // the path's primal radiance and dLoss/dI / spp
  int depth;    // the vertex the next hit is
  int pix;
  float gp[3];  // the path's global gradient
};

// The camera ray of pixel (x, y) in pass `pass` (path_tracing.cu:29-48).
HM_D void adj_start(const SceneView& sc, int x, int y, int pass, AdjPath& q) {
  q.rng = path_seed((uint32_t)(pass + 10007), (uint32_t)x, (uint32_t)y, (uint32_t)sc.width);
  const float jx = draw(q.rng), jy = draw(q.rng);
  const float sx = ((float)x + jx) / (float)sc.width, sy = ((float)y + jy) / (float)sc.height;
  q.d = normalize(sc.cam_d00 + sx * sc.cam_du + sy * sc.cam_dv);
  q.o = sc.cam_pos;
  q.beta = v3(1.f);
  q.L = v3(0.f);
  q.depth = 0;
  q.gp[0] = q.gp[1] = q.gp[2] = 0.f;
}

// The vertex the path's ray hit (t, prim, u): its direct lighting with its
// two shadow rays, the gradient of every factor there that depends on the
// parameters, Russian roulette and the continuation. Returns whether the
// path goes on (q then holds the next ray). Given the path's primal
// radiance Lt, the continuation's gradient uses the radiance still to come,
// Lt - L.
HM_D bool adj_vertex(const SceneView& sc, const AdjArgs& a, TraceStack& st, float t, int prim, float u,
                     AdjPath& q) {
  TraceCounters cnt{0, 0, 0};
  const int depth = q.depth;
  Hit h;
  setup_hit(sc, q.o, q.d, t, prim, u, h);
  const bool hair = !h.surface;
  if (depth == 0 ? a.dl_zero != 0 : depth >= a.dl_from) {
    DirectOut dl;
    direct_lighting(sc, h, q.rng, dl);
    const bool v1 = dl.has1 && !(sc.curves ? occluded<false, true>(sc, dl.o1, dl.d1, st, cnt)
                                           : occluded<false>(sc, dl.o1, dl.d1, st, cnt));
    const bool v2 = dl.has2 && !(sc.curves ? occluded<false, true>(sc, dl.o2, dl.d2, st, cnt)
                                           : occluded<false>(sc, dl.o2, dl.d2, st, cnt));
    V3 sum = v3(0.f);
    if (v1) sum = sum + dl.c1;
    if (v2) sum = sum + dl.c2;
    const bool bad = any_nan(sum);
    if (bad) sum = v3(0.f);
    if (hair && !bad) {
      const V3 sigma = strand_sigma_of(sc, prim);
      const V3 W = q.adj * q.beta;  // dLoss / dc per channel
      float r[kAdParams];
      if (v1) {
        Dual f[3], pd;
        hair_eval_dual(a.hcd, h.wo_local, dl.wl1, h.hw.h, sigma, f, &pd);
        const float pb = dl.bp1, pl = dl.lp1;
#pragma unroll
        for (int k = 0; k < kAdParams; ++k) r[k] = dl.mis1 ? -2.f * pb * pd.d[k] / (pl * pl + pb * pb) : 0.f;
        add_grad(sc, a, prim, f, W * safe_ratio(dl.c1, dl.f1), dot(W, dl.c1), r, q.gp);
      }
      if (v2) {
        Dual f[3], pd;
        hair_eval_dual(a.hcd, h.wo_local, dl.wl2, h.hw.h, sigma, f, &pd);
        const float pb = dl.bp2, pl = dl.lp2;
#pragma unroll
        for (int k = 0; k < kAdParams; ++k) r[k] = pb > 0.f ? 2.f * pl * pl * pd.d[k] / (pb * (pb * pb + pl * pl)) : 0.f;
        add_grad(sc, a, prim, f, W * safe_ratio(dl.c2, dl.f2), dot(W, dl.c2), r, q.gp);
      }
    }
    q.L = q.L + q.beta * sum;
  }
  if (depth >= 1) {
    const float rr = fmaxf(0.05f, 1.f - luminance(q.beta));
    if (draw(q.rng) < rr) return false;
    q.beta = q.beta / (1.f - rr);
  }
  if (depth + 1 > a.max_vertex) return false;
  const BsdfSample bs = sample_bsdf(sc, h, q.rng);
  const V3 wl = bs.wl, wi = bs.wi, f = bs.f;
  const float pdf = bs.pdf;
  V3 mul = (pdf == 0.f) ? f : f / pdf;
  const bool mul_bad = any_nan(mul);
  if (mul_bad) mul = v3(1.f);
  if (hair && !mul_bad) {
    const V3 rem = q.Lt - q.L;  // radiance gathered after this vertex
    Dual fd[3], pd;
    hair_eval_dual(a.hcd, h.wo_local, wl, h.hw.h, strand_sigma_of(sc, prim), fd, &pd);
    float r[kAdParams];
#pragma unroll
    for (int k = 0; k < kAdParams; ++k) r[k] = 0.f;
    add_grad(sc, a, prim, fd, q.adj * safe_ratio(rem, f), 0.f, r, q.gp);
  }
  q.beta = q.beta * mul;
  V3 p = h.p, nd = h.n;
  if (dot(wi, h.n) < 0.f && hair) {
    nd = -1.f * h.n;
    p = p + (2.f * h.radius) * nd;
  }
  q.o = p + 1e-3f * nd;
  q.d = wi;
  q.depth = depth + 1;
  return true;
}

// Waves per SIMD the replay kernel is compiled for. At 512x512 C1, 16 spp,
// with path regeneration the replay ran at 55.4 M paths/s with 2 waves, 59.3
// with 3, 62-63 with 4, 54.0 with 5 and 49.8 with 6 (the registers beyond
// the budget spill to scratch: 784 bytes per lane at 4). Without
// regeneration (each lane's path run to its end while the wave waited for
// its longest path) the best was 48.3 with 3 waves.
#ifndef HM_ADJ_MIN_WAVES
#define HM_ADJ_MIN_WAVES 4
#endif
__global__ void __launch_bounds__(kTraceBlock, HM_ADJ_MIN_WAVES) k_adjoint(SceneView sc, AdjArgs a) {
  __shared__ int lds_stack[kStackLds * kTraceBlock];
  TraceStack st;
  st.lds = lds_stack + threadIdx.x;
  st.spill = a.spill + blockIdx.x * blockDim.x + threadIdx.x;
  st.spill_stride = a.spill_stride;
  st.spill_levels = a.spill_levels;
  st.sp = 0;
  const int npx = sc.width * sc.height;
  const long long total = (long long)npx * a.npass;
  float gg[3] = {0.f, 0.f, 0.f};
  const long long stride = (long long)gridDim.x * blockDim.x;
  // Path regeneration: each lane replays its paths i, i + stride, ... one
  // vertex per iteration, and a lane whose path has ended starts its next
  // one in the same iteration, so the wave's lanes stay busy instead of
  // waiting for the longest path of the wave.
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x - stride;
  bool live = false;
  AdjPath q;
  for (;;) {
    if (!live) {
      for (i += stride; i < total; i += stride) {
        const int pix = (int)(i % npx), pass = a.pass0 + (int)(i / npx);
        const float4 g4 = a.dl_dimg[pix];
        const V3 adj = a.inv_spp * v3(g4.x, g4.y, g4.z);
        if (adj.x == 0.f && adj.y == 0.f && adj.z == 0.f) continue;
        const float4 l = a.ring[i];  // i = (pass - pass0) * npx + pix
        const V3 Lt = v3(l.x, l.y, l.z);
        if (any_nan(Lt)) continue;  // writePixel drops NaN samples (utils.cuh:20-21)
        adj_start(sc, pix % sc.width, pix / sc.width, pass, q);
        q.adj = adj;
        q.Lt = Lt;
        q.pix = pix;
        live = true;
        break;
      }
    }
    if (!__any(live)) break;
    if (!live) continue;
    float t, u;
    int prim;
    TraceCounters cnt{0, 0, 0};
    if (sc.curves)
      closest_hit<false, true>(sc, q.o, q.d, st, t, prim, u, cnt);
    else
      closest_hit<false>(sc, q.o, q.d, st, t, prim, u, cnt);
    live = prim != kPrimMiss && adj_vertex(sc, a, st, t, prim, u, q);
    if (!live) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gg[k] += q.gp[k];
        if (a.g_px && q.gp[k] != 0.f) atomicAdd(&a.g_px[3 * (size_t)q.pix + k], q.gp[k]);
      }
    }
  }
  // the global gradients: wave sums, block sums in LDS, one FP64 atomic per
  // block and parameter
  __shared__ float part[kTraceBlock / 64][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float v = gg[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0) part[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < kTraceBlock / 64; ++w) v += (double)part[w][threadIdx.x];
    if (v != 0.0) atomicAdd(&a.g_glob[threadIdx.x], v);
  }
}

}  // namespace

}  // namespace hm

using namespace hm;

struct hmsnn_ctx {
  Context c;
};

extern "C" {

int hmsnn_set_hair_params(hmsnn_ctx* ctx, float beta_m, float beta_n, float alpha) {
  if (!ctx || !(beta_m > 0.f) || !(beta_n > 0.f)) return fail(HMSNN_EINVAL, "hmsnn_set_hair_params: bad arguments");
  Context& c = ctx->c;
  c.scene.beta_m = beta_m;
  c.scene.beta_n = beta_n;
  c.scene.alpha = alpha;
  const float gains[4] = {c.scene.R_G, c.scene.TT_G, c.scene.TRT_G, c.scene.TRRT_G};
  c.view.hc = make_hair_consts(beta_m, beta_n, alpha, gains);
  return HMSNN_OK;
}

int hmsnn_adjoint_render(hmsnn_ctx* ctx, int spp, const float* dl_dimg, double* grad_sigma, double grad_glob[3],
                         float* grad_px) {
  if (!ctx || spp <= 0 || !dl_dimg || !grad_glob) return fail(HMSNN_EINVAL, "hmsnn_adjoint_render: bad arguments");
  Context& c = ctx->c;
  if (!c.scene.has_hair) return fail(HMSNN_EINVAL, "scene has no hair");
  std::string err;
  int rc = c.ensure_path_buffers(false, err);
  if (rc) return fail(rc, err);
  const size_t npx = (size_t)c.scene.width * c.scene.height;
  const size_t ns = (size_t)c.hair.num_strands;
  float4* d_adj = nullptr;
  double* d_gs = nullptr;
  double* d_gg = nullptr;
  float* d_px = nullptr;
  rc = dev_alloc((void**)&d_adj, sizeof(float4) * npx, err);
  if (!rc) rc = dev_alloc((void**)&d_gs, sizeof(double) * 3 * ns, err);
  if (!rc) rc = dev_alloc((void**)&d_gg, sizeof(double) * 4, err);
  if (!rc && grad_px) rc = dev_alloc((void**)&d_px, sizeof(float) * 3 * npx, err);
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c.stream);
    dev_free(d_adj);
    dev_free(d_gs);
    dev_free(d_gg);
    dev_free(d_px);
  };
  if (rc) {
    cleanup();
    return fail(rc, err);
  }
  hipError_t e = hipMemcpyAsync(d_adj, dl_dimg, sizeof(float4) * npx, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_gs, 0, sizeof(double) * 3 * ns, c.stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_gg, 0, sizeof(double) * 4, c.stream);
  if (e == hipSuccess && d_px) e = hipMemsetAsync(d_px, 0, sizeof(float) * 3 * npx, c.stream);
  if (e != hipSuccess) {
    cleanup();
    return fail(HMSNN_EDEVICE, hipGetErrorString(e));
  }
  AdjArgs a;
  const float gains[4] = {c.scene.R_G, c.scene.TT_G, c.scene.TRT_G, c.scene.TRRT_G};
  a.hcd = make_hair_consts_ad(c.scene.beta_m, c.scene.beta_n, c.scene.alpha, gains);
  a.dl_dimg = d_adj;
  a.inv_spp = 1.f / (float)spp;
  const PassConfig pc = c.pt_pass_config();
  a.dl_zero = pc.dl_zero;
  a.dl_from = pc.dl_from;
  a.max_vertex = pc.max_vertex;
  a.g_sigma = d_gs;
  a.g_glob = d_gg;
  a.g_px = d_px;
  a.spill = c.pb.spill;
  a.spill_stride = c.pb.spill_stride;
  a.spill_levels = c.pb.spill_levels;
  // The primal radiance of every path comes from the wavefront path tracer
  // (the same paths: same seeds, same operations), kStreamRing passes at a
  // time into the context's radiance ring; the replay kernel then only runs
  // the adjoint pass over them.
  float4* tmp_acc = nullptr;
  rc = dev_alloc((void**)&tmp_acc, 2 * sizeof(float4) * (size_t)npx, err);
  for (int p0 = 0; !rc && p0 < spp; p0 += kStreamRing) {
    const int np = std::min(kStreamRing, spp - p0);
    rc = stream_passes(c, p0, np, tmp_acc, tmp_acc + npx, nullptr, err);
    if (rc) break;
    a.ring = c.stream_ring;
    a.pass0 = c.seed_offset + p0;
    a.npass = np;
    hipLaunchKernelGGL(k_adjoint, dim3(c.trace_blocks), dim3(kTraceBlock), 0, c.stream, c.view, a);
    rc = check_launch("adjoint", err);
  }
  (void)hipStreamSynchronize(c.stream);
  dev_free(tmp_acc);
  if (!rc) {
    double gg[4];
    e = hipMemcpyAsync(gg, d_gg, sizeof(gg), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess && grad_sigma)
      e = hipMemcpyAsync(grad_sigma, d_gs, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess && grad_px)
      e = hipMemcpyAsync(grad_px, d_px, sizeof(float) * 3 * npx, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) rc = HMSNN_EDEVICE, err = hipGetErrorString(e);
    else for (int k = 0; k < 3; ++k) grad_glob[k] = gg[k];
  }
  cleanup();
  return rc ? fail(rc, err) : HMSNN_OK;
}

}  // extern "C"
