// Device-side BVH view and traversal (closest hit and any hit).
//
// Layout in HBM: the Karras LBVH is built binary (64-byte nodes holding both
// children's boxes) and then collapsed into 4-wide nodes of 128 bytes
// (BvhNode4) that traversal reads; segments are 32-byte records in Morton
// order. A child index >= 0 is an inner node, < 0 is ~leaf where leaf is the
// position in the Morton-ordered segment array.
//
// Traversal is one ray per lane (wave64). The stack's top kStackLds entries
// live in LDS, one column per lane (conflict-free), deeper entries spill to a
// per-thread global area sized from the 4-wide tree's depth at build time.
#pragma once

#include "hm_common.h"
#include "hm_geom.h"

namespace hm {

struct BvhNode {
  float4 a;  // child0 lo.xyz, child0 hi.x
  float4 b;  // child0 hi.yz,  child1 lo.xy
  float4 c;  // child1 lo.z,   child1 hi.xyz
  int4 d;    // child0, child1, unused, unused
};

// 4-wide node (128 bytes): child boxes as structure-of-arrays, so one node
// fetch is eight 16-byte loads that decide all four children. child[k] >= 0
// is an inner node, < 0 is ~segment position (leaf), kEmptyChild is an
// unused slot (its box is never tested).
struct BvhNode4 {
  float4 lox, loy, loz;
  float4 hix, hiy, hiz;
  int4 child;
  int4 pad;
};
constexpr int kEmptyChild = 0x7fffffff;

struct BvhView {
  const BvhNode4* nodes;  // 4-wide tree, root at 0
  const float4* segs;     // 2 float4 per segment: (p0, r), (p1, id bits)
  int num_segs;
};

struct SphereView {
  int enabled;
  float cx, cy, cz, r;
};

constexpr int kStackLds = 16;
constexpr int kTraceBlock = 256;

// Launch bounds of the traversal kernels: 256 threads, and a floor of
// HM_TRACE_MIN_WAVES waves per SIMD (the register budget follows from it).
#ifndef HM_TRACE_MIN_WAVES
#define HM_TRACE_MIN_WAVES 8
#endif
#define HM_TRACE_BOUNDS __launch_bounds__(kTraceBlock, HM_TRACE_MIN_WAVES)

struct TraceCounters {
  uint32_t nodes;
  uint32_t segs;
  uint32_t rays;
};

#if defined(__HIPCC__)

struct RayInv {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz, oix, oiy, oiz;
};

HM_D float safe_inv(float d) {
  const float e = 1e-20f;
  return 1.f / (fabsf(d) > e ? d : copysignf(e, d));
}

HM_D RayInv make_ray(float ox, float oy, float oz, float dx, float dy, float dz) {
  RayInv r;
  r.ox = ox; r.oy = oy; r.oz = oz;
  r.dx = dx; r.dy = dy; r.dz = dz;
  r.ix = safe_inv(dx); r.iy = safe_inv(dy); r.iz = safe_inv(dz);
  r.oix = ox * r.ix; r.oiy = oy * r.iy; r.oiz = oz * r.iz;
  return r;
}

// Slab test against [0, tmax]; returns the entry distance, or NaN on a miss
// so that every ordered comparison with it is false (a +inf sentinel would
// pass `t <= best_t` while nothing has been hit yet).
HM_D float box_enter(const RayInv& r, float lx, float ly, float lz, float hx, float hy, float hz,
                     float tmax) {
  const float tx0 = fmaf(lx, r.ix, -r.oix), tx1 = fmaf(hx, r.ix, -r.oix);
  const float ty0 = fmaf(ly, r.iy, -r.oiy), ty1 = fmaf(hy, r.iy, -r.oiy);
  const float tz0 = fmaf(lz, r.iz, -r.oiz), tz1 = fmaf(hz, r.iz, -r.oiz);
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return tn <= tf ? tn : __builtin_nanf("");
}

struct TraceStack {
  int* lds;       // this lane's column base: lds[k * kTraceBlock]
  int* spill;     // global spill base for this thread
  int spill_stride;
  int spill_levels;  // capacity of the spill area (entries per thread)
  int sp;
  HM_D void push(int v) {
    if (sp < kStackLds) lds[sp * kTraceBlock] = v;
    else if (sp - kStackLds < spill_levels) spill[(sp - kStackLds) * spill_stride] = v;
    else return;  // unreachable for a tree within the sized depth; drop rather than overrun
    ++sp;
  }
  HM_D int pop() {
    --sp;
    return sp < kStackLds ? lds[sp * kTraceBlock] : spill[(sp - kStackLds) * spill_stride];
  }
};

// Traversal: the current entry is an inner node (>= 0) or a leaf (~index <
// 0); each iteration either visits one 4-wide node (nearest hit child next,
// the other hit children pushed far to near, leaves included) or tests one
// segment. Keeping the two steps apart bounds the code a wave executes per
// iteration to one node visit plus one segment test. The cap bounds the loop
// for a malformed tree only: a valid tree visits each node and leaf at most
// once per ray.
constexpr int kNone = 0x7fffffff;

// Entry distances of the four child boxes against [0, tmax]; +inf for a miss
// or an empty slot (so that sorting puts them last).
HM_D void node4_hits(const BvhNode4& n, const RayInv& r, float tmax, float t[4]) {
  const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w}, ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w};
  const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w}, hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
  const float hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w}, hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
  const int ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float e = box_enter(r, lx[k], ly[k], lz[k], hx[k], hy[k], hz[k], tmax);
    t[k] = (ch[k] != kEmptyChild && e <= tmax) ? e : __builtin_inff();
  }
}

HM_D void cswap(float& ta, int& ca, float& tb, int& cb) {
  const bool s = tb < ta;
  const float t = s ? tb : ta;
  tb = s ? ta : tb;
  ta = t;
  const int c = s ? cb : ca;
  cb = s ? ca : cb;
  ca = c;
}

// Closest hit over hair segments: minimum (t, id) lexicographically.
template <bool kCount>
HM_D void trace_closest(const BvhView& bvh, const RayInv& r, TraceStack& st, float& best_t, int& best_id,
                        float& best_u, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return;
  st.sp = 0;
  int cur = 0;
  for (int iter = 0; iter < 2 * bvh.num_segs + 64 && cur != kNone; ++iter) {
    if (cur < 0) {
      const int leaf = ~cur;
      const float4 s0 = bvh.segs[2 * leaf];
      const float4 s1 = bvh.segs[2 * leaf + 1];
      if (kCount) cnt.segs++;
      float t, u;
      if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, s0.w, &t, &u)) {
        const int id = __float_as_int(s1.w);
        if (t < best_t || (t == best_t && id < best_id)) {
          best_t = t;
          best_id = id;
          best_u = u;
        }
      }
      cur = st.sp ? st.pop() : kNone;
    } else {
      const BvhNode4 n = bvh.nodes[cur];
      if (kCount) cnt.nodes++;
      float t[4];
      node4_hits(n, r, best_t, t);
      int c[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
      // sorting network for four: nearest first, misses (+inf) last
      cswap(t[0], c[0], t[1], c[1]);
      cswap(t[2], c[2], t[3], c[3]);
      cswap(t[0], c[0], t[2], c[2]);
      cswap(t[1], c[1], t[3], c[3]);
      cswap(t[1], c[1], t[2], c[2]);
      const float inf = __builtin_inff();
      if (t[3] < inf) st.push(c[3]);
      if (t[2] < inf) st.push(c[2]);
      if (t[1] < inf) st.push(c[1]);
      cur = t[0] < inf ? c[0] : (st.sp ? st.pop() : kNone);
    }
  }
}

// Any hit with t > 0 over hair segments (binary visibility); children are
// taken in slot order.
template <bool kCount>
HM_D bool trace_any(const BvhView& bvh, const RayInv& r, TraceStack& st, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return false;
  const float tinf = __builtin_inff();
  st.sp = 0;
  int cur = 0;
  for (int iter = 0; iter < 2 * bvh.num_segs + 64 && cur != kNone; ++iter) {
    if (cur < 0) {
      const int leaf = ~cur;
      const float4 s0 = bvh.segs[2 * leaf];
      const float4 s1 = bvh.segs[2 * leaf + 1];
      if (kCount) cnt.segs++;
      float t, u;
      if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, s0.w, &t, &u))
        return true;
      cur = st.sp ? st.pop() : kNone;
    } else {
      const BvhNode4 n = bvh.nodes[cur];
      if (kCount) cnt.nodes++;
      float t[4];
      node4_hits(n, r, tinf, t);
      const int c[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
      int next = kNone;
#pragma unroll
      for (int k = 3; k >= 0; --k) {
        if (t[k] < tinf) {
          if (next != kNone) st.push(next);
          next = c[k];
        }
      }
      cur = next != kNone ? next : (st.sp ? st.pop() : kNone);
    }
  }
  return false;
}

#endif  // __HIPCC__

}  // namespace hm
