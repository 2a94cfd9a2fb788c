// Device-side BVH view and traversal (closest hit and any hit).
//
// Layout in HBM: binary nodes of 64 bytes that hold both children's boxes
// (so one node fetch decides both children), and 32-byte segment records in
// Morton order. Child index >= 0 is an internal node, < 0 is ~leaf where leaf
// is the position in the Morton-ordered segment array.
//
// Traversal is one ray per lane (wave64). The stack's top kStackLds entries
// live in LDS, one column per lane (conflict-free), deeper entries spill to a
// per-thread global area sized from the tree depth at build time.
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

struct BvhView {
  const BvhNode* nodes;
  const float4* segs;  // 2 float4 per segment: (p0, r), (p1, id bits)
  int num_segs;
};

struct SphereView {
  int enabled;
  float cx, cy, cz, r;
};

constexpr int kStackLds = 16;
constexpr int kTraceBlock = 256;

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
// 0); each iteration either visits one node (nearer hit child next, the other
// pushed, leaves included) or tests one segment. Keeping the two steps apart
// bounds the per-iteration code a wave executes to one node visit plus one
// segment test (testing leaf children inside the node visit costs up to two
// segment tests per iteration under divergence). The cap bounds the loop for
// a malformed tree only: a valid tree visits each node and leaf at most once.
constexpr int kNone = 0x7fffffff;
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
      const BvhNode n = bvh.nodes[cur];
      if (kCount) cnt.nodes++;
      const float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, best_t);
      const float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, best_t);
      const bool h0 = t0 <= best_t, h1 = t1 <= best_t;
      const bool sw = h0 && h1 && t1 < t0;
      const int ca = sw ? n.d.y : n.d.x, cb = sw ? n.d.x : n.d.y;
      const bool ha = sw ? h1 : h0, hb = sw ? h0 : h1;
      if (ha && hb) {
        st.push(cb);
        cur = ca;
      } else if (ha) {
        cur = ca;
      } else if (hb) {
        cur = cb;
      } else {
        cur = st.sp ? st.pop() : kNone;
      }
    }
  }
}

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
      const BvhNode n = bvh.nodes[cur];
      if (kCount) cnt.nodes++;
      const float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, tinf);
      const float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, tinf);
      const bool h0 = t0 < tinf, h1 = t1 < tinf;
      const bool sw = h0 && h1 && t1 < t0;
      const int ca = sw ? n.d.y : n.d.x, cb = sw ? n.d.x : n.d.y;
      const bool ha = sw ? h1 : h0, hb = sw ? h0 : h1;
      if (ha && hb) {
        st.push(cb);
        cur = ca;
      } else if (ha) {
        cur = ca;
      } else if (hb) {
        cur = cb;
      } else {
        cur = st.sp ? st.pop() : kNone;
      }
    }
  }
  return false;
}

#endif  // __HIPCC__

}  // namespace hm
