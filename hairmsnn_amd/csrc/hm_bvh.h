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

// Slab test; returns the entry distance or +inf on a miss.
HM_D float box_enter(const RayInv& r, float lx, float ly, float lz, float hx, float hy, float hz,
                     float tmax) {
  const float tx0 = fmaf(lx, r.ix, -r.oix), tx1 = fmaf(hx, r.ix, -r.oix);
  const float ty0 = fmaf(ly, r.iy, -r.oiy), ty1 = fmaf(hy, r.iy, -r.oiy);
  const float tz0 = fmaf(lz, r.iz, -r.oiz), tz1 = fmaf(hz, r.iz, -r.oiz);
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return tn <= tf ? tn : __builtin_inff();
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

// Closest hit over hair segments: minimum (t, id) lexicographically.
template <bool kCount>
HM_D void trace_closest(const BvhView& bvh, const RayInv& r, TraceStack& st, float& best_t,
                        int& best_id, float& best_u, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return;
  st.sp = 0;
  int node = 0;
  // A valid tree visits each node at most once per ray; the cap only bounds
  // the loop if the tree were malformed.
  for (int iter = 0; iter < bvh.num_segs + 64; ++iter) {
    const BvhNode n = bvh.nodes[node];
    if (kCount) cnt.nodes++;
    float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, best_t);
    float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, best_t);
    const int c0 = n.d.x, c1 = n.d.y;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int c = k == 0 ? c0 : c1;
      const float tk = k == 0 ? t0 : t1;
      if (c < 0 && tk <= best_t) {
        const int leaf = ~c;
        const float4 s0 = bvh.segs[2 * leaf];
        const float4 s1 = bvh.segs[2 * leaf + 1];
        if (kCount) cnt.segs++;
        float t, u;
        if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z,
                              s0.w, &t, &u)) {
          const int id = __float_as_int(s1.w);
          if (t < best_t || (t == best_t && id < best_id)) {
            best_t = t;
            best_id = id;
            best_u = u;
          }
        }
      }
    }
    const bool h0 = c0 >= 0 && t0 <= best_t;
    const bool h1 = c1 >= 0 && t1 <= best_t;
    if (h0 && h1) {
      const bool first0 = t0 <= t1;
      st.push(first0 ? c1 : c0);
      node = first0 ? c0 : c1;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (st.sp == 0) break;
      node = st.pop();
    }
  }
}

// Any hit with t > 0 over hair segments (binary visibility).
template <bool kCount>
HM_D bool trace_any(const BvhView& bvh, const RayInv& r, TraceStack& st, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return false;
  const float tinf = __builtin_inff();
  st.sp = 0;
  int node = 0;
  for (int iter = 0; iter < bvh.num_segs + 64; ++iter) {
    const BvhNode n = bvh.nodes[node];
    if (kCount) cnt.nodes++;
    const float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, tinf);
    const float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, tinf);
    const int c0 = n.d.x, c1 = n.d.y;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int c = k == 0 ? c0 : c1;
      const float tk = k == 0 ? t0 : t1;
      if (c < 0 && tk < tinf) {
        const int leaf = ~c;
        const float4 s0 = bvh.segs[2 * leaf];
        const float4 s1 = bvh.segs[2 * leaf + 1];
        if (kCount) cnt.segs++;
        float t, u;
        if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z,
                              s0.w, &t, &u))
          return true;
      }
    }
    const bool h0 = c0 >= 0 && t0 < tinf;
    const bool h1 = c1 >= 0 && t1 < tinf;
    if (h0 && h1) {
      const bool first0 = t0 <= t1;
      st.push(first0 ? c1 : c0);
      node = first0 ? c0 : c1;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (st.sp == 0) break;
      node = st.pop();
    }
  }
  return false;
}

#endif  // __HIPCC__

}  // namespace hm
