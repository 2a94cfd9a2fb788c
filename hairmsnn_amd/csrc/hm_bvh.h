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

// Traversal runs in two phases per loop iteration so the expensive segment
// test executes with many lanes at once instead of inside every node visit:
//   inner step: a lane with an inner node to visit fetches it and tests both
//     child boxes; a hit leaf child becomes the lane's pending leaf (one
//     slot) or, if the slot is taken, goes on the stack as ~leaf; inner
//     children continue near-first;
//   leaf step: when at least kLeafBatch lanes hold a pending leaf, or no lane
//     has an inner node left, every lane with a pending leaf tests it.
// Stack entries are inner nodes (>= 0) or leaves (~index < 0). A lane never
// sits with an empty node, an empty leaf slot and a non-empty stack, so it is
// done exactly when both are empty; every iteration advances at least one
// lane, and the iteration cap bounds the loop for a malformed tree only.
constexpr int kLeafBatch = 16;
constexpr int kNone = 0x7fffffff;

// Takes the stack top into node or leaf (the leaf slot must be empty).
HM_D void refill(TraceStack& st, int& node, int& leaf) {
  if (st.sp == 0) return;
  const int e = st.pop();
  if (e < 0) leaf = ~e;
  else node = e;
}

HM_D int wave_popc(bool p) { return __popcll(__ballot(p)); }

// Closest hit over hair segments: minimum (t, id) lexicographically.
template <bool kCount>
HM_D void trace_closest(const BvhView& bvh, const RayInv& r, TraceStack& st, float& best_t,
                        int& best_id, float& best_u, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return;
  st.sp = 0;
  int node = 0, leaf = kNone;
  const long long cap = 128ll * bvh.num_segs + 256;
  for (long long iter = 0; iter < cap; ++iter) {
    if (node != kNone) {
      const BvhNode n = bvh.nodes[node];
      if (kCount) cnt.nodes++;
      const float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, best_t);
      const float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, best_t);
      const bool h0 = t0 <= best_t, h1 = t1 <= best_t;  // false on a miss (NaN)
      const bool sw = h0 && h1 && t1 < t0;
      const int ca = sw ? n.d.y : n.d.x, cb = sw ? n.d.x : n.d.y;
      const bool ha = sw ? h1 : h0, hb = sw ? h0 : h1;
      int next = kNone;
      if (ha) {
        if (ca >= 0) next = ca;
        else if (leaf == kNone) leaf = ~ca;
        else st.push(ca);
      }
      if (hb) {
        if (cb >= 0) {
          if (next == kNone) next = cb;
          else st.push(cb);
        } else if (leaf == kNone) {
          leaf = ~cb;
        } else {
          st.push(cb);
        }
      }
      node = next;
      if (node == kNone && leaf == kNone) refill(st, node, leaf);
    }
    const int pending = wave_popc(leaf != kNone);
    if (pending >= kLeafBatch || (pending > 0 && wave_popc(node != kNone) == 0)) {
      if (leaf != kNone) {
        const float4 s0 = bvh.segs[2 * leaf];
        const float4 s1 = bvh.segs[2 * leaf + 1];
        if (kCount) cnt.segs++;
        float t, u;
        if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, s0.w, &t,
                              &u)) {
          const int id = __float_as_int(s1.w);
          if (t < best_t || (t == best_t && id < best_id)) {
            best_t = t;
            best_id = id;
            best_u = u;
          }
        }
        leaf = kNone;
        if (node == kNone) refill(st, node, leaf);
      }
    }
    if (wave_popc(node != kNone || leaf != kNone) == 0) break;
  }
}

// Any hit with t > 0 over hair segments (binary visibility).
template <bool kCount>
HM_D bool trace_any(const BvhView& bvh, const RayInv& r, TraceStack& st, TraceCounters& cnt) {
  if (bvh.num_segs <= 0) return false;
  const float tinf = __builtin_inff();
  st.sp = 0;
  int node = 0, leaf = kNone;
  bool hit = false;
  const long long cap = 128ll * bvh.num_segs + 256;
  for (long long iter = 0; iter < cap; ++iter) {
    if (node != kNone) {
      const BvhNode n = bvh.nodes[node];
      if (kCount) cnt.nodes++;
      const float t0 = box_enter(r, n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, tinf);
      const float t1 = box_enter(r, n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, tinf);
      const bool h0 = t0 < tinf, h1 = t1 < tinf;
      const bool sw = h0 && h1 && t1 < t0;
      const int ca = sw ? n.d.y : n.d.x, cb = sw ? n.d.x : n.d.y;
      const bool ha = sw ? h1 : h0, hb = sw ? h0 : h1;
      int next = kNone;
      if (ha) {
        if (ca >= 0) next = ca;
        else if (leaf == kNone) leaf = ~ca;
        else st.push(ca);
      }
      if (hb) {
        if (cb >= 0) {
          if (next == kNone) next = cb;
          else st.push(cb);
        } else if (leaf == kNone) {
          leaf = ~cb;
        } else {
          st.push(cb);
        }
      }
      node = next;
      if (node == kNone && leaf == kNone) refill(st, node, leaf);
    }
    const int pending = wave_popc(leaf != kNone);
    if (pending >= kLeafBatch || (pending > 0 && wave_popc(node != kNone) == 0)) {
      if (leaf != kNone) {
        const float4 s0 = bvh.segs[2 * leaf];
        const float4 s1 = bvh.segs[2 * leaf + 1];
        if (kCount) cnt.segs++;
        float t, u;
        leaf = kNone;
        if (intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, s0.w, &t,
                              &u)) {
          hit = true;
          node = kNone;
          st.sp = 0;
        } else if (node == kNone) {
          refill(st, node, leaf);
        }
      }
    }
    if (wave_popc(node != kNone || leaf != kNone) == 0) break;
  }
  return hit;
}

#endif  // __HIPCC__

}  // namespace hm
