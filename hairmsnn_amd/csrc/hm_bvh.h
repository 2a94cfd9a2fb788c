// Device-side BVH view and traversal (closest hit and any hit).
//
// Layout in HBM: the Karras LBVH is built binary (64-byte nodes holding both
// children's boxes) and then collapsed into wide nodes (BvhNodeWide, 4 or 8
// children) that traversal reads; segments are 32-byte records in Morton
// order. A child index >= 0 is an inner node, < 0 codes a leaf: a run of
// consecutive positions of the segment array (leaf_code).
//
// Traversal is one ray per lane (wave64). The stack's top kStackLds entries
// live in LDS, one column per lane (conflict-free), deeper entries spill to a
// per-thread global area sized from the wide tree's depth at build time.
#pragma once

#include "hm_common.h"
#include "hm_curve.h"
#include "hm_geom.h"

namespace hm {

struct BvhNode {
  float4 a;  // child0 lo.xyz, child0 hi.x
  float4 b;  // child0 hi.yz,  child1 lo.xy
  float4 c;  // child1 lo.z,   child1 hi.xyz
  int4 d;    // child0, child1, unused, unused
};

// Wide node of kBvhWidth children (4: 128 bytes, 8: 224 bytes): child boxes
// as structure-of-arrays, so one node fetch is a run of 16-byte loads that
// decides all children. child[k] >= 0 is an inner node, < 0 is ~segment
// position (leaf), kEmptyChild an unused slot (its box is never tested).
#ifndef HM_BVH_WIDTH
#define HM_BVH_WIDTH 4
#endif
constexpr int kBvhWidth = HM_BVH_WIDTH;
static_assert(kBvhWidth == 4 || kBvhWidth == 8, "4- or 8-wide nodes");
struct __attribute__((aligned(16))) BvhNodeWide {
  float lox[kBvhWidth], loy[kBvhWidth], loz[kBvhWidth];
  float hix[kBvhWidth], hiy[kBvhWidth], hiz[kBvhWidth];
  int child[kBvhWidth];
  int pad[kBvhWidth == 4 ? 4 : 0];  // 4-wide: one 128-byte line; 8-wide: 224 bytes
};
constexpr int kEmptyChild = 0x7fffffff;


struct BvhView {
  const BvhNodeWide* nodes;  // wide tree, root at 0
  const float4* segs;     // 2 float4 per segment: (p0, r), (p1, id bits)
  int num_segs;
};

struct SphereView {
  int enabled;
  float cx, cy, cz, r;
};

// OBJ head mesh: BVH over the triangles' boxes, vertices (3 float4 per
// triangle) for intersection, and per triangle 3 vertex normals + the diffuse
// colour (4 float4) for shading.
struct MeshView {
  BvhView bvh;
  const float4* tri;
  const float4* shade;
  int num;
};

// Stack entries per lane kept in LDS (4 bytes x 256 lanes each per block).
#ifndef HM_STACK_LDS
#define HM_STACK_LDS 16
#endif
constexpr int kStackLds = HM_STACK_LDS;
constexpr int kTraceBlock = 256;

// Launch bounds of the traversal kernels: 256 threads, and a floor of
// HM_TRACE_MIN_WAVES waves per SIMD (the register budget follows from it).
#ifndef HM_TRACE_MIN_WAVES
#define HM_TRACE_MIN_WAVES 8
#endif
#define HM_TRACE_BOUNDS __launch_bounds__(kTraceBlock, HM_TRACE_MIN_WAVES)
// Traversal kernels over round curves: the curve iteration needs more
// registers than the segment test, fewer waves.
#ifndef HM_CURVE_MIN_WAVES
#define HM_CURVE_MIN_WAVES 4
#endif
#define HM_TRACE_BOUNDS_T(curves) __launch_bounds__(kTraceBlock, (curves) ? HM_CURVE_MIN_WAVES : HM_TRACE_MIN_WAVES)

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
// 0); each iteration either visits one wide node (nearest hit child next,
// the other hit children pushed far to near, leaves included) or tests one
// segment. Keeping the two steps apart bounds the code a wave executes per
// iteration to one node visit plus one segment test. The cap bounds the loop
// for a malformed tree only: a valid tree visits each node and leaf at most
// once per ray.
constexpr int kNone = 0x7fffffff;

// Entry distances of the child boxes against [0, tmax]; +inf for a miss or
// an empty slot (so that sorting puts them last).
HM_D void node_hits(const BvhNodeWide& n, const RayInv& r, float tmax, float t[kBvhWidth]) {
#pragma unroll
  for (int k = 0; k < kBvhWidth; ++k) {
    const float e = box_enter(r, n.lox[k], n.loy[k], n.loz[k], n.hix[k], n.hiy[k], n.hiz[k], tmax);
    t[k] = (n.child[k] != kEmptyChild && e <= tmax) ? e : __builtin_inff();
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

// Per-lane traversal state, advanced one step (one wide node or one segment)
// at a time so that a kernel can hand an idle lane a new ray while the rest of
// its wave is still traversing.
struct ClosestState {
  float best_t, best_u;
  int best_id;
  int cur;   // current entry: node (>= 0), ~leaf (< 0) or kNone when done
  int iter;  // steps taken (cap for a malformed tree)
};

HM_D int trace_cap(const BvhView& bvh) { return 2 * bvh.num_segs + 64; }

// Starts a closest-hit query with (best_t, best_id) preset by the caller
// (e.g. a hit on the head sphere).
HM_D void closest_begin(const BvhView& bvh, TraceStack& st, ClosestState& s) {
  st.sp = 0;
  s.cur = bvh.num_segs > 0 ? 0 : kNone;
  s.iter = 0;
}

// Leaf tests: a leaf record is two float4s in Morton order whose second w
// holds the primitive id. Hair: the record is the segment itself. Triangle
// meshes: the record holds the triangle's box, the vertices are read by id.
struct SegLeaf {
  HM_D bool operator()(const BvhView& bvh, const RayInv& r, int leaf, float* t, float* u, int* id) const {
    const float4 s0 = bvh.segs[2 * leaf];
    const float4 s1 = bvh.segs[2 * leaf + 1];
    *id = __float_as_int(s1.w);
    return intersect_segment(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, s0.w, t, u);
  }
};

struct TriLeaf {
  const float4* tri;  // 3 float4 per triangle, original order
  HM_D bool operator()(const BvhView& bvh, const RayInv& r, int leaf, float* t, float* u, int* id) const {
    const int k = __float_as_int(bvh.segs[2 * leaf + 1].w);
    *id = k;
    const float4 a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
    float v;
    return intersect_triangle(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t,
                              u, &v);
  }
};

// Round Catmull-Rom hair: the record holds the segment's box, the control
// points (5 float4 per segment: q0..q3, then strand and end flags) are read
// by id; caps: end caps on (the flags then say which ends have one).
struct CurveLeaf {
  const float4* crv;
  int caps;
  HM_D bool operator()(const BvhView& bvh, const RayInv& r, int leaf, float* t, float* u, int* id) const {
    const int k = __float_as_int(bvh.segs[2 * leaf + 1].w);
    *id = k;
    const float4* q = crv + 5 * (size_t)k;
    const int flags = caps ? (int)q[4].y : 0;
    return intersect_curve(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, q[0], q[1], q[2], q[3], flags, t, u);
  }
};

// Closest hit over the tree's primitives: minimum (t, id) lexicographically.
template <bool kCount, class Leaf = SegLeaf>
HM_D void closest_step(const BvhView& bvh, const RayInv& r, TraceStack& st, ClosestState& s, TraceCounters& cnt,
                       const Leaf& leaf_test = Leaf{}) {
  if (s.cur < 0) {
    const int code = ~s.cur, leaf = code & kLeafPosMask;
    if (kCount) cnt.segs++;
    float t, u;
    int id;
    if (leaf_test(bvh, r, leaf, &t, &u, &id)) {
      if (t < s.best_t || (t == s.best_t && id < s.best_id)) {
        s.best_t = t;
        s.best_id = id;
        s.best_u = u;
      }
    }
    // the rest of a leaf run, one primitive per step
    const int rest = leaf_rest(code);
    s.cur = rest ? rest : (st.sp ? st.pop() : kNone);
  } else {
    const BvhNodeWide n = bvh.nodes[s.cur];
    if (kCount) cnt.nodes++;
    float t[kBvhWidth];
    node_hits(n, r, s.best_t, t);
    int c[kBvhWidth];
#pragma unroll
    for (int k = 0; k < kBvhWidth; ++k) c[k] = n.child[k];
    // sorting network (Batcher): nearest first, misses (+inf) last
    if (kBvhWidth == 4) {
      cswap(t[0], c[0], t[1], c[1]);
      cswap(t[2], c[2], t[3], c[3]);
      cswap(t[0], c[0], t[2], c[2]);
      cswap(t[1], c[1], t[3], c[3]);
      cswap(t[1], c[1], t[2], c[2]);
    } else {
      constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                  {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
      for (int q = 0; q < 19; ++q) cswap(t[net[q][0]], c[net[q][0]], t[net[q][1]], c[net[q][1]]);
    }
    const float inf = __builtin_inff();
#pragma unroll
    for (int k = kBvhWidth - 1; k >= 1; --k)
      if (t[k] < inf) st.push(c[k]);
    s.cur = t[0] < inf ? c[0] : (st.sp ? st.pop() : kNone);
  }
  if (++s.iter >= trace_cap(bvh)) s.cur = kNone;
}

template <bool kCount, class Leaf = SegLeaf>
HM_D void trace_closest(const BvhView& bvh, const RayInv& r, TraceStack& st, float& best_t, int& best_id,
                        float& best_u, TraceCounters& cnt, const Leaf& leaf_test = Leaf{}) {
  ClosestState s;
  s.best_t = best_t;
  s.best_id = best_id;
  s.best_u = best_u;
  closest_begin(bvh, st, s);
  while (s.cur != kNone) closest_step<kCount, Leaf>(bvh, r, st, s, cnt, leaf_test);
  best_t = s.best_t;
  best_id = s.best_id;
  best_u = s.best_u;
}

// Any hit with t > 0 over hair segments (binary visibility); children are
// taken in slot order.
struct AnyState {
  int cur;
  int iter;
  bool hit;
};

HM_D void any_begin(const BvhView& bvh, TraceStack& st, AnyState& s) {
  st.sp = 0;
  s.cur = bvh.num_segs > 0 ? 0 : kNone;
  s.iter = 0;
  s.hit = false;
}

template <bool kCount, class Leaf = SegLeaf>
HM_D void any_step(const BvhView& bvh, const RayInv& r, TraceStack& st, AnyState& s, TraceCounters& cnt,
                   const Leaf& leaf_test = Leaf{}) {
  const float tinf = __builtin_inff();
  if (s.cur < 0) {
    const int code = ~s.cur, leaf = code & kLeafPosMask;
    if (kCount) cnt.segs++;
    float t, u;
    int id;
    if (leaf_test(bvh, r, leaf, &t, &u, &id)) {
      s.hit = true;
      s.cur = kNone;
    } else {
      const int rest = leaf_rest(code);
      s.cur = rest ? rest : (st.sp ? st.pop() : kNone);
    }
  } else {
    const BvhNodeWide n = bvh.nodes[s.cur];
    if (kCount) cnt.nodes++;
    float t[kBvhWidth];
    node_hits(n, r, tinf, t);
    int next = kNone;
#pragma unroll
    for (int k = kBvhWidth - 1; k >= 0; --k) {
      if (t[k] < tinf) {
        if (next != kNone) st.push(next);
        next = n.child[k];
      }
    }
    s.cur = next != kNone ? next : (st.sp ? st.pop() : kNone);
  }
  if (++s.iter >= trace_cap(bvh)) s.cur = kNone;
}

template <bool kCount, class Leaf = SegLeaf>
HM_D bool trace_any(const BvhView& bvh, const RayInv& r, TraceStack& st, TraceCounters& cnt,
                    const Leaf& leaf_test = Leaf{}) {
  AnyState s;
  any_begin(bvh, st, s);
  while (s.cur != kNone) any_step<kCount, Leaf>(bvh, r, st, s, cnt, leaf_test);
  return s.hit;
}

#endif  // __HIPCC__

}  // namespace hm
