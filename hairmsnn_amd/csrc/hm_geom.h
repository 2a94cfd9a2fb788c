// Hair-segment and head-sphere primitives: ray intersection and hit geometry.
//
// A hair segment is a finite open cylinder of constant radius r around the
// axis p0 -> p1 (the linear-curve branch of the reference,
// cuda_headers/curve_utils.cuh:12-49 and :242-255; the reference hands the
// intersection itself to OptiX's closed curve intersector, so the primitive is
// defined here and documented). A ray hits the segment at the smallest root
// t > 0 of the cylinder quadric whose axial coordinate u lies in [0, 1].
//
// The intersection is written as a fixed sequence of IEEE FP32 operations
// (+ - * / sqrt, no fused multiply-add) so that the GPU traversal and a CPU
// brute-force search produce bit-identical (t, u) for every (ray, segment)
// pair. The closest hit is the minimum of (t, segment id) in lexicographic
// order, which makes ties at shared strand joints deterministic.
#pragma once

#include "hm_common.h"

namespace hm {

// Segment record as stored in HBM: two float4s (32 bytes).
struct SegRec {
  float p0x, p0y, p0z, r;
  float p1x, p1y, p1z;
  int32_t id;  // original segment index
};

#if defined(__HIPCC__)
#define HM_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define HM_NOCONTRACT
#endif

// Ray / finite-cylinder test. Returns true and writes (t, u) for the nearest
// root with t > 0 and u in [0, 1]. The quadratic is solved around the ray
// point closest to the segment centre, which keeps the constant term at the
// segment's scale instead of the scene's.
HM_HD bool intersect_segment(float ox, float oy, float oz, float dx, float dy, float dz,
                             float p0x, float p0y, float p0z, float p1x, float p1y, float p1z,
                             float r, float* t_out, float* u_out) {
  HM_NOCONTRACT
  const float ax = p1x - p0x, ay = p1y - p0y, az = p1z - p0z;
  const float aa = ax * ax + ay * ay + az * az;
  const float dd = dx * dx + dy * dy + dz * dz;
  // shift the origin to the ray point nearest the segment centre
  const float cx = (p0x + p1x) * 0.5f, cy = (p0y + p1y) * 0.5f, cz = (p0z + p1z) * 0.5f;
  const float tc = ((cx - ox) * dx + (cy - oy) * dy + (cz - oz) * dz) / dd;
  const float qx = (ox + tc * dx) - p0x, qy = (oy + tc * dy) - p0y, qz = (oz + tc * dz) - p0z;
  const float da = dx * ax + dy * ay + dz * az;
  const float qa = qx * ax + qy * ay + qz * az;
  const float A = dd * aa - da * da;
  const float B = (dx * qx + dy * qy + dz * qz) * aa - da * qa;
  const float C = (qx * qx + qy * qy + qz * qz) * aa - qa * qa - r * r * aa;
  if (!(A > 0.f)) return false;  // ray parallel to the axis (or degenerate)
  const float disc = B * B - A * C;
  if (!(disc >= 0.f)) return false;
  const float sq = sqrtf(disc);
  const float qq = (B >= 0.f) ? -(B + sq) : -(B - sq);
  float s0 = qq / A;
  float s1 = (qq != 0.f) ? C / qq : s0;
  if (s1 < s0) {
    const float tmp = s0;
    s0 = s1;
    s1 = tmp;
  }
  const float inv_aa = 1.f / aa;
  // nearer root first
  float t = tc + s0;
  float u = (qa + s0 * da) * inv_aa;
  if (t > 0.f && u >= 0.f && u <= 1.f) {
    *t_out = t;
    *u_out = u;
    return true;
  }
  t = tc + s1;
  u = (qa + s1 * da) * inv_aa;
  if (t > 0.f && u >= 0.f && u <= 1.f) {
    *t_out = t;
    *u_out = u;
    return true;
  }
  return false;
}

// Ray / sphere (head proxy). Nearest root with t > 0.
// Ray / triangle (Moller-Trumbore, double-sided): nearest t > 0 with
// barycentrics (u, v) of vertices 1 and 2, in fixed FP32 operations like the
// segment test.
HM_HD bool intersect_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay,
                              float az, float bx, float by, float bz, float cx, float cy, float cz, float* t_out,
                              float* u_out, float* v_out) {
  HM_NOCONTRACT
  const float e1x = bx - ax, e1y = by - ay, e1z = bz - az;
  const float e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
  const float px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
  const float det = e1x * px + e1y * py + e1z * pz;
  if (!(det != 0.f)) return false;
  const float inv = 1.f / det;
  const float sx = ox - ax, sy = oy - ay, sz = oz - az;
  const float u = (sx * px + sy * py + sz * pz) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  const float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
  const float v = (dx * qx + dy * qy + dz * qz) * inv;
  if (!(v >= 0.f && u + v <= 1.f)) return false;
  const float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (!(t > 0.f)) return false;
  *t_out = t;
  *u_out = u;
  *v_out = v;
  return true;
}

HM_HD bool intersect_sphere(V3 o, V3 d, V3 c, float radius, float* t_out) {
  HM_NOCONTRACT
  const float dd = d.x * d.x + d.y * d.y + d.z * d.z;
  const float tc = ((c.x - o.x) * d.x + (c.y - o.y) * d.y + (c.z - o.z) * d.z) / dd;
  const float qx = (o.x + tc * d.x) - c.x, qy = (o.y + tc * d.y) - c.y, qz = (o.z + tc * d.z) - c.z;
  const float C = (qx * qx + qy * qy + qz * qz) - radius * radius;
  const float disc = -dd * C;  // B = dot(d, q) = 0 at the closest point
  if (!(disc >= 0.f)) return false;
  const float s = sqrtf(disc) / dd;
  if (tc - s > 0.f) {
    *t_out = tc - s;
    return true;
  }
  if (tc + s > 0.f) {
    *t_out = tc + s;
    return true;
  }
  return false;
}

// Hit geometry on a segment at axial coordinate u (curve_utils.cuh:159-197
// for the linear evaluator): the point is projected onto the radius-r
// surface, the normal is radial (flat end caps at exactly u = 0 or 1), the
// tangent is normalize(p1 - p0) (curve_utils.cuh:253-254).
struct SegHitGeom {
  V3 p, n, tangent;
};

HM_HD SegHitGeom segment_hit_geometry(V3 p0, V3 p1, float r, float u, V3 ps) {
  const V3 dv = p1 - p0;
  SegHitGeom g;
  g.p = ps;
  if (u == 0.f) {
    g.n = normalize(-dv);
  } else if (u == 1.f) {
    g.n = normalize(dv);
  } else {
    const V3 pa = p0 + u * dv;
    const float dd = dot(dv, dv);
    V3 o1 = ps - pa;
    o1 = o1 - (dot(o1, dv) / dd) * dv;
    o1 = o1 * (r / length(o1));
    g.p = pa + o1;
    g.n = normalize(dd * o1);
  }
  g.tangent = normalize(dv);
  return g;
}

}  // namespace hm
