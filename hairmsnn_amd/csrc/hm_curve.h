// Round cubic Catmull-Rom hair curves: ray intersection and hit geometry.
//
// The reference draws .hair strands as OptiX round Catmull-Rom curves
// (owlCurvesSetDegree(…, 3, forceCaps), render_path_tracing.cu:223,
// render_hair_msnn.cu:346, render_nrc.cu:409) whose intersector is closed.
// The primitive is defined here from what the reference's own code states:
// a segment is the cubic P(u), u in [0, 1], interpolating four control
// points (x, y, z, radius) with the Catmull-Rom matrix of
// curve_utils.cuh:57-71, and its surface is the swept circle: the points at
// distance r(u) from P(u) in the plane normal to P'(u) (the surface whose
// normal curve_utils.cuh:159-197 computes). With end caps (forceCaps, only
// at the first and last segment of a strand) flat disks close the strand at
// u = 0 and u = 1.
//
// Intersection, in a fixed sequence of IEEE FP32 operations (no fused
// multiply-add) so that a CPU restatement gives the same bits: the segment is
// split into kCurveSeeds pieces; a piece whose capsule bound (chord, radius
// bound plus the chord's deviation bound) the ray misses is skipped;
// otherwise, from the parameter u of the ray's closest approach to the
// chord, the hit is iterated: intersect the ray with the cylinder tangent to
// the tube at u (axis P(u) + s P'(u), radius r(u)), take its entry root (the
// exit root when the entry lies behind the origin; the ray point nearest the
// axis when the ray passes outside), then move u toward the parameter whose
// normal plane holds that point (a Newton step, then secant steps). At the
// fixed point the hit lies in the normal plane at u at distance r(u) from
// P(u), i.e. exactly on the swept circle. Fixed points with u outside [0, 1]
// are rejected: a ray that enters the tube's continuation past an end meets
// the neighbouring segment of its strand there, and an open strand end is no
// way into the tube. The nearest accepted hit over the pieces and the caps
// is returned.
#pragma once

#include "hm_common.h"
#include "hm_geom.h"

namespace hm {

constexpr int kCurveSeeds = 4;       // pieces of a segment that seed the iteration
constexpr int kCurveIters = 12;      // local-cylinder iterations per seed
constexpr float kCurveUTol = 1e-5f;  // converged when the Newton step in u is below this
constexpr int kCurveCap0 = 1;        // end flags: first / last segment of its strand
constexpr int kCurveCap1 = 2;

// Power-basis coefficients of one coordinate: ((a u + b) u + c) u + d.
struct Cubic {
  float a, b, c, d;
};

// Catmull-Rom to power basis (curve_utils.cuh:61-71).
HM_HD Cubic catrom_cubic(float q0, float q1, float q2, float q3) {
  HM_NOCONTRACT
  Cubic k;
  k.a = ((q3 - q0) + 3.f * (q1 - q2)) * 0.5f;
  k.b = ((2.f * q0 - 5.f * q1) + (4.f * q2 - q3)) * 0.5f;
  k.c = (q2 - q0) * 0.5f;
  k.d = q1;
  return k;
}

HM_HD float cubic_at(const Cubic& k, float u) {
  HM_NOCONTRACT
  return ((k.a * u + k.b) * u + k.c) * u + k.d;
}
HM_HD float cubic_d1(const Cubic& k, float u) {
  HM_NOCONTRACT
  return (3.f * k.a * u + 2.f * k.b) * u + k.c;
}
HM_HD float cubic_d2(const Cubic& k, float u) {
  HM_NOCONTRACT
  return 6.f * k.a * u + 2.f * k.b;
}

struct CurveSeg {
  Cubic x, y, z, r;
};

// Control points q[4] as (x, y, z, radius).
HM_HD CurveSeg make_curve(const float4 q0, const float4 q1, const float4 q2, const float4 q3) {
  CurveSeg s;
  s.x = catrom_cubic(q0.x, q1.x, q2.x, q3.x);
  s.y = catrom_cubic(q0.y, q1.y, q2.y, q3.y);
  s.z = catrom_cubic(q0.z, q1.z, q2.z, q3.z);
  s.r = catrom_cubic(q0.w, q1.w, q2.w, q3.w);
  return s;
}

HM_HD V3 curve_pos(const CurveSeg& s, float u) { return v3(cubic_at(s.x, u), cubic_at(s.y, u), cubic_at(s.z, u)); }
HM_HD V3 curve_vel(const CurveSeg& s, float u) { return v3(cubic_d1(s.x, u), cubic_d1(s.y, u), cubic_d1(s.z, u)); }
HM_HD V3 curve_acc(const CurveSeg& s, float u) { return v3(cubic_d2(s.x, u), cubic_d2(s.y, u), cubic_d2(s.z, u)); }

// Entry root (exit root when the entry is at t <= 0) of the ray with the
// infinite cylinder of radius r around the line c + s v, solved around the
// ray point nearest c. Returns false without a root at t > 0.
HM_HD bool ray_cylinder_first(float ox, float oy, float oz, float dx, float dy, float dz, V3 c, V3 v, float r,
                              float* t_out) {
  HM_NOCONTRACT
  const float vv = v.x * v.x + v.y * v.y + v.z * v.z;
  const float dd = dx * dx + dy * dy + dz * dz;
  const float tc = ((c.x - ox) * dx + (c.y - oy) * dy + (c.z - oz) * dz) / dd;
  const float qx = (ox + tc * dx) - c.x, qy = (oy + tc * dy) - c.y, qz = (oz + tc * dz) - c.z;
  const float dv = dx * v.x + dy * v.y + dz * v.z;
  const float qv = qx * v.x + qy * v.y + qz * v.z;
  const float A = dd * vv - dv * dv;
  const float B = (dx * qx + dy * qy + dz * qz) * vv - dv * qv;
  const float C = (qx * qx + qy * qy + qz * qz) * vv - qv * qv - r * r * vv;
  if (!(A > 0.f)) return false;
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
  if (tc + s0 > 0.f) {
    *t_out = tc + s0;
    return true;
  }
  if (tc + s1 > 0.f) {
    *t_out = tc + s1;
    return true;
  }
  return false;
}

// The point the iteration moves toward at parameter u: the entry root of the
// ray with the local cylinder (the exit root when the entry is at t <= 0),
// or, when the ray passes outside the cylinder, the ray point nearest its
// axis. False when the axis is parallel to the ray or the cylinder lies
// behind the origin.
HM_HD bool ray_cylinder_iterate(float ox, float oy, float oz, float dx, float dy, float dz, V3 c, V3 v, float r,
                                float* t_out) {
  HM_NOCONTRACT
  const float vv = v.x * v.x + v.y * v.y + v.z * v.z;
  const float dd = dx * dx + dy * dy + dz * dz;
  const float tc = ((c.x - ox) * dx + (c.y - oy) * dy + (c.z - oz) * dz) / dd;
  const float qx = (ox + tc * dx) - c.x, qy = (oy + tc * dy) - c.y, qz = (oz + tc * dz) - c.z;
  const float dv = dx * v.x + dy * v.y + dz * v.z;
  const float qv = qx * v.x + qy * v.y + qz * v.z;
  const float A = dd * vv - dv * dv;
  const float B = (dx * qx + dy * qy + dz * qz) * vv - dv * qv;
  const float C = (qx * qx + qy * qy + qz * qz) * vv - qv * qv - r * r * vv;
  if (!(A > 0.f)) return false;
  const float disc = B * B - A * C;
  if (!(disc >= 0.f)) {
    *t_out = tc - B / A;
    return true;
  }
  const float sq = sqrtf(disc);
  const float qq = (B >= 0.f) ? -(B + sq) : -(B - sq);
  float s0 = qq / A;
  float s1 = (qq != 0.f) ? C / qq : s0;
  if (s1 < s0) {
    const float tmp = s0;
    s0 = s1;
    s1 = tmp;
  }
  if (tc + s0 > 0.f) {
    *t_out = tc + s0;
    return true;
  }
  if (tc + s1 > 0.f) {
    *t_out = tc + s1;
    return true;
  }
  return false;
}

// Upper bound of the radius over [0, 1]: the largest Bezier control value of
// the radius cubic (the Bezier form of a Catmull-Rom segment), at least 0.
HM_HD float curve_radius_bound(float w0, float w1, float w2, float w3) {
  HM_NOCONTRACT
  const float b1 = w1 + (w2 - w0) * (1.f / 6.f);
  const float b2 = w2 - (w3 - w1) * (1.f / 6.f);
  return fmaxf(fmaxf(fmaxf(w1, w2), fmaxf(b1, b2)), 0.f);
}

// Flat end cap: the disk of radius r at c in the plane normal to n.
HM_HD bool ray_disk(float ox, float oy, float oz, float dx, float dy, float dz, V3 c, V3 n, float r, float* t_out) {
  HM_NOCONTRACT
  const float dn = dx * n.x + dy * n.y + dz * n.z;
  if (!(dn != 0.f)) return false;
  const float t = ((c.x - ox) * n.x + (c.y - oy) * n.y + (c.z - oz) * n.z) / dn;
  if (!(t > 0.f)) return false;
  const float px = (ox + t * dx) - c.x, py = (oy + t * dy) - c.y, pz = (oz + t * dz) - c.z;
  if (!(px * px + py * py + pz * pz <= r * r)) return false;
  *t_out = t;
  return true;
}

// Ray / round Catmull-Rom segment. q: the four control points; flags:
// kCurveCap0 / kCurveCap1 when caps are on and the segment starts / ends its
// strand. Nearest hit with t > 0: (t, u), u in [0, 1] (exactly 0 or 1 on a
// cap).
HM_HD bool intersect_curve(float ox, float oy, float oz, float dx, float dy, float dz, const float4 q0,
                           const float4 q1, const float4 q2, const float4 q3, int flags, float* t_out,
                           float* u_out) {
  HM_NOCONTRACT
  const CurveSeg s = make_curve(q0, q1, q2, q3);
  const float rmax = curve_radius_bound(q0.w, q1.w, q2.w, q3.w);
  const float dd = dx * dx + dy * dy + dz * dz;
  float best_t = __builtin_inff(), best_u = 0.f;
  const float h = 1.f / (float)kCurveSeeds;
  V3 e0 = curve_pos(s, 0.f);
  V3 a0 = curve_acc(s, 0.f);
#pragma unroll 1
  for (int k = 0; k < kCurveSeeds; ++k) {
    const float u0 = (float)k * h, u1 = (float)(k + 1) * h;
    const V3 e1 = curve_pos(s, u1);
    const V3 a1 = curve_acc(s, u1);
    // capsule bound of the piece: the curve stays within h^2/8 max|P''| of
    // its chord (P'' is linear in u, so its norm peaks at an end)
    const float m = fmaxf(length(a0), length(a1));
    const float R = rmax + (h * h * 0.125f) * m;
    // closest approach of the ray line to the chord segment e0 -> e1
    const V3 ev = e1 - e0;
    const float ee = dot(ev, ev);
    const float de = dx * ev.x + dy * ev.y + dz * ev.z;
    const V3 w0 = v3(ox - e0.x, oy - e0.y, oz - e0.z);
    const float dw = dx * w0.x + dy * w0.y + dz * w0.z;
    const float ew = dot(ev, w0);
    const float den = dd * ee - de * de;
    float sc = den > 0.f ? (dd * ew - de * dw) / den : 0.f;
    sc = fminf(fmaxf(sc, 0.f), 1.f);
    const float tcl = (sc * de - dw) / dd;
    const float px = (ox + tcl * dx) - (e0.x + sc * ev.x);
    const float py = (oy + tcl * dy) - (e0.y + sc * ev.y);
    const float pz = (oz + tcl * dz) - (e0.z + sc * ev.z);
    const float dist2 = px * px + py * py + pz * pz;
    const bool near = dist2 <= R * R;
    e0 = e1;
    a0 = a1;
    if (!near) continue;
    float u = u0 + sc * h;
    float t = 0.f;
    bool ok = false;
    // f(u) = (X(u) - P(u)) . P'(u) with X(u) the ray's point on the local
    // cylinder: a Newton step (X held fixed) first, secant steps after
    float up = 0.f, fp = 0.f;
    bool prev = false;
#pragma unroll 1
    for (int it = 0; it < kCurveIters; ++it) {
      const V3 c = curve_pos(s, u), v = curve_vel(s, u);
      const float r = cubic_at(s.r, u);
      if (!ray_cylinder_iterate(ox, oy, oz, dx, dy, dz, c, v, r, &t)) {
        ok = false;
        break;
      }
      const V3 x = v3((ox + t * dx) - c.x, (oy + t * dy) - c.y, (oz + t * dz) - c.z);
      const float f = dot(x, v);
      float du;
      if (prev && f != fp) {
        du = -f * (u - up) / (f - fp);
      } else {
        const float j = dot(x, curve_acc(s, u)) - dot(v, v);
        if (!(j < 0.f)) {
          ok = false;
          break;
        }
        du = -f / j;
      }
      up = u;
      fp = f;
      prev = true;
      u = u + du;
      ok = fabsf(du) <= kCurveUTol;
      if (ok || !(u >= -h && u <= 1.f + h)) break;
    }
    if (!ok || !(u >= 0.f && u <= 1.f)) continue;
    // the hit at the converged parameter
    const V3 c = curve_pos(s, u), v = curve_vel(s, u);
    if (!ray_cylinder_first(ox, oy, oz, dx, dy, dz, c, v, cubic_at(s.r, u), &t)) continue;
    if (t < best_t) {
      best_t = t;
      best_u = u;
    }
  }
  if (flags & kCurveCap0) {
    float t;
    if (ray_disk(ox, oy, oz, dx, dy, dz, curve_pos(s, 0.f), curve_vel(s, 0.f), cubic_at(s.r, 0.f), &t) &&
        t < best_t) {
      best_t = t;
      best_u = 0.f;
    }
  }
  if (flags & kCurveCap1) {
    float t;
    if (ray_disk(ox, oy, oz, dx, dy, dz, curve_pos(s, 1.f), curve_vel(s, 1.f), cubic_at(s.r, 1.f), &t) &&
        t < best_t) {
      best_t = t;
      best_u = 1.f;
    }
  }
  if (!(best_t < __builtin_inff())) return false;
  *t_out = best_t;
  *u_out = best_u;
  return true;
}

// Hit geometry at parameter u (computeCurveIntersection and surfaceNormal,
// curve_utils.cuh:159-197 and :218-241): the hit point ps is moved onto the
// swept circle in the normal plane at u, the normal follows the swept
// surface (the radius derivative and the curvature enter it); on a cap
// (u exactly 0 or 1) the normal is -+P'(u). The tangent is normalize(P'(u)).
// Velocities at u = 0 and 1 are taken at 1e-6 and 1 - 1e-6
// (curve_utils.cuh:95-101).
struct CurveHitGeom {
  V3 p, n, tangent, center;
  float radius;
};

HM_HD V3 curve_vel_safe(const CurveSeg& s, float u) {
  if (u == 0.f) u = 0.000001f;
  if (u == 1.f) u = 0.999999f;
  return curve_vel(s, u);
}

HM_HD CurveHitGeom curve_hit_geometry(const CurveSeg& s, float u, V3 ps) {
  CurveHitGeom g;
  g.center = curve_pos(s, u);
  g.radius = cubic_at(s.r, u);
  g.p = ps;
  const V3 d = curve_vel_safe(s, u);
  if (u == 0.f) {
    g.n = normalize(-1.f * d);
  } else if (u == 1.f) {
    g.n = normalize(d);
  } else {
    const float dr = cubic_d1(s.r, u);
    float dd = dot(d, d);
    V3 o1 = ps - g.center;
    o1 = o1 - (dot(o1, d) / dd) * d;
    o1 = o1 * (g.radius / length(o1));
    g.p = g.center + o1;
    dd = dd - dot(curve_acc(s, u), o1);
    g.n = normalize(dd * o1 - (dr * g.radius) * d);
  }
  g.tangent = normalize(d);
  return g;
}

}  // namespace hm
