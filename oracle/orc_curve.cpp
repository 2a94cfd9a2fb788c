// ORACLE - test infrastructure only (see orc_math.h).
//
// Round cubic Catmull-Rom hair curves. The reference traces .hair strands
// as OptiX round Catmull-Rom curves (owlCurvesSetDegree(…, 3, forceCaps),
// render_path_tracing.cu:223, render_hair_msnn.cu:346, render_nrc.cu:409),
// whose intersector is closed. The primitive follows the reference's own
// curve code: the cubic through four control points (x, y, z, radius) with
// the Catmull-Rom matrix of curve_utils.cuh:57-71, and the swept-circle
// surface whose normal curve_utils.cuh:159-197 computes (points at distance
// r(u) from P(u) in the plane normal to P'(u)); flat end caps close a strand
// where forceCaps is set.
//
//   crv_intersect  the GPU build's iteration restated in the same IEEE FP32
//                  operations (compiled without contraction), so traversal
//                  results can be compared bit for bit;
//   crv_exact      an independent FP64 search for the first crossing of the
//                  surface along the ray (the closest centre-line point of
//                  every ray point decides inside / outside), which pins the
//                  iteration to the surface's definition.
#include <algorithm>
#include <cmath>
#include <limits>

#include "orc_scene.h"

namespace orc {

namespace {

constexpr int kSeeds = 4;
constexpr int kIters = 12;
constexpr float kUTol = 1e-5f;

struct Cub {
  float a, b, c, d;
};

Cub catrom(float q0, float q1, float q2, float q3) {
  Cub k;
  k.a = ((q3 - q0) + 3.f * (q1 - q2)) * 0.5f;
  k.b = ((2.f * q0 - 5.f * q1) + (4.f * q2 - q3)) * 0.5f;
  k.c = (q2 - q0) * 0.5f;
  k.d = q1;
  return k;
}
float at(const Cub& k, float u) { return ((k.a * u + k.b) * u + k.c) * u + k.d; }
float d1(const Cub& k, float u) { return (3.f * k.a * u + 2.f * k.b) * u + k.c; }
float d2(const Cub& k, float u) { return 6.f * k.a * u + 2.f * k.b; }

struct F3 {
  float x, y, z;
};
struct Seg {
  Cub x, y, z, r;
  F3 pos(float u) const { return {at(x, u), at(y, u), at(z, u)}; }
  F3 vel(float u) const { return {d1(x, u), d1(y, u), d1(z, u)}; }
  F3 acc(float u) const { return {d2(x, u), d2(y, u), d2(z, u)}; }
};
Seg make_seg(const float* q) {
  Seg s;
  s.x = catrom(q[0], q[4], q[8], q[12]);
  s.y = catrom(q[1], q[5], q[9], q[13]);
  s.z = catrom(q[2], q[6], q[10], q[14]);
  s.r = catrom(q[3], q[7], q[11], q[15]);
  return s;
}
float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float lenf(F3 a) { return std::sqrt(dotf(a, a)); }

bool cyl_first(const float o[3], const float d[3], F3 c, F3 v, float r, float* t_out) {
  const float vv = v.x * v.x + v.y * v.y + v.z * v.z;
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float tc = ((c.x - o[0]) * d[0] + (c.y - o[1]) * d[1] + (c.z - o[2]) * d[2]) / dd;
  const float qx = (o[0] + tc * d[0]) - c.x, qy = (o[1] + tc * d[1]) - c.y, qz = (o[2] + tc * d[2]) - c.z;
  const float dv = d[0] * v.x + d[1] * v.y + d[2] * v.z;
  const float qv = qx * v.x + qy * v.y + qz * v.z;
  const float A = dd * vv - dv * dv;
  const float B = (d[0] * qx + d[1] * qy + d[2] * qz) * vv - dv * qv;
  const float C = (qx * qx + qy * qy + qz * qz) * vv - qv * qv - r * r * vv;
  if (!(A > 0.f)) return false;
  const float disc = B * B - A * C;
  if (!(disc >= 0.f)) return false;
  const float sq = std::sqrt(disc);
  const float qq = (B >= 0.f) ? -(B + sq) : -(B - sq);
  float s0 = qq / A;
  float s1 = (qq != 0.f) ? C / qq : s0;
  if (s1 < s0) std::swap(s0, s1);
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

// cyl_first, except that a ray passing outside the cylinder gives the point
// nearest its axis (the iteration's target, not a hit)
bool cyl_iter(const float o[3], const float d[3], F3 c, F3 v, float r, float* t_out) {
  const float vv = v.x * v.x + v.y * v.y + v.z * v.z;
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float tc = ((c.x - o[0]) * d[0] + (c.y - o[1]) * d[1] + (c.z - o[2]) * d[2]) / dd;
  const float qx = (o[0] + tc * d[0]) - c.x, qy = (o[1] + tc * d[1]) - c.y, qz = (o[2] + tc * d[2]) - c.z;
  const float dv = d[0] * v.x + d[1] * v.y + d[2] * v.z;
  const float qv = qx * v.x + qy * v.y + qz * v.z;
  const float A = dd * vv - dv * dv;
  const float B = (d[0] * qx + d[1] * qy + d[2] * qz) * vv - dv * qv;
  const float C = (qx * qx + qy * qy + qz * qz) * vv - qv * qv - r * r * vv;
  if (!(A > 0.f)) return false;
  const float disc = B * B - A * C;
  if (!(disc >= 0.f)) {
    *t_out = tc - B / A;
    return true;
  }
  const float sq = std::sqrt(disc);
  const float qq = (B >= 0.f) ? -(B + sq) : -(B - sq);
  float s0 = qq / A;
  float s1 = (qq != 0.f) ? C / qq : s0;
  if (s1 < s0) std::swap(s0, s1);
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

bool disk(const float o[3], const float d[3], F3 c, F3 n, float r, float* t_out) {
  const float dn = d[0] * n.x + d[1] * n.y + d[2] * n.z;
  if (!(dn != 0.f)) return false;
  const float t = ((c.x - o[0]) * n.x + (c.y - o[1]) * n.y + (c.z - o[2]) * n.z) / dn;
  if (!(t > 0.f)) return false;
  const float px = (o[0] + t * d[0]) - c.x, py = (o[1] + t * d[1]) - c.y, pz = (o[2] + t * d[2]) - c.z;
  if (!(px * px + py * py + pz * pz <= r * r)) return false;
  *t_out = t;
  return true;
}

}  // namespace

float crv_radius_bound(const float* q) {
  const float w0 = q[3], w1 = q[7], w2 = q[11], w3 = q[15];
  const float b1 = w1 + (w2 - w0) * (1.f / 6.f);
  const float b2 = w2 - (w3 - w1) * (1.f / 6.f);
  return std::fmax(std::fmax(std::fmax(w1, w2), std::fmax(b1, b2)), 0.f);
}

void crv_box(const float* q, float lo[3], float hi[3]) {
  const float rmax = crv_radius_bound(q);
  for (int a = 0; a < 3; ++a) {
    const float b0 = q[4 + a], b3 = q[8 + a];
    const float b1 = b0 + (q[8 + a] - q[a]) * (1.f / 6.f), b2 = b3 - (q[12 + a] - q[4 + a]) * (1.f / 6.f);
    const float l = std::min(std::min(b0, b1), std::min(b2, b3));
    const float h = std::max(std::max(b0, b1), std::max(b2, b3));
    const float pad = 1.01f * rmax + 1e-5f * (std::fabs(l) + std::fabs(h)) + 1e-7f;
    lo[a] = l - pad;
    hi[a] = h + pad;
  }
}

bool crv_intersect(const float o[3], const float d[3], const float* q, int flags, float* t_out, float* u_out) {
  const Seg s = make_seg(q);
  const float rmax = crv_radius_bound(q);
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  float best_t = std::numeric_limits<float>::infinity(), best_u = 0.f;
  const float h = 1.f / (float)kSeeds;
  F3 e0 = s.pos(0.f), a0 = s.acc(0.f);
  for (int k = 0; k < kSeeds; ++k) {
    const float u0 = (float)k * h, u1 = (float)(k + 1) * h;
    const F3 e1 = s.pos(u1), a1 = s.acc(u1);
    const float m = std::fmax(lenf(a0), lenf(a1));
    const float R = rmax + (h * h * 0.125f) * m;
    const F3 ev{e1.x - e0.x, e1.y - e0.y, e1.z - e0.z};
    const float ee = dotf(ev, ev);
    const float de = d[0] * ev.x + d[1] * ev.y + d[2] * ev.z;
    const F3 w0{o[0] - e0.x, o[1] - e0.y, o[2] - e0.z};
    const float dw = d[0] * w0.x + d[1] * w0.y + d[2] * w0.z;
    const float ew = dotf(ev, w0);
    const float den = dd * ee - de * de;
    float sc = den > 0.f ? (dd * ew - de * dw) / den : 0.f;
    sc = std::fmin(std::fmax(sc, 0.f), 1.f);
    const float tcl = (sc * de - dw) / dd;
    const float px = (o[0] + tcl * d[0]) - (e0.x + sc * ev.x);
    const float py = (o[1] + tcl * d[1]) - (e0.y + sc * ev.y);
    const float pz = (o[2] + tcl * d[2]) - (e0.z + sc * ev.z);
    const float dist2 = px * px + py * py + pz * pz;
    const bool near = dist2 <= R * R;
    e0 = e1;
    a0 = a1;
    if (!near) continue;
    float u = u0 + sc * h;
    float t = 0.f;
    bool ok = false;
    float up = 0.f, fp = 0.f;
    bool prev = false;
    for (int it = 0; it < kIters; ++it) {
      const F3 c = s.pos(u), v = s.vel(u);
      const float r = at(s.r, u);
      if (!cyl_iter(o, d, c, v, r, &t)) {
        ok = false;
        break;
      }
      const F3 x{(o[0] + t * d[0]) - c.x, (o[1] + t * d[1]) - c.y, (o[2] + t * d[2]) - c.z};
      const float f = dotf(x, v);
      float du;
      if (prev && f != fp) {
        du = -f * (u - up) / (f - fp);
      } else {
        const float j = dotf(x, s.acc(u)) - dotf(v, v);
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
      ok = std::fabs(du) <= kUTol;
      if (ok || !(u >= -h && u <= 1.f + h)) break;
    }
    if (!ok || !(u >= 0.f && u <= 1.f)) continue;
    if (!cyl_first(o, d, s.pos(u), s.vel(u), at(s.r, u), &t)) continue;
    if (t < best_t) {
      best_t = t;
      best_u = u;
    }
  }
  if (flags & 1) {
    float t;
    if (disk(o, d, s.pos(0.f), s.vel(0.f), at(s.r, 0.f), &t) && t < best_t) {
      best_t = t;
      best_u = 0.f;
    }
  }
  if (flags & 2) {
    float t;
    if (disk(o, d, s.pos(1.f), s.vel(1.f), at(s.r, 1.f), &t) && t < best_t) {
      best_t = t;
      best_u = 1.f;
    }
  }
  if (!(best_t < std::numeric_limits<float>::infinity())) return false;
  *t_out = best_t;
  *u_out = best_u;
  return true;
}

// ---- FP64 brute force ----------------------------------------------------
namespace {

struct CubD {
  double a, b, c, d;
};
CubD catrom_d(double q0, double q1, double q2, double q3) {
  // P(u) = 1/2 [(-q0 + 3q1 - 3q2 + q3) u^3 + (2q0 - 5q1 + 4q2 - q3) u^2 + (q2 - q0) u + 2 q1]
  return {0.5 * (-q0 + 3 * q1 - 3 * q2 + q3), 0.5 * (2 * q0 - 5 * q1 + 4 * q2 - q3), 0.5 * (q2 - q0), q1};
}
struct SegD {
  CubD k[4];  // x, y, z, r
  void pos(double u, double p[3]) const {
    for (int a = 0; a < 3; ++a) p[a] = ((k[a].a * u + k[a].b) * u + k[a].c) * u + k[a].d;
  }
  void vel(double u, double p[3]) const {
    for (int a = 0; a < 3; ++a) p[a] = (3 * k[a].a * u + 2 * k[a].b) * u + k[a].c;
  }
  void acc(double u, double p[3]) const {
    for (int a = 0; a < 3; ++a) p[a] = 6 * k[a].a * u + 2 * k[a].b;
  }
  double r(double u) const { return ((k[3].a * u + k[3].b) * u + k[3].c) * u + k[3].d; }
};

// Parameter range of the extended cubic the inside test uses: entries
// through the side wall are found on the tube continued past the segment's
// ends and accepted when their foot lies in [0, 1], so that an open end is not
// a way into the tube (a neighbouring segment continues the strand there).
constexpr double kExtLo = -0.25, kExtHi = 1.25;

// The centre-line parameter closest to X over the extended range (a grid of
// 193 points, then Newton on the stationarity condition, clamped).
double foot(const SegD& s, const double X[3], double* dist2) {
  double best_u = 0, best = std::numeric_limits<double>::infinity();
  for (int i = 0; i <= 192; ++i) {
    const double u = kExtLo + (kExtHi - kExtLo) * i / 192.0;
    double p[3];
    s.pos(u, p);
    const double g = (X[0] - p[0]) * (X[0] - p[0]) + (X[1] - p[1]) * (X[1] - p[1]) + (X[2] - p[2]) * (X[2] - p[2]);
    if (g < best) {
      best = g;
      best_u = u;
    }
  }
  double u = best_u;
  for (int it = 0; it < 30; ++it) {
    double p[3], v[3], a[3];
    s.pos(u, p);
    s.vel(u, v);
    s.acc(u, a);
    const double w[3] = {X[0] - p[0], X[1] - p[1], X[2] - p[2]};
    const double g1 = -(w[0] * v[0] + w[1] * v[1] + w[2] * v[2]);
    const double g2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - (w[0] * a[0] + w[1] * a[1] + w[2] * a[2]);
    if (!(g2 > 0)) break;
    const double nu = std::min(kExtHi, std::max(kExtLo, u - g1 / g2));
    const bool done = std::fabs(nu - u) < 1e-15;
    u = nu;
    if (done) break;
  }
  double p[3];
  s.pos(u, p);
  *dist2 = (X[0] - p[0]) * (X[0] - p[0]) + (X[1] - p[1]) * (X[1] - p[1]) + (X[2] - p[2]) * (X[2] - p[2]);
  return u;
}

// Signed clearance of X from the extended tube: distance to the centre line
// minus the radius at the foot point; +inf where the foot is clamped.
double clearance(const SegD& s, const double X[3], double* uf) {
  double d2;
  const double u = foot(s, X, &d2);
  *uf = u;
  if (u <= kExtLo || u >= kExtHi) return std::numeric_limits<double>::infinity();
  return std::sqrt(d2) - s.r(u);
}

// Ray / flat end cap at parameter ue (0 or 1).
bool cap_hit(const SegD& s, const double o[3], const double d[3], double ue, double* t) {
  double c[3], n[3];
  s.pos(ue, c);
  s.vel(ue, n);
  const double dn = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
  if (dn == 0) return false;
  const double tt = ((c[0] - o[0]) * n[0] + (c[1] - o[1]) * n[1] + (c[2] - o[2]) * n[2]) / dn;
  if (!(tt > 0)) return false;
  double r2 = 0;
  for (int a = 0; a < 3; ++a) r2 += (o[a] + tt * d[a] - c[a]) * (o[a] + tt * d[a] - c[a]);
  if (!(r2 <= s.r(ue) * s.r(ue))) return false;
  *t = tt;
  return true;
}

}  // namespace

void crv_eval(const float* q, double u, int prec, double out[8]) {
  if (prec == 1) {
    const Seg s = make_seg(q);
    const float uf = (float)u;
    const F3 p = s.pos(uf), v = s.vel(uf);
    const double o[8] = {p.x, p.y, p.z, at(s.r, uf), v.x, v.y, v.z, d1(s.r, uf)};
    for (int i = 0; i < 8; ++i) out[i] = o[i];
    return;
  }
  SegD s;
  for (int a = 0; a < 4; ++a) s.k[a] = catrom_d(q[a], q[4 + a], q[8 + a], q[12 + a]);
  double p[3], v[3];
  s.pos(u, p);
  s.vel(u, v);
  const double dr = (3 * s.k[3].a * u + 2 * s.k[3].b) * u + s.k[3].c;
  const double o[8] = {p[0], p[1], p[2], s.r(u), v[0], v[1], v[2], dr};
  for (int i = 0; i < 8; ++i) out[i] = o[i];
}

bool crv_exact(const double o[3], const double d[3], const float* q, int flags, double* t_out, double* u_out,
               double* min_clear) {
  SegD s;
  for (int a = 0; a < 4; ++a) s.k[a] = catrom_d(q[a], q[4 + a], q[8 + a], q[12 + a]);
  double best_t = std::numeric_limits<double>::infinity(), best_u = 0;
  if (min_clear) *min_clear = std::numeric_limits<double>::infinity();
  for (int e = 0; e < 2; ++e) {
    double tc;
    if ((flags & (1 << e)) && cap_hit(s, o, d, (double)e, &tc) && tc < best_t) {
      best_t = tc;
      best_u = e;
    }
  }
  // the ray span inside a box around the extended segment: sampled points
  // of the extended centre line, padded by the largest radius there
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmin = 1e300, rmax = 0;
  for (int i = 0; i <= 256; ++i) {
    const double u = kExtLo + (kExtHi - kExtLo) * i / 256.0;
    double p[3];
    s.pos(u, p);
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
    if (u >= 0 && u <= 1) rmin = std::min(rmin, s.r(u));
    rmax = std::max(rmax, s.r(u));
  }
  double t0 = 0.0, t1 = std::numeric_limits<double>::infinity();
  for (int a = 0; a < 3; ++a) {
    const double l = lo[a] - 2 * rmax, h = hi[a] + 2 * rmax;
    if (d[a] == 0.0) {
      if (o[a] < l || o[a] > h) t1 = -1;
      continue;
    }
    double ta = (l - o[a]) / d[a], tb = (h - o[a]) / d[a];
    if (ta > tb) std::swap(ta, tb);
    t0 = std::max(t0, ta);
    t1 = std::min(t1, tb);
  }
  if (t0 < t1 && rmin > 0) {
    const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double step = rmin / 32.0 / dl;
    auto inside = [&](double t, double* cl, double* uf) {
      const double x[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
      *cl = clearance(s, x, uf);
      return *cl < 0;
    };
    double cl, uf;
    double tp = t0;
    const bool start_in = inside(tp, &cl, &uf) && t0 == 0.0;
    bool in_p = cl < 0;
    if (min_clear && std::isfinite(cl)) *min_clear = std::min(*min_clear, cl);
    const int n = (int)std::ceil((t1 - t0) / step) + 1;
    bool first_event = true;
    for (int i = 1; i <= n && tp < best_t; ++i) {
      const double t = std::min(t1, t0 + i * step);
      const bool in = inside(t, &cl, &uf);
      if (min_clear && std::isfinite(cl)) *min_clear = std::min(*min_clear, cl);
      if (in != in_p) {
        double a = tp, b = t;
        for (int k = 0; k < 80; ++k) {
          const double m = 0.5 * (a + b);
          double c2, u2;
          if (inside(m, &c2, &u2) == in_p) a = m;
          else b = m;
        }
        const double tc = 0.5 * (a + b);
        double c3, uc;
        inside(in ? b : a, &c3, &uc);  // foot on the inner side of the crossing
        // an entry, or the first exit of a ray that starts inside
        const bool counts = in || (start_in && first_event);
        first_event = false;
        if (counts && tc > 0 && uc >= 0 && uc <= 1 && tc < best_t) {
          best_t = tc;
          best_u = uc;
        }
      }
      tp = t;
      in_p = in;
      if (t >= t1) break;
    }
  }
  if (!(best_t < std::numeric_limits<double>::infinity())) return false;
  *t_out = best_t;
  *u_out = best_u;
  return true;
}

}  // namespace orc
