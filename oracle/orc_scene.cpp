// ORACLE - test infrastructure only (see orc_math.h).
#include "orc_scene.h"

#include <algorithm>
#include <cmath>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// scene.cpp:349-425
void env_tables(EnvTables& e, int mode) {
  const int W = e.W, H = e.H, cw = W + 1;
  e.cpdf.assign((size_t)cw * H, 0.f);
  e.ccdf.assign((size_t)cw * H, 0.f);
  e.mpdf.assign(H + 1, 0.f);
  e.mcdf.assign(H + 1, 0.f);
  for (int y = 0; y < H; ++y) {
    const float sinTheta = sinf(3.14159f * (y + 0.5f) / H);
    float* pdf = &e.cpdf[(size_t)y * cw];
    float* cdf = &e.ccdf[(size_t)y * cw];
    for (int x = 0; x < W; ++x) {
      const float* t = &e.rgba[4 * ((size_t)y * W + x)];
      pdf[x] = (t[0] + t[1] + t[2]) * (1.0f / 3.0f) * sinTheta;
    }
    if (mode == 1) {
      cdf[0] = 0.f;
      for (int x = 1; x < W; ++x) cdf[x] = cdf[x - 1] + pdf[x - 1] / W;
      const float total = cdf[W - 1] + pdf[W - 1] / W;
      pdf[W] = total;
      if (total > 0.f) {
        const float inv = 1.0f / total;
        for (int x = 1; x < W; ++x) cdf[x] *= inv;
      }
    } else {
      std::vector<double> run(W);
      double acc = 0.0;
      for (int x = 0; x < W; ++x) {
        run[x] = acc;
        acc += (double)pdf[x] / (double)W;
      }
      const double total = acc;
      pdf[W] = (float)total;
      cdf[0] = 0.f;
      for (int x = 1; x < W; ++x) cdf[x] = (float)(total > 0.0 ? run[x] / total : run[x]);
    }
    cdf[W] = 1.f;
  }
  float* mp = e.mpdf.data();
  float* mc = e.mcdf.data();
  for (int y = 0; y < H; ++y) mp[y] = e.cpdf[(size_t)y * cw + W];
  if (mode == 1) {
    mc[0] = 0.f;
    for (int i = 1; i < H; ++i) mc[i] = mc[i - 1] + mp[i - 1] / H;
    const float total = mc[H - 1] + mp[H - 1] / H;
    mp[H] = total;
    if (total > 0.f)
      for (int i = 1; i < H; ++i) mc[i] /= total;
  } else {
    std::vector<double> run(H);
    double acc = 0.0;
    for (int y = 0; y < H; ++y) {
      run[y] = acc;
      acc += (double)mp[y] / (double)H;
    }
    mp[H] = (float)acc;
    mc[0] = 0.f;
    for (int y = 1; y < H; ++y) mc[y] = (float)(acc > 0.0 ? run[y] / acc : run[y]);
  }
  mc[H] = 1.f;
}

// ---------------------------------------------------------------------------
// Intersection primitives (FP32, compiled with -ffp-contract=off).
bool seg_intersect(const float o[3], const float d[3], const float* s, float* t_out, float* u_out) {
  const float p0x = s[0], p0y = s[1], p0z = s[2], r = s[3];
  const float p1x = s[4], p1y = s[5], p1z = s[6];
  const float ax = p1x - p0x, ay = p1y - p0y, az = p1z - p0z;
  const float aa = ax * ax + ay * ay + az * az;
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float cx = (p0x + p1x) * 0.5f, cy = (p0y + p1y) * 0.5f, cz = (p0z + p1z) * 0.5f;
  const float tc = ((cx - o[0]) * d[0] + (cy - o[1]) * d[1] + (cz - o[2]) * d[2]) / dd;
  const float qx = (o[0] + tc * d[0]) - p0x, qy = (o[1] + tc * d[1]) - p0y, qz = (o[2] + tc * d[2]) - p0z;
  const float da = d[0] * ax + d[1] * ay + d[2] * az;
  const float qa = qx * ax + qy * ay + qz * az;
  const float A = dd * aa - da * da;
  const float B = (d[0] * qx + d[1] * qy + d[2] * qz) * aa - da * qa;
  const float C = (qx * qx + qy * qy + qz * qz) * aa - qa * qa - r * r * aa;
  if (!(A > 0.f)) return false;
  const float disc = B * B - A * C;
  if (!(disc >= 0.f)) return false;
  const float sq = sqrtf(disc);
  const float qq = (B >= 0.f) ? -(B + sq) : -(B - sq);
  float s0 = qq / A;
  float s1 = (qq != 0.f) ? C / qq : s0;
  if (s1 < s0) std::swap(s0, s1);
  const float inv_aa = 1.f / aa;
  const float roots[2] = {s0, s1};
  for (float sr : roots) {
    const float t = tc + sr;
    const float u = (qa + sr * da) * inv_aa;
    if (t > 0.f && u >= 0.f && u <= 1.f) {
      *t_out = t;
      *u_out = u;
      return true;
    }
  }
  return false;
}

bool sphere_intersect(const float o[3], const float d[3], const float c[3], float radius, float* t_out) {
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float tc = ((c[0] - o[0]) * d[0] + (c[1] - o[1]) * d[1] + (c[2] - o[2]) * d[2]) / dd;
  const float qx = (o[0] + tc * d[0]) - c[0], qy = (o[1] + tc * d[1]) - c[1], qz = (o[2] + tc * d[2]) - c[2];
  const float C = (qx * qx + qy * qy + qz * qz) - radius * radius;
  const float disc = -dd * C;
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

bool tri_intersect(const float o[3], const float d[3], const float* v, float* t_out, float* u_out, float* v_out) {
  const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
  const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
  const float px = d[1] * e2z - d[2] * e2y, py = d[2] * e2x - d[0] * e2z, pz = d[0] * e2y - d[1] * e2x;
  const float det = e1x * px + e1y * py + e1z * pz;
  if (!(det != 0.f)) return false;
  const float inv = 1.f / det;
  const float sx = o[0] - v[0], sy = o[1] - v[1], sz = o[2] - v[2];
  const float u = (sx * px + sy * py + sz * pz) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  const float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
  const float w = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
  if (!(w >= 0.f && u + w <= 1.f)) return false;
  const float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (!(t > 0.f)) return false;
  *t_out = t;
  *u_out = u;
  *v_out = w;
  return true;
}

// ---------------------------------------------------------------------------
// CPU BVH (median split, leaves of up to 4 segments). Boxes are conservative
// (padded) and pruning keeps every candidate with entry distance <= best t,
// so the result equals the brute-force minimum over (t, id).
void Bvh::build(const float* s, int n) {
  seg = s;
  nseg = n;
  nodes.clear();
  idx.resize(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  if (n == 0) return;
  std::vector<float> cen(3 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) cen[3 * (size_t)i + k] = 0.5f * (s[8 * (size_t)i + k] + s[8 * (size_t)i + 4 + k]);
  struct Item {
    int node, first, count;
  };
  std::vector<Item> work;
  nodes.push_back(Node{});
  work.push_back({0, 0, n});
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    Node nd{};
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = std::numeric_limits<float>::infinity();
      nd.hi[k] = -std::numeric_limits<float>::infinity();
    }
    float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
    for (int j = it.first; j < it.first + it.count; ++j) {
      const float* q = s + 8 * (size_t)idx[j];
      const float r = q[3];
      for (int k = 0; k < 3; ++k) {
        const float pad = r + 2e-6f * (std::fabs(q[k]) + std::fabs(q[4 + k]) + r) + 1e-6f;
        nd.lo[k] = std::min(nd.lo[k], std::min(q[k], q[4 + k]) - pad);
        nd.hi[k] = std::max(nd.hi[k], std::max(q[k], q[4 + k]) + pad);
        clo[k] = std::min(clo[k], cen[3 * (size_t)idx[j] + k]);
        chi[k] = std::max(chi[k], cen[3 * (size_t)idx[j] + k]);
      }
    }
    if (it.count <= 4) {
      nd.first = it.first;
      nd.count = it.count;
      nd.left = nd.right = -1;
      nodes[it.node] = nd;
      continue;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    const int mid = it.first + it.count / 2;
    std::nth_element(idx.begin() + it.first, idx.begin() + mid, idx.begin() + it.first + it.count,
                     [&](int a, int b) { return cen[3 * (size_t)a + axis] < cen[3 * (size_t)b + axis]; });
    const int l = (int)nodes.size();
    nodes.push_back(Node{});
    const int r = (int)nodes.size();
    nodes.push_back(Node{});
    nd.left = l;
    nd.right = r;
    nd.count = 0;
    nodes[it.node] = nd;
    work.push_back({l, it.first, mid - it.first});
    work.push_back({r, mid, it.first + it.count - mid});
  }
}

void Bvh::build_curves(const float* c, int n) {
  boxes.assign((size_t)8 * n, 0.f);
  for (int k = 0; k < n; ++k) {
    float lo[3], hi[3];
    crv_box(c + (size_t)kCrvFloats * k, lo, hi);
    float* b = &boxes[8 * (size_t)k];
    for (int a = 0; a < 3; ++a) {
      b[a] = lo[a];
      b[4 + a] = hi[a];
    }
    b[7] = (float)k;
  }
  build(boxes.data(), n);
  crv = c;
}

bool Bvh::leaf(int id, const float o[3], const float d[3], int caps, float* t, float* u) const {
  if (crv) {
    const float* q = crv + (size_t)kCrvFloats * id;
    return crv_intersect(o, d, q, caps ? (int)q[17] : 0, t, u);
  }
  return seg_intersect(o, d, seg + 8 * (size_t)id, t, u);
}

static inline bool box_hit(const Bvh::Node& n, const double o[3], const double inv[3], double tmax, double* tn) {
  double t0 = 0.0, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double a = (n.lo[k] - o[k]) * inv[k], b = (n.hi[k] - o[k]) * inv[k];
    if (a > b) std::swap(a, b);
    t0 = std::max(t0, a);
    t1 = std::min(t1, b);
  }
  *tn = t0;
  return t0 <= t1 * (1.0 + 1e-9) + 1e-12;
}

bool Bvh::closest(const float o[3], const float d[3], float* t, int* id, float* u, int caps) const {
  if (nseg == 0) return false;
  double od[3] = {o[0], o[1], o[2]}, inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (d[k] != 0.f ? (double)d[k] : 1e-300);
  float best_t = std::numeric_limits<float>::infinity(), best_u = 0.f;
  int best_id = -1;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node& n = nodes[stack[--sp]];
    double tn;
    if (!box_hit(n, od, inv, std::isinf(best_t) ? 1e300 : (double)best_t * (1.0 + 1e-6) + 1e-6, &tn)) continue;
    if (n.count > 0) {
      for (int j = n.first; j < n.first + n.count; ++j) {
        const int sid = idx[j];
        float tt, uu;
        if (leaf(sid, o, d, caps, &tt, &uu)) {
          if (tt < best_t || (tt == best_t && sid < best_id)) {
            best_t = tt;
            best_id = sid;
            best_u = uu;
          }
        }
      }
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  if (best_id < 0) return false;
  *t = best_t;
  *id = best_id;
  *u = best_u;
  return true;
}

bool Bvh::any(const float o[3], const float d[3], int caps) const {
  if (nseg == 0) return false;
  double od[3] = {o[0], o[1], o[2]}, inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (d[k] != 0.f ? (double)d[k] : 1e-300);
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node& n = nodes[stack[--sp]];
    double tn;
    if (!box_hit(n, od, inv, 1e300, &tn)) continue;
    if (n.count > 0) {
      for (int j = n.first; j < n.first + n.count; ++j) {
        float tt, uu;
        if (leaf(idx[j], o, d, caps, &tt, &uu)) return true;
      }
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// Integrator
template <class R>
struct Inter {
  bool hit = false, surface = false;
  Vec<R> p, n, wo, wi, wo_local, wi_local;
  Vec<R> to_local[3];
  R radius = 0;
  Vec<R> color;  // sigma_a (hair) or albedo (surface)
  R alpha = 0;
  Vec<R> Le;
  int prim = -1;  // segment id, -2 for the head sphere
  Vec<R> color0;  // sigma_a used for sampling (frozen mode), else = color
};

template <class R>
struct Ctx {
  const SceneDesc& s;
  const Bvh& bvh;
  HairParams<R> hp;
  HairShading<R> hs;
  bool frozen = false;  // sampling uses hp0 / hs0 / color0
  HairParams<R> hp0;
  HairShading<R> hs0;
  int caps = 1;  // round curves' end caps: plain path tracing only (render_path_tracing.cu:223)
};

template <class R> inline Vec<R> apply_mat(const Vec<R> m[3], Vec<R> v) { return mk(dot(m[0], v), dot(m[1], v), dot(m[2], v)); }
template <class R> inline Vec<R> apply_T(const Vec<R> m[3], Vec<R> v) {  // to_world = transpose
  return mk(m[0].x * v.x + m[1].x * v.y + m[2].x * v.z, m[0].y * v.x + m[1].y * v.y + m[2].y * v.z,
            m[0].z * v.x + m[1].z * v.y + m[2].z * v.z);
}

// optix_common.cuh:10-24 (bilinear, clamp)
template <class R>
Vec<R> env_radiance(const SceneDesc& s, Vec<R> dir) {
  const EnvTables& e = *s.env;
  const R theta = std::acos(std::fmin((R)1, std::fmax((R)-1, dir.z)));
  R phi = std::atan2(dir.y, dir.x);
  if (phi < 0) phi += (R)2 * Pi<R>();
  phi += (R)s.env_rot;
  if (phi > TwoPiEnv<R>()) phi -= TwoPiEnv<R>();
  const R x = phi / ((R)2 * Pi<R>()), y = theta / Pi<R>();
  const R xb = x * (R)e.W - (R)0.5, yb = y * (R)e.H - (R)0.5;
  const R fx = std::floor(xb), fy = std::floor(yb);
  const R a = xb - fx, b = yb - fy;
  const int x0 = (int)fx, y0 = (int)fy;
  auto cl = [](int v, int hi) { return std::min(std::max(v, 0), hi - 1); };
  const int xa = cl(x0, e.W), xb1 = cl(x0 + 1, e.W), ya = cl(y0, e.H), yb1 = cl(y0 + 1, e.H);
  auto T = [&](int xx, int yy, int c) { return (R)e.rgba[4 * ((size_t)yy * e.W + xx) + c]; };
  Vec<R> r;
  R* rc[3] = {&r.x, &r.y, &r.z};
  for (int c = 0; c < 3; ++c)
    *rc[c] = ((R)1 - a) * ((R)1 - b) * T(xa, ya, c) + a * ((R)1 - b) * T(xb1, ya, c) + ((R)1 - a) * b * T(xa, yb1, c) +
             a * b * T(xb1, yb1, c);
  return (R)s.env_scale * r;
}

// optix_common.cuh:26-64
template <class R>
R env_pdf(const SceneDesc& s, Vec<R> wi) {
  if (!s.env_pdf) return (R)1 / ((R)4 * Pi<R>());
  const EnvTables& e = *s.env;
  const R theta = std::acos(std::fmin((R)1, std::fmax((R)-1, wi.z)));
  R phi = std::atan2(wi.y, wi.x);
  if (phi < 0) phi += (R)2 * Pi<R>();
  phi += (R)s.env_rot;
  if (phi > TwoPiEnv<R>()) phi -= TwoPiEnv<R>();
  const R u = phi / ((R)2 * Pi<R>()), v = theta / Pi<R>();
  const int iu = std::min(std::max((int)(u * e.W), 0), e.W - 1);
  const int iv = std::min(std::max((int)(v * e.H), 0), e.H - 1);
  const int cw = e.W + 1;
  const R last_u = e.cpdf[(size_t)iv * cw + e.W], last_v = e.mpdf[e.H];
  const R denom = ((R)2 * Pi<R>() * Pi<R>() * std::sin(theta)) * last_u * last_v;
  const R cu = e.cpdf[(size_t)iv * cw + iu], cv = e.mpdf[iv];
  return denom != 0 ? (cu * cv) / denom : (R)0;
}

static int lower_bound_m1(const float* cdf, int count, float u) {
  int first = 0;
  while (count > 0) {
    const int step = count >> 1, mid = first + step;
    if (cdf[mid] < u) {
      first = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return std::max(0, first - 1);
}

// optix_common.cuh:66-149
template <class R>
Vec<R> env_sample(const SceneDesc& s, Lcg& rng, Vec<R>* wi, R* pdf) {
  const float ux = rng.next(), uy = rng.next();
  if (!s.env_pdf) {
    const R z = (R)1 - (R)2 * (R)ux;
    const R r = std::sqrt(std::fmax((R)0, (R)1 - z * z));
    const R ph = (R)2 * Pi<R>() * (R)uy;
    *wi = nrm(mk(r * std::cos(ph), r * std::sin(ph), z));
    *pdf = (R)1 / ((R)4 * Pi<R>());
    return env_radiance(s, *wi);
  }
  const EnvTables& e = *s.env;
  const int cw = e.W + 1;
  const int iv = lower_bound_m1(e.mcdf.data(), e.H, uy);
  const R cv = e.mcdf[iv], cnv = e.mcdf[iv + 1];
  const R v = ((R)iv + (cnv - (R)uy) / (cnv - cv)) / (R)e.H;
  const float* row = &e.ccdf[(size_t)iv * cw];
  const int iu = lower_bound_m1(row, e.W, ux);
  const R cu = row[iu], cnu = row[iu + 1];
  const R u = ((R)iu + (cnu - (R)ux) / (cnu - cu)) / (R)e.W;
  // radiance at the sampled texture coordinate (u, v): reuse env_radiance's
  // filter by building the direction's texture coordinates directly
  const R xb = u * (R)e.W - (R)0.5, yb = v * (R)e.H - (R)0.5;
  const R fx = std::floor(xb), fy = std::floor(yb);
  const R a = xb - fx, b = yb - fy;
  auto cl = [](int vv, int hi) { return std::min(std::max(vv, 0), hi - 1); };
  const int xa = cl((int)fx, e.W), xb1 = cl((int)fx + 1, e.W), ya = cl((int)fy, e.H), yb1 = cl((int)fy + 1, e.H);
  auto T = [&](int xx, int yy, int c) { return (R)e.rgba[4 * ((size_t)yy * e.W + xx) + c]; };
  Vec<R> L;
  R* lc[3] = {&L.x, &L.y, &L.z};
  for (int c = 0; c < 3; ++c)
    *lc[c] = ((R)1 - a) * ((R)1 - b) * T(xa, ya, c) + a * ((R)1 - b) * T(xb1, ya, c) + ((R)1 - a) * b * T(xa, yb1, c) +
             a * b * T(xb1, yb1, c);
  L = (R)s.env_scale * L;
  const R last_u = e.cpdf[(size_t)iv * cw + e.W], last_v = e.mpdf[e.H];
  const R pu = e.cpdf[(size_t)iv * cw + iu], pv = e.mpdf[iv];
  const R theta = Pi<R>() * v;
  const R st = std::sin(theta);
  const R denom = ((R)2 * Pi<R>() * Pi<R>() * st) * last_u * last_v;
  *pdf = denom != 0 ? (pu * pv) / denom : (R)0;
  R phi = (R)2 * Pi<R>() * u - (R)s.env_rot;
  if (phi < 0) phi = TwoPiEnv<R>() + phi;
  *wi = mk(st * std::cos(phi), st * std::sin(phi), std::cos(theta));
  return L;
}

template <class R>
void setup_hit(const Ctx<R>& c, Vec<R> o, Vec<R> d, float t, int prim, float u, Inter<R>& si) {
  const SceneDesc& s = c.s;
  si.hit = true;
  si.prim = prim;
  si.wo = (R)-1 * d;
  Vec<R> p = o + (R)t * d;
  if (prim == -2 || prim <= -3) {  // head sphere or mesh triangle (triangleMeshCH, optix_common.cuh:527-556)
    if (prim == -2) {
      const Vec<R> cc = mk((R)s.sphere[0], (R)s.sphere[1], (R)s.sphere[2]);
      si.n = nrm(p - cc);
      si.p = cc + (R)s.sphere[3] * si.n;
      si.color = mk((R)s.sphere_albedo[0], (R)s.sphere_albedo[1], (R)s.sphere_albedo[2]);
      si.alpha = (R)s.sphere_alpha;
    } else {
      // barycentric position and normal (barycentricInterpolate, utils.cuh:92-101)
      const int k = -3 - prim;
      const float* tv = s.tri_v + 9 * (size_t)k;
      const float* tn = s.tri_n + 9 * (size_t)k;
      const float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
      float tt = 0.f, bu = 0.f, bv = 0.f;
      (void)tri_intersect(of, df, tv, &tt, &bu, &bv);
      const R w0 = (R)(1.f - bu - bv), w1 = (R)bu, w2 = (R)bv;
      si.p = mk(w0 * (R)tv[0] + w1 * (R)tv[3] + w2 * (R)tv[6], w0 * (R)tv[1] + w1 * (R)tv[4] + w2 * (R)tv[7],
                w0 * (R)tv[2] + w1 * (R)tv[5] + w2 * (R)tv[8]);
      si.n = nrm(mk(w0 * (R)tn[0] + w1 * (R)tn[3] + w2 * (R)tn[6], w0 * (R)tn[1] + w1 * (R)tn[4] + w2 * (R)tn[7],
                    w0 * (R)tn[2] + w1 * (R)tn[5] + w2 * (R)tn[8]));
      si.color = mk((R)s.tri_kd[3 * k], (R)s.tri_kd[3 * k + 1], (R)s.tri_kd[3 * k + 2]);
      si.alpha = (R)1;  // mesh->alpha = 1 (model.cpp:323), clamped to [0.01, 1] and squared
    }
    const Vec<R> n = si.n;
    Vec<R> c1, c2;
    if (n.z < (R)-0.999999f) {  // orthonormalBasis, utils.cuh:232-259
      c1 = mk((R)0, (R)-1, (R)0);
      c2 = mk((R)-1, (R)0, (R)0);
    } else {
      const R a = (R)1 / ((R)1 + n.z);
      const R b = -n.x * n.y * a;
      c1 = nrm(mk((R)1 - n.x * n.x * a, b, -n.x));
      c2 = nrm(mk(b, (R)1 - n.y * n.y * a, -n.y));
    }
    si.to_local[0] = c1;
    si.to_local[1] = c2;
    si.to_local[2] = n;
    si.wo_local = nrm(apply_mat(si.to_local, si.wo));
    si.color0 = si.color;
    si.surface = true;
    si.radius = 0;
    return;
  }
  // hairCH + computeCurveIntersection (optix_common.cuh:558-599,
  // curve_utils.cuh:218-255, surfaceNormal :159-197): linear branch, or the
  // cubic Catmull-Rom branch for round curves
  const float* q = s.seg + 8 * (size_t)prim;
  Vec<R> n, X;
  R r;
  if (s.crv) {
    const float* w = s.crv + (size_t)kCrvFloats * prim;
    R A[4][4];  // power basis per coordinate (x, y, z, radius): a, b, c, d
    for (int k = 0; k < 4; ++k) {
      const R q0 = (R)w[k], q1 = (R)w[4 + k], q2 = (R)w[8 + k], q3 = (R)w[12 + k];
      A[k][0] = (-q0 + (R)3 * q1 - (R)3 * q2 + q3) / (R)2;
      A[k][1] = ((R)2 * q0 - (R)5 * q1 + (R)4 * q2 - q3) / (R)2;
      A[k][2] = (q2 - q0) / (R)2;
      A[k][3] = q1;
    }
    const R uu = (R)u;
    auto P = [&](int k, R x) { return ((A[k][0] * x + A[k][1]) * x + A[k][2]) * x + A[k][3]; };
    auto V = [&](int k, R x) {  // velocity4 with the triple-knot guard (curve_utils.cuh:95-101)
      if (x == (R)0) x = (R)0.000001f;
      if (x == (R)1) x = (R)0.999999f;
      return ((R)3 * A[k][0] * x + (R)2 * A[k][1]) * x + A[k][2];
    };
    auto Acc = [&](int k, R x) { return (R)6 * A[k][0] * x + (R)2 * A[k][1]; };
    const Vec<R> cp = mk(P(0, uu), P(1, uu), P(2, uu));
    const Vec<R> dv = mk(V(0, uu), V(1, uu), V(2, uu));
    r = P(3, uu);
    if (u == 0.f) {
      n = nrm((R)-1 * dv);
    } else if (u == 1.f) {
      n = nrm(dv);
    } else {
      const R dr = V(3, uu);
      R dd = dot(dv, dv);
      Vec<R> o1 = p - cp;
      o1 = o1 - (dot(o1, dv) / dd) * dv;
      o1 = (r / len(o1)) * o1;
      p = cp + o1;
      dd = dd - dot(mk(Acc(0, uu), Acc(1, uu), Acc(2, uu)), o1);
      n = nrm(dd * o1 - (dr * r) * dv);
    }
    X = nrm(dv);
  } else {
    const Vec<R> p0 = mk((R)q[0], (R)q[1], (R)q[2]), p1 = mk((R)q[4], (R)q[5], (R)q[6]);
    const Vec<R> dv = p1 - p0;
    r = (R)q[3];
    if (u == 0.f) {
      n = nrm((R)-1 * dv);
    } else if (u == 1.f) {
      n = nrm(dv);
    } else {
      const Vec<R> pa = p0 + (R)u * dv;
      const R dd = dot(dv, dv);
      Vec<R> o1 = p - pa;
      o1 = o1 - (dot(o1, dv) / dd) * dv;
      o1 = (r / len(o1)) * o1;
      p = pa + o1;
      n = nrm(dd * o1);
    }
    X = nrm(dv);
  }
  const Vec<R> Y = nrm(cross(si.wo, X));
  const Vec<R> Z = nrm(cross(X, Y));
  si.p = p;
  si.n = n;
  si.to_local[0] = X;
  si.to_local[1] = Y;
  si.to_local[2] = Z;
  si.wo_local = nrm(apply_mat(si.to_local, si.wo));
  si.radius = r;
  if (s.strand_sigma) {
    const int strand = (int)q[7];
    si.color = mk((R)s.strand_sigma[3 * strand], (R)s.strand_sigma[3 * strand + 1], (R)s.strand_sigma[3 * strand + 2]);
  } else {
    si.color = mk((R)s.sigma_a[0], (R)s.sigma_a[1], (R)s.sigma_a[2]);
  }
  si.color0 = si.color;
  if (c.frozen && s.strand_sigma0) {
    const int strand = (int)q[7];
    si.color0 = mk((R)s.strand_sigma0[3 * strand], (R)s.strand_sigma0[3 * strand + 1], (R)s.strand_sigma0[3 * strand + 2]);
  }
  si.surface = false;
}

template <class R>
Vec<R> bsdf_eval(const Ctx<R>& c, const Inter<R>& si, R* pdf) {
  if (si.surface) {
    *pdf = ggx_pdf(si.alpha, si.wo_local, nrm(si.wo_local + si.wi_local));
    return frostbite(si.wo_local, si.wi_local, si.color, si.alpha);
  }
  HairQuery<R> q{si.wo_local, si.wi_local, dot(si.to_local[1], si.n), si.color};
  return hair_eval(c.hp, c.hs, q, pdf);
}

// sample_disney_hair / sample_GGX with the draw order of
// optix_common.cuh:154-170 and :314-329. Returns f; *pdf is the sampling
// pdf. In frozen mode the direction and *pdf come from the sampling
// parameters, *pdf_eval is the pdf under the evaluation parameters and *f0
// the BSDF value under the sampling ones; otherwise all coincide.
template <class R>
Vec<R> bsdf_sample(const Ctx<R>& c, Inter<R>& si, Lcg& rng, R* pdf, R* pdf_eval = nullptr, Vec<R>* f0 = nullptr) {
  const float r0 = rng.next(), r1 = rng.next();
  if (!si.surface) {
    const float r2 = rng.next(), r3 = rng.next();
    int lobe;
    const R hh = dot(si.to_local[1], si.n);
    si.wi_local = hair_sample(c.frozen ? c.hs0 : c.hs, si.wo_local, hh, c.frozen ? si.color0 : si.color, (R)r0, (R)r1,
                              (R)r2, (R)r3, &lobe);
    si.wi = nrm(apply_T(si.to_local, si.wi_local));
    HairQuery<R> q{si.wo_local, si.wi_local, hh, si.color};
    R pe;
    const Vec<R> f = hair_eval(c.hp, c.hs, q, &pe);
    *pdf = pe;
    if (pdf_eval) *pdf_eval = pe;
    if (f0) *f0 = f;
    if (c.frozen) {
      HairQuery<R> q0{si.wo_local, si.wi_local, hh, si.color0};
      const Vec<R> fs = hair_eval(c.hp0, c.hs0, q0, pdf);
      if (f0) *f0 = fs;
    }
    return f;
  }
  si.wi_local = ggx_sample(si.alpha, si.wo_local, (R)r0, (R)r1, pdf);
  si.wi = nrm(apply_T(si.to_local, si.wi_local));
  if (pdf_eval) *pdf_eval = *pdf;
  const Vec<R> f = frostbite(si.wo_local, si.wi_local, si.color, si.alpha);
  if (f0) *f0 = f;
  return f;
}

template <class R>
bool occluded(const Ctx<R>& c, Vec<R> o, Vec<R> d) {
  const float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
  if (c.s.has_sphere) {
    float t;
    if (sphere_intersect(of, df, c.s.sphere, c.s.sphere[3], &t)) return true;
  }
  for (int k = 0; k < c.s.ntri; ++k) {
    float t, u, v;
    if (tri_intersect(of, df, c.s.tri_v + 9 * (size_t)k, &t, &u, &v)) return true;
  }
  return c.bvh.any(of, df, c.caps);
}

template <class R>
void trace(const Ctx<R>& c, Vec<R> o, Vec<R> d, Inter<R>& si) {
  const float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
  float best = std::numeric_limits<float>::infinity();
  int prim = -1;
  float u = 0.f;
  if (c.s.has_sphere) {
    float t;
    if (sphere_intersect(of, df, c.s.sphere, c.s.sphere[3], &t)) {
      best = t;
      prim = -2;
    }
  }
  // mesh triangles: minimum (t, triangle) by brute force
  for (int k = 0; k < c.s.ntri; ++k) {
    float tt, uu, vv;
    if (tri_intersect(of, df, c.s.tri_v + 9 * (size_t)k, &tt, &uu, &vv) && (tt < best || (tt == best && prim == -2))) {
      best = tt;
      prim = -3 - k;
    }
  }
  float th, uh;
  int id;
  if (c.bvh.closest(of, df, &th, &id, &uh, c.caps) && th <= best) {
    best = th;
    prim = id;
    u = uh;
  }
  if (prim == -1) {
    si.hit = false;
    si.Le = c.s.has_env ? env_radiance<R>(c.s, d) : mk((R)0, (R)0, (R)0);
    return;
  }
  setup_hit(c, o, d, best, prim, u, si);
}

// What the adjoint needs of one direct-lighting sample that reached a light:
// the direction, the radiance, both pdfs and whether the MIS weight applies.
template <class R>
struct DlRec {
  bool valid = false, mis = false;
  Vec<R> wi_local, emit;
  R light_pdf = 0, bsdf_pdf = 0;
};

// sampleBSDF (optix_common.cuh:151-203); si is the direct-lighting copy.
template <class R>
Vec<R> sample_bsdf_light(const Ctx<R>& c, Inter<R>& si, Lcg& rng, DlRec<R>* rec = nullptr) {
  R bsdfPdf = 1, pdfEval = 1;
  const Vec<R> bsdf = bsdf_sample(c, si, rng, &bsdfPdf, &pdfEval);
  const bool lower = dot(si.wi, si.n) < 0;
  Vec<R> nd = si.n;
  if (si.surface && lower) return mk((R)0, (R)0, (R)0);
  if (!si.surface && lower) {
    nd = (R)-1 * si.n;
    si.p = si.p + ((R)2 * si.radius) * nd;
  }
  if (occluded(c, si.p + (R)1e-3f * nd, si.wi)) return mk((R)0, (R)0, (R)0);
  const Vec<R> emit = env_radiance<R>(c.s, si.wi);
  const R lightPdf = env_pdf<R>(c.s, si.wi) * (R)1 / (R)(c.s.ndl + (c.s.has_env ? 1 : 0));
  if (rec) *rec = DlRec<R>{true, c.s.mis != 0, si.wi_local, emit, lightPdf, bsdfPdf};
  const R f2 = pdfEval * pdfEval, g2 = lightPdf * lightPdf;
  if (c.s.mis) return (f2 / (f2 + g2) / bsdfPdf) * (bsdf * emit);
  return (bsdf * emit) / bsdfPdf;
}

// sampleLights (optix_common.cuh:205-277)
template <class R>
Vec<R> sample_lights(const Ctx<R>& c, Inter<R>& si, Lcg& rng, bool& isDelta, DlRec<R>* rec = nullptr) {
  isDelta = false;
  const int ntot = c.s.ndl + (c.s.has_env ? 1 : 0);
  const float rsel = rng.next();
  int sel = (int)std::floor(rsel * (float)ntot);
  if (!c.s.has_env && sel >= c.s.ndl) sel = c.s.ndl - 1;
  R lightPdf = (R)1 * (R)1 / (R)ntot;
  Vec<R> emit;
  if (sel >= c.s.ndl) {
    R envPdf = 1;
    Vec<R> wi;
    emit = env_sample<R>(c.s, rng, &wi, &envPdf);
    si.wi = wi;
    si.wi_local = nrm(apply_mat(si.to_local, si.wi));
    lightPdf = lightPdf * envPdf;
  } else {
    emit = mk((R)c.s.dl_emit[sel][0], (R)c.s.dl_emit[sel][1], (R)c.s.dl_emit[sel][2]);
    si.wi = nrm(mk((R)c.s.dl_from[sel][0], (R)c.s.dl_from[sel][1], (R)c.s.dl_from[sel][2]));
    si.wi_local = nrm(apply_mat(si.to_local, si.wi));
    isDelta = true;
  }
  const bool lower = dot(si.wi, si.n) < 0;
  Vec<R> nd = si.n;
  if (si.surface && lower) return mk((R)0, (R)0, (R)0);
  if (!si.surface && lower) {
    nd = (R)-1 * si.n;
    si.p = si.p + ((R)2 * si.radius) * nd;
  }
  if (occluded(c, si.p + (R)1e-3f * nd, si.wi)) return mk((R)0, (R)0, (R)0);
  R bsdfPdf = 1;
  const Vec<R> bsdf = bsdf_eval(c, si, &bsdfPdf);
  if (rec) *rec = DlRec<R>{true, c.s.mis && !isDelta, si.wi_local, emit, lightPdf, bsdfPdf};
  if (c.s.mis && !isDelta) {
    const R f2 = lightPdf * lightPdf, g2 = bsdfPdf * bsdfPdf;
    return (f2 / (f2 + g2) / lightPdf) * (bsdf * emit);
  }
  return (bsdf * emit) / lightPdf;
}

// directLighting (optix_common.cuh:279-309): si by value.
template <class R>
Vec<R> direct_lighting(const Ctx<R>& c, Inter<R> si, Lcg& rng, DlRec<R>* r1 = nullptr, DlRec<R>* r2 = nullptr) {
  Vec<R> out;
  if (c.s.mis) {
    bool isDelta = false;
    const Vec<R> ls = sample_lights(c, si, rng, isDelta, r1);
    Vec<R> bs = mk((R)0, (R)0, (R)0);
    if (!isDelta) bs = sample_bsdf_light(c, si, rng, r2);
    out = ls + bs;
  } else {
    bool isDelta = false;
    out = sample_lights(c, si, rng, isDelta, r1);
  }
  if (has_nan(out)) {
    out = mk((R)0, (R)0, (R)0);
    if (r1) r1->valid = false;
    if (r2) r2->valid = false;
  }
  return out;
}

// The continuation's hair query, BSDF value and vertex, for the adjoint.
template <class R>
struct NvRec {
  bool hair = false, mul_bad = false;
  int prim = -1;
  HairQuery<R> q;
  Vec<R> f;
};

// nextPathVertex (optix_common.cuh:311-360); returns mulFac, and in *mul0
// the factor under the sampling parameters (frozen mode; else the same).
template <class R>
Vec<R> next_vertex(const Ctx<R>& c, Inter<R>& si, Lcg& rng, NvRec<R>* rec = nullptr, Vec<R>* mul0 = nullptr) {
  R pdf = 1;
  Vec<R> fs;
  const Vec<R> f = bsdf_sample(c, si, rng, &pdf, (R*)nullptr, &fs);
  Vec<R> mul = pdf == 0 ? f : f / pdf;
  if (mul0) {
    *mul0 = pdf == 0 ? fs : fs / pdf;
    if (has_nan(*mul0)) *mul0 = mk((R)1, (R)1, (R)1);
  }
  if (rec) {
    rec->hair = !si.surface;
    rec->mul_bad = has_nan(mul);
    rec->prim = si.prim;
    rec->q = HairQuery<R>{si.wo_local, si.wi_local, dot(si.to_local[1], si.n), si.color};
    rec->f = f;
  }
  if (has_nan(mul)) mul = mk((R)1, (R)1, (R)1);
  const Vec<R> wi = si.wi;
  Vec<R> nd = si.n;
  if (dot(wi, si.n) < 0 && !si.surface) {
    nd = (R)-1 * si.n;
    si.p = si.p + ((R)2 * si.radius) * nd;
  }
  const Vec<R> o = si.p + (R)1e-3f * nd;
  Inter<R> nx;
  nx.wi = wi;
  trace(c, o, wi, nx);
  si = nx;
  return mul;
}

template <class R> inline R lum(Vec<R> v) { return (R)0.212671f * v.x + (R)0.715160f * v.y + (R)0.072169f * v.z; }

// pathTrace (optix_common.cuh:362-409)
// In frozen mode the roulette follows the throughput beta0 under the
// sampling parameters; otherwise beta0 == beta.
template <class R>
Vec<R> path_trace(const Ctx<R>& c, Inter<R> si, Lcg& rng, int v1, int v2) {
  if (v2 < v1) return mk((R)0, (R)0, (R)0);
  Vec<R> beta = mk((R)1, (R)1, (R)1), beta0 = beta, col = mk((R)0, (R)0, (R)0);
  if (v1 == 0) col = direct_lighting(c, si, rng);
  for (int b = 1; b <= v2; ++b) {
    Vec<R> mul0;
    beta = beta * next_vertex(c, si, rng, (NvRec<R>*)nullptr, &mul0);
    beta0 = beta0 * mul0;
    if (!si.hit) break;
    if (b >= v1) col = col + beta * direct_lighting(c, si, rng);
    const R q = std::fmax((R)0.05f, (R)1 - lum(beta0));
    const float eps = rng.next();
    if ((R)eps < q) break;
    beta = beta / ((R)1 - q);
    beta0 = beta0 / ((R)1 - q);
  }
  return col;
}

// msnnPathTrace (hair_msnn.cu:134-185)
template <class R>
Vec<R> msnn_path(const Ctx<R>& c, Inter<R> si, Lcg& rng, int v2, int kbeta) {
  Vec<R> beta = mk((R)1, (R)1, (R)1);
  Vec<R> col = direct_lighting(c, si, rng);
  for (int b = 1; b <= v2; ++b) {
    beta = beta * next_vertex(c, si, rng);
    if (!si.hit) break;
    col = col + beta * direct_lighting(c, si, rng);
    const R q = std::fmax((R)0.05f, (R)1 - lum(beta));
    const float eps = rng.next();
    if ((R)eps < q || b > kbeta) break;
    beta = beta / ((R)1 - q);
  }
  return col;
}

// msnnTrainingPath (hair_msnn.cu:67-132)
template <class R>
Vec<R> msnn_training(const Ctx<R>& c, Inter<R> si, Lcg& rng, int v2, int kbeta, Vec<R>& shortc) {
  Vec<R> beta = mk((R)1, (R)1, (R)1), bs = beta;
  Vec<R> col = direct_lighting(c, si, rng);
  shortc = col;
  for (int b = 1; b <= v2; ++b) {
    const Vec<R> mul = next_vertex(c, si, rng);
    beta = beta * mul;
    bs = bs * mul;
    if (!si.hit) break;
    const Vec<R> dl = direct_lighting(c, si, rng);
    col = col + beta * dl;
    shortc = shortc + bs * dl;
    const R q = std::fmax((R)0.05f, (R)1 - lum(beta));
    const R qs = std::fmax((R)0.05f, (R)1 - lum(bs));
    const float eps = rng.next();
    if ((R)eps < qs || b > kbeta) bs = mk((R)0, (R)0, (R)0);
    if ((R)eps < q) break;
    beta = beta / ((R)1 - q);
    if (!(bs.x == 0 && bs.y == 0 && bs.z == 0)) bs = bs / ((R)1 - qs);
  }
  return col;
}

// nextPathVertexNRC (cuda/nrc.cu:16-67): nextPathVertex that also returns
// the sampled direction's pdf.
template <class R>
Vec<R> nrc_next_vertex(const Ctx<R>& c, Inter<R>& si, Lcg& rng, R* pdf_out) {
  R pdf = 1;
  const Vec<R> f = bsdf_sample(c, si, rng, &pdf);
  Vec<R> mul = pdf == 0 ? f : f / pdf;
  if (has_nan(mul)) mul = mk((R)1, (R)1, (R)1);
  const Vec<R> wi = si.wi;
  Vec<R> nd = si.n;
  if (dot(wi, si.n) < 0 && !si.surface) {
    nd = (R)-1 * si.n;
    si.p = si.p + ((R)2 * si.radius) * nd;
  }
  const Vec<R> o = si.p + (R)1e-3f * nd;
  Inter<R> nx;
  nx.wi = wi;
  trace(c, o, wi, nx);
  si = nx;
  *pdf_out = pdf;
  return mul;
}

template <class R>
inline R len2(Vec<R> v) {
  const R l = std::sqrt(dot(v, v));
  return l * l;
}

// nrcTracePaths (cuda/nrc.cu:135-343), one pixel sample from its primary hit.
template <class R>
void nrc_trace(const Ctx<R>& c, Inter<R> si, Lcg& rng, Vec<R> cam, bool training, bool unbiased, R cthr,
               R scale, NrcPixelOut& o) {
  bool suffix = false, recalc = false;
  R a0 = len2(si.p - cam) / ((R)4 * Pi<R>()) / std::fabs(si.wo_local.z);
  Vec<R> prev = mk((R)0, (R)0, (R)0);
  R prev_pdf = 0, spread = 0;
  Vec<R> col = mk((R)0, (R)0, (R)0), beta = mk((R)1, (R)1, (R)1);
  auto put = [](double* d, Vec<R> v) { d[0] = (double)v.x; d[1] = (double)v.y; d[2] = (double)v.z; };
  auto putf = [](float* d, Vec<R> v) { d[0] = (float)v.x; d[1] = (float)v.y; d[2] = (float)v.z; };
  for (int b = 0; b < kNrcMaxBounces; ++b) {
    const Vec<R> dl = direct_lighting(c, si, rng);
    col = col + beta * dl;
    const Vec<R> pwo = si.wo, pn = si.n;
    prev = si.p;
    const Vec<R> mul = nrc_next_vertex(c, si, rng, &prev_pdf);
    beta = beta * mul;
    if (training) {
      if (recalc) {
        a0 = len2(si.p - prev) / ((R)4 * Pi<R>()) / std::fabs(si.wo_local.z);
        recalc = false;
      }
      double* r = o.rec + (size_t)b * 15;
      put(r, prev / scale);
      put(r + 3, pwo);
      put(r + 6, pn);
      put(r + 9, dl);
      put(r + 12, mul);
    }
    spread = spread + std::sqrt(len2(si.p - prev) / prev_pdf / std::fabs(si.wo_local.z));
    const bool cond = spread * spread > cthr * a0;
    if (!si.hit || b == kNrcMaxBounces - 1) {
      if (!suffix) {
        put(o.rad, col);
        put(o.beta, mk((R)0, (R)0, (R)0));
      }
      o.bounces = b;
      o.tb_hit = 0;
      break;
    } else if (cond && !suffix) {
      put(o.rad, col);
      put(o.beta, beta);
      o.query = 1;
      putf(o.q, si.p / scale);
      putf(o.q + 3, si.wo);
      putf(o.q + 6, si.n);
      if (training) {
        suffix = true;
        recalc = true;
        spread = 0;
      } else {
        break;
      }
    } else if (cond && suffix) {
      putf(o.sq, si.p / scale);
      putf(o.sq + 3, si.wo);
      putf(o.sq + 6, si.n);
      o.bounces = b;
      o.tb_hit = 1;
      if (!unbiased) break;
      cthr = (R)1e30f;
    }
  }
}

template <class R>
void nrc_gbuffer(const SceneDesc& s, const Bvh& bvh, int accum_id, const uint8_t* flags, float cthr, float scale,
                 int threads, NrcPixelOut* out) {
  Ctx<R> c{s, bvh, {}, {}};
  c.caps = 0;  // NRC: no end caps (render_nrc.cu:409)
  c.hp = HairParams<R>{(R)s.beta_m, (R)s.beta_n, (R)s.alpha, (R)s.gains[0], (R)s.gains[1], (R)s.gains[2], (R)s.gains[3]};
  c.hs = setup_hair(c.hp);
  c.hp0 = c.hp;
  c.hs0 = c.hs;
  const Vec<R> cp = mk((R)s.cam[0], (R)s.cam[1], (R)s.cam[2]);
  const Vec<R> d00 = mk((R)s.cam[3], (R)s.cam[4], (R)s.cam[5]);
  const Vec<R> du = mk((R)s.cam[6], (R)s.cam[7], (R)s.cam[8]);
  const Vec<R> dv = mk((R)s.cam[9], (R)s.cam[10], (R)s.cam[11]);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int py = 0; py < s.height; ++py) {
    for (int px = 0; px < s.width; ++px) {
      const size_t i = (size_t)py * s.width + px;
      NrcPixelOut& o = out[i];
      Lcg rng{get_rng((uint32_t)(accum_id + 10007), (uint32_t)px, (uint32_t)py, (uint32_t)s.width)};
      const float jx = rng.next(), jy = rng.next();
      const float sx = ((float)px + jx) / (float)s.width, sy = ((float)py + jy) / (float)s.height;
      const Vec<R> d = nrm(d00 + (R)sx * du + (R)sy * dv);
      Inter<R> si;
      si.wi = d;
      trace(c, cp, d, si);
      if (!si.hit || (si.surface && si.wo_local.z < 0)) {
        const Vec<R> le = si.hit ? mk((R)0, (R)0, (R)0) : si.Le;
        o.rad[0] = (double)le.x;
        o.rad[1] = (double)le.y;
        o.rad[2] = (double)le.z;
        continue;
      }
      o.gb_hit = 1;
      nrc_trace(c, si, rng, cp, (flags[i] & 1) != 0, (flags[i] & 2) != 0, (R)cthr, (R)scale, o);
    }
  }
}

template void nrc_gbuffer<float>(const SceneDesc&, const Bvh&, int, const uint8_t*, float, float, int, NrcPixelOut*);
template void nrc_gbuffer<double>(const SceneDesc&, const Bvh&, int, const uint8_t*, float, float, int, NrcPixelOut*);

template <class R>
void render_pixels(const SceneDesc& s, const Bvh& bvh, int mode, int kbeta, int accum_id0, int spp, int x0, int y0,
                   int w, int h, int threads, double* out_full, double* out_short, uint8_t* out_hit) {
  Ctx<R> c{s, bvh, {}, {}};
  c.caps = mode == 0 ? 1 : 0;  // HairMSNN modes have no end caps (render_hair_msnn.cu:346)
  c.hp = HairParams<R>{(R)s.beta_m, (R)s.beta_n, (R)s.alpha, (R)s.gains[0], (R)s.gains[1], (R)s.gains[2], (R)s.gains[3]};
  c.hs = setup_hair(c.hp);
  c.frozen = s.frozen != 0 && mode == 0;
  c.hp0 = c.hp;
  if (c.frozen) {
    c.hp0.beta_m = (R)s.beta_m0;
    c.hp0.beta_n = (R)s.beta_n0;
    c.hp0.alpha = (R)s.alpha0;
  }
  c.hs0 = setup_hair(c.hp0);
  const Vec<R> cp = mk((R)s.cam[0], (R)s.cam[1], (R)s.cam[2]);
  const Vec<R> d00 = mk((R)s.cam[3], (R)s.cam[4], (R)s.cam[5]);
  const Vec<R> du = mk((R)s.cam[6], (R)s.cam[7], (R)s.cam[8]);
  const Vec<R> dv = mk((R)s.cam[9], (R)s.cam[10], (R)s.cam[11]);
  const int v1 = s.path_v1 - 1, v2 = s.path_v2 - 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int yy = 0; yy < h; ++yy) {
    for (int xx = 0; xx < w; ++xx) {
      const int px = x0 + xx, py = y0 + yy;
      double acc[3] = {0, 0, 0}, accs[3] = {0, 0, 0};
      uint8_t hitflag = 0;
      for (int k = 0; k < spp; ++k) {
        Lcg rng{get_rng((uint32_t)(accum_id0 + k + 10007), (uint32_t)px, (uint32_t)py, (uint32_t)s.width)};
        const float jx = rng.next(), jy = rng.next();
        const float sx = ((float)px + jx) / (float)s.width, sy = ((float)py + jy) / (float)s.height;
        const Vec<R> d = nrm(d00 + (R)sx * du + (R)sy * dv);
        Inter<R> si;
        si.wi = d;
        trace(c, cp, d, si);
        Vec<R> col = mk((R)0, (R)0, (R)0), sc = col;
        if (!si.hit) {
          col = si.Le;
          sc = col;
        } else {
          hitflag = (uint8_t)(1 | (si.surface ? 2 : 0));
          if (mode == 0) col = path_trace(c, si, rng, v1, v2);
          else if (mode == 1) col = msnn_path(c, si, rng, v2, kbeta);
          else col = msnn_training(c, si, rng, v2, kbeta, sc);
        }
        if (has_nan(col)) col = mk((R)0, (R)0, (R)0);  // writePixel scrub (utils.cuh:20-21)
        acc[0] += (double)col.x;
        acc[1] += (double)col.y;
        acc[2] += (double)col.z;
        accs[0] += (double)sc.x;
        accs[1] += (double)sc.y;
        accs[2] += (double)sc.z;
      }
      const size_t o = (size_t)yy * w + xx;
      for (int k = 0; k < 3; ++k) {
        out_full[3 * o + k] = acc[k] / spp;
        if (out_short) out_short[3 * o + k] = accs[k] / spp;
      }
      if (out_hit) out_hit[o] = hitflag;
    }
  }
}

template <class R>
void msnn_records(const SceneDesc& s, const Bvh& bvh, int accum_id, int kbeta, int n, const int* pixels,
                  float scene_scale, int threads, MsnnRecordOut* out) {
  Ctx<R> c{s, bvh, {}, {}};
  c.caps = 0;  // HairMSNN: no end caps (render_hair_msnn.cu:346)
  c.hp = HairParams<R>{(R)s.beta_m, (R)s.beta_n, (R)s.alpha, (R)s.gains[0], (R)s.gains[1], (R)s.gains[2], (R)s.gains[3]};
  c.hs = setup_hair(c.hp);
  c.hp0 = c.hp;
  c.hs0 = c.hs;
  const Vec<R> cp = mk((R)s.cam[0], (R)s.cam[1], (R)s.cam[2]);
  const Vec<R> d00 = mk((R)s.cam[3], (R)s.cam[4], (R)s.cam[5]);
  const Vec<R> du = mk((R)s.cam[6], (R)s.cam[7], (R)s.cam[8]);
  const Vec<R> dv = mk((R)s.cam[9], (R)s.cam[10], (R)s.cam[11]);
  const int v2 = s.path_v2 - 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int i = 0; i < n; ++i) {
    const int px = pixels[i] % s.width, py = pixels[i] / s.width;
    // camera ray of the G_BUFFER pass (hair_msnn.cu:193-220)
    Lcg rng{get_rng((uint32_t)(accum_id + 10007), (uint32_t)px, (uint32_t)py, (uint32_t)s.width)};
    const float jx = rng.next(), jy = rng.next();
    const float sx = ((float)px + jx) / (float)s.width, sy = ((float)py + jy) / (float)s.height;
    const Vec<R> d = nrm(d00 + (R)sx * du + (R)sy * dv);
    Inter<R> si;
    si.wi = d;
    trace(c, cp, d, si);
    MsnnRecordOut& o = out[i];
    o = MsnnRecordOut{};
    Vec<R> col = mk((R)0, (R)0, (R)0), sc = col;
    if (si.hit) {
      col = msnn_training(c, si, rng, v2, kbeta, sc);
      o.hit = (uint8_t)(1 | (si.surface ? 2 : 0));
      const Vec<R> p = si.p / (R)scene_scale;
      const Vec<R> t = si.to_local[0];
      const float rec[12] = {(float)p.x, (float)p.y, (float)p.z, (float)si.wo.x, (float)si.wo.y, (float)si.wo.z,
                             (float)t.x, (float)t.y, (float)t.z, 0.f, 0.f, 0.f};
      std::copy(rec, rec + 12, o.input);
      // channels 9-11: the hit's absorption (the reference leaves them unset)
      const Vec<R> sig = si.surface ? mk((R)s.sigma_a[0], (R)s.sigma_a[1], (R)s.sigma_a[2]) : si.color;
      o.input[9] = (float)sig.x;
      o.input[10] = (float)sig.y;
      o.input[11] = (float)sig.z;
    } else {
      o.input[3] = (float)-d.x;
      o.input[4] = (float)-d.y;
      o.input[5] = (float)-d.z;
    }
    auto has_inf = [](Vec<R> v) { return std::isinf(v.x) || std::isinf(v.y) || std::isinf(v.z); };
    if (has_nan(col)) col = mk((R)0, (R)0, (R)0);
    if (has_inf(col)) col = mk((R)1e5f, (R)1e5f, (R)1e5f);
    if (has_nan(sc)) sc = mk((R)0.01f, (R)0.01f, (R)0.01f);
    if (has_inf(sc)) sc = mk((R)1e5f, (R)1e5f, (R)1e5f);
    const R tg[3] = {col.x - sc.x, col.y - sc.y, col.z - sc.z};
    const R fc[3] = {col.x, col.y, col.z}, scc[3] = {sc.x, sc.y, sc.z};
    for (int k = 0; k < 3; ++k) {
      o.target[k] = (double)tg[k];
      o.full[k] = (double)fc[k];
      o.shortc[k] = (double)scc[k];
    }
  }
}
template void msnn_records<float>(const SceneDesc&, const Bvh&, int, int, int, const int*, float, int, MsnnRecordOut*);
template void msnn_records<double>(const SceneDesc&, const Bvh&, int, int, int, const int*, float, int, MsnnRecordOut*);

// ---------------------------------------------------------------------------
// Path-replay adjoint in FP64 with the BSDF derivatives taken by central
// differences of the FP64 hair BSDF (an independent route to the GPU's dual
// numbers). Each (pixel, pass) path of pathTrace is run twice from its seed:
// first for its radiance, then again accumulating, at every hair vertex, the
// parameter derivative of
//   the light sample      f * Le * w(light_pdf, bsdf_pdf(theta)) / light_pdf,
//   the BSDF env sample   f * Le * w(bsdf_pdf(theta), light_pdf) / bsdf_pdf0,
//   the continuation      rem * f(theta) / f0 (rem: radiance still to come),
// with the sampled directions, the sampling pdfs and Russian roulette fixed.
namespace {

struct AdjFd {
  HairParams<double> hp[7];  // base, then -/+ for beta_m, beta_n, alpha
  HairShading<double> hs[7];
  double h;
};

inline double pw(double a, double b) { return a * a / (a * a + b * b); }

// g[k] += d/d theta_k of sum_c W_c * C_c(f(theta), pdf(theta)) by central
// differences; parameters 3..5 are the query's sigma_a channels.
template <class F>
void fd_grad(const AdjFd& a, const HairQuery<double>& q, Vec<double> W, F contrib, double g[6]) {
  auto S = [&](int j, const HairQuery<double>& qq) {
    double pdf;
    const Vec<double> f = hair_eval(a.hp[j], a.hs[j], qq, &pdf);
    const Vec<double> C = contrib(f, pdf);
    return W.x * C.x + W.y * C.y + W.z * C.z;
  };
  for (int k = 0; k < 3; ++k) g[k] += (S(2 + 2 * k, q) - S(1 + 2 * k, q)) / (2 * a.h);
  for (int ch = 0; ch < 3; ++ch) {
    HairQuery<double> qp = q, qm = q;
    double* cp = ch == 0 ? &qp.sigma_a.x : ch == 1 ? &qp.sigma_a.y : &qp.sigma_a.z;
    double* cm = ch == 0 ? &qm.sigma_a.x : ch == 1 ? &qm.sigma_a.y : &qm.sigma_a.z;
    *cp += a.h;
    *cm -= a.h;
    g[3 + ch] += (S(0, qp) - S(0, qm)) / (2 * a.h);
  }
}

struct AdjAcc {
  double glob[3] = {0, 0, 0};
  std::vector<double> sigma;
};

void add_vertex_grad(const SceneDesc& s, int prim, const double g[6], AdjAcc& acc, double* px) {
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(g[k])) return;
  for (int k = 0; k < 3; ++k) {
    acc.glob[k] += g[k];
    px[k] += g[k];
  }
  const int strand = (int)s.seg[8 * (size_t)prim + 7];
  for (int ch = 0; ch < 3; ++ch) acc.sigma[3 * (size_t)strand + ch] += g[3 + ch];
}

// pathTrace (optix_common.cuh:362-409) with the adjoint records.
Vec<double> path_adjoint(const Ctx<double>& c, const AdjFd& a, Inter<double> si, Lcg& rng, int v1, int v2,
                         bool replay, Vec<double> Ltot, Vec<double> adj, AdjAcc& acc, double* px) {
  const Vec<double> zero = mk(0.0, 0.0, 0.0);
  if (v2 < v1) return zero;
  Vec<double> beta = mk(1.0, 1.0, 1.0), col = zero;
  auto dl_at = [&](const Inter<double>& v) {
    DlRec<double> r1, r2;
    const Vec<double> out = direct_lighting(c, v, rng, &r1, &r2);
    if (replay && !v.surface) {
      const Vec<double> W = adj * beta;
      const double hh = dot(v.to_local[1], v.n);
      double g[6] = {0, 0, 0, 0, 0, 0};
      if (r1.valid) {
        const HairQuery<double> q{v.wo_local, r1.wi_local, hh, v.color};
        fd_grad(a, q, W, [&](Vec<double> f, double pdf) {
          const double w = r1.mis ? pw(r1.light_pdf, pdf) : 1.0;
          return (w / r1.light_pdf) * (f * r1.emit);
        }, g);
      }
      if (r2.valid) {
        const HairQuery<double> q{v.wo_local, r2.wi_local, hh, v.color};
        const double pb0 = r2.bsdf_pdf;
        fd_grad(a, q, W, [&](Vec<double> f, double pdf) { return (pw(pdf, r2.light_pdf) / pb0) * (f * r2.emit); }, g);
      }
      if (r1.valid || r2.valid) add_vertex_grad(c.s, v.prim, g, acc, px);
    }
    return out;
  };
  if (v1 == 0) col = dl_at(si);
  for (int b = 1; b <= v2; ++b) {
    NvRec<double> nv;
    const Vec<double> mul = next_vertex(c, si, rng, &nv);
    if (replay && nv.hair && !nv.mul_bad) {
      const Vec<double> rem = Ltot - col;
      const Vec<double> f0 = nv.f;
      double g[6] = {0, 0, 0, 0, 0, 0};
      fd_grad(a, nv.q, adj, [&](Vec<double> f, double) {
        return mk(f0.x != 0 ? rem.x * f.x / f0.x : 0.0, f0.y != 0 ? rem.y * f.y / f0.y : 0.0,
                  f0.z != 0 ? rem.z * f.z / f0.z : 0.0);
      }, g);
      add_vertex_grad(c.s, nv.prim, g, acc, px);
    }
    beta = beta * mul;
    if (!si.hit) break;
    if (b >= v1) col = col + beta * dl_at(si);
    const double q = std::fmax((double)0.05f, 1.0 - lum(beta));
    const float eps = rng.next();
    if ((double)eps < q) break;
    beta = beta / (1.0 - q);
  }
  return col;
}

}  // namespace

void adjoint_render(const SceneDesc& s, const Bvh& bvh, int spp, const float* dl_dimg, int num_strands, int threads,
                    double fd_h, double* grad_sigma, double* grad_glob, double* grad_px) {
  Ctx<double> c{s, bvh, {}, {}};
  c.hp = HairParams<double>{(double)s.beta_m, (double)s.beta_n, (double)s.alpha, (double)s.gains[0],
                            (double)s.gains[1], (double)s.gains[2], (double)s.gains[3]};
  c.hs = setup_hair(c.hp);
  AdjFd a;
  a.h = fd_h;
  for (int j = 0; j < 7; ++j) {
    a.hp[j] = c.hp;
    if (j > 0) {
      double* t = j <= 2 ? &a.hp[j].beta_m : j <= 4 ? &a.hp[j].beta_n : &a.hp[j].alpha;
      *t += (j % 2 == 0) ? fd_h : -fd_h;
    }
    a.hs[j] = setup_hair(a.hp[j]);
  }
  const Vec<double> cp = mk((double)s.cam[0], (double)s.cam[1], (double)s.cam[2]);
  const Vec<double> d00 = mk((double)s.cam[3], (double)s.cam[4], (double)s.cam[5]);
  const Vec<double> du = mk((double)s.cam[6], (double)s.cam[7], (double)s.cam[8]);
  const Vec<double> dv = mk((double)s.cam[9], (double)s.cam[10], (double)s.cam[11]);
  const int v1 = s.path_v1 - 1, v2 = s.path_v2 - 1;
  const int W = s.width, H = s.height;
  for (int k = 0; k < 3; ++k) grad_glob[k] = 0;
  std::fill(grad_sigma, grad_sigma + 3 * (size_t)num_strands, 0.0);
  if (grad_px) std::fill(grad_px, grad_px + 3 * (size_t)W * H, 0.0);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
  {
    AdjAcc acc;
    acc.sigma.assign(3 * (size_t)num_strands, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int py = 0; py < H; ++py) {
      for (int px = 0; px < W; ++px) {
        const float* g4 = dl_dimg + 4 * ((size_t)py * W + px);
        const Vec<double> adj = (1.0 / spp) * mk((double)g4[0], (double)g4[1], (double)g4[2]);
        if (adj.x == 0 && adj.y == 0 && adj.z == 0) continue;
        double pxg[3] = {0, 0, 0};
        for (int k = 0; k < spp; ++k) {
          Vec<double> Lt = mk(0.0, 0.0, 0.0);
          for (int pass = 0; pass < 2; ++pass) {
            Lcg rng{get_rng((uint32_t)(k + 10007), (uint32_t)px, (uint32_t)py, (uint32_t)W)};
            const float jx = rng.next(), jy = rng.next();
            const float sx = ((float)px + jx) / (float)W, sy = ((float)py + jy) / (float)H;
            const Vec<double> d = nrm(d00 + (double)sx * du + (double)sy * dv);
            Inter<double> si;
            si.wi = d;
            trace(c, cp, d, si);
            if (!si.hit) break;
            if (pass == 0) {
              Lt = path_adjoint(c, a, si, rng, v1, v2, false, Lt, adj, acc, pxg);
              if (has_nan(Lt)) break;
            } else {
              path_adjoint(c, a, si, rng, v1, v2, true, Lt, adj, acc, pxg);
            }
          }
        }
        if (grad_px)
          for (int k = 0; k < 3; ++k) grad_px[3 * ((size_t)py * W + px) + k] = pxg[k];
      }
    }
#pragma omp critical
    {
      for (int k = 0; k < 3; ++k) grad_glob[k] += acc.glob[k];
      for (size_t i = 0; i < acc.sigma.size(); ++i) grad_sigma[i] += acc.sigma[i];
    }
  }
}

template void render_pixels<float>(const SceneDesc&, const Bvh&, int, int, int, int, int, int, int, int, int, double*,
                                   double*, uint8_t*);
template void render_pixels<double>(const SceneDesc&, const Bvh&, int, int, int, int, int, int, int, int, int, double*,
                                    double*, uint8_t*);

// ---- probes of single integrator functions --------------------------------

template <class R>
static Ctx<R> probe_ctx(const SceneDesc& s, const Bvh& bvh) {
  Ctx<R> c{s, bvh, {}, {}};
  c.caps = 1;
  c.hp = HairParams<R>{(R)s.beta_m, (R)s.beta_n, (R)s.alpha, (R)s.gains[0], (R)s.gains[1], (R)s.gains[2], (R)s.gains[3]};
  c.hs = setup_hair(c.hp);
  c.hp0 = c.hp;
  c.hs0 = c.hs;
  return c;
}

template <class R>
static void put3(double* o, Vec<R> v) {
  o[0] = (double)v.x;
  o[1] = (double)v.y;
  o[2] = (double)v.z;
}

template <class R>
void probe_env(const SceneDesc& s, int n, const float* dirs, double* rgb, double* pdf) {
  for (int i = 0; i < n; ++i) {
    const float* d = dirs + 3 * (size_t)i;
    const Vec<R> w = mk((R)d[0], (R)d[1], (R)d[2]);
    put3(rgb + 3 * (size_t)i, env_radiance<R>(s, w));
    pdf[i] = (double)env_pdf<R>(s, w);
  }
}

template <class R>
void probe_env_sample(const SceneDesc& s, int n, const uint32_t* seeds, double* wi, double* rgb, double* pdf) {
  for (int i = 0; i < n; ++i) {
    Lcg rng{seeds[i]};
    Vec<R> w;
    R p = 0;
    const Vec<R> L = env_sample<R>(s, rng, &w, &p);
    put3(wi + 3 * (size_t)i, w);
    put3(rgb + 3 * (size_t)i, L);
    pdf[i] = (double)p;
  }
}

template <class R>
void probe_hit(const SceneDesc& s, const Bvh& bvh, int n, const float* rays, const float* t, const int* prim,
               const float* u, double* out) {
  const Ctx<R> c = probe_ctx<R>(s, bvh);
  for (int i = 0; i < n; ++i) {
    const float* r = rays + 8 * (size_t)i;
    Inter<R> si;
    setup_hit(c, mk((R)r[0], (R)r[1], (R)r[2]), mk((R)r[4], (R)r[5], (R)r[6]), t[i], prim[i], u[i], si);
    double* o = out + kProbeHit * (size_t)i;
    put3(o, si.p);
    put3(o + 3, si.n);
    put3(o + 6, si.to_local[0]);
    put3(o + 9, si.to_local[1]);
    put3(o + 12, si.to_local[2]);
    put3(o + 15, si.wo_local);
    o[18] = (double)dot(si.to_local[1], si.n);
    o[19] = (double)si.radius;
  }
}

template <class R>
void probe_direct(const SceneDesc& s, const Bvh& bvh, const float ray[8], float t, int prim, float u, int n,
                  const uint32_t* seeds, int threads, double* mean_rgb) {
  const Ctx<R> c = probe_ctx<R>(s, bvh);
  Inter<R> si;
  si.wi = mk((R)ray[4], (R)ray[5], (R)ray[6]);
  setup_hit(c, mk((R)ray[0], (R)ray[1], (R)ray[2]), si.wi, t, prim, u, si);
  double acc[6] = {0, 0, 0, 0, 0, 0};
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static) reduction(+ : acc[:6])
#endif
  for (int i = 0; i < n; ++i) {
    Lcg rng{seeds[i]};
    const Vec<R> L = direct_lighting(c, si, rng);
    const double l[3] = {(double)L.x, (double)L.y, (double)L.z};
    for (int k = 0; k < 3; ++k) {
      acc[k] += l[k];
      acc[3 + k] += l[k] * l[k];
    }
  }
  for (int k = 0; k < 6; ++k) mean_rgb[k] = acc[k] / (n > 0 ? n : 1);
}

template void probe_env<double>(const SceneDesc&, int, const float*, double*, double*);
template void probe_env<float>(const SceneDesc&, int, const float*, double*, double*);
template void probe_env_sample<double>(const SceneDesc&, int, const uint32_t*, double*, double*, double*);
template void probe_env_sample<float>(const SceneDesc&, int, const uint32_t*, double*, double*, double*);
template void probe_hit<double>(const SceneDesc&, const Bvh&, int, const float*, const float*, const int*,
                                const float*, double*);
template void probe_hit<float>(const SceneDesc&, const Bvh&, int, const float*, const float*, const int*,
                               const float*, double*);
template void probe_direct<double>(const SceneDesc&, const Bvh&, const float*, float, int, float, int,
                                   const uint32_t*, int, double*);
template void probe_direct<float>(const SceneDesc&, const Bvh&, const float*, float, int, float, int,
                                  const uint32_t*, int, double*);

}  // namespace orc
