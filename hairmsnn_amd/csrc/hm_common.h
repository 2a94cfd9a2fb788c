// Shared host/device vector math for the HairMSNN MI355X path tracer.
//
// Everything here compiles under hipcc (device + host) and under plain g++
// (host-only tools: scene parser, procedural generators, CLIs).
#pragma once

#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HM_HD __host__ __device__ __forceinline__
#define HM_D __device__ __forceinline__
// Large routines called from several places of one kernel: one out-of-line
// copy keeps the shading kernels' machine code small enough for the
// instruction cache.
#define HM_HD_CALL __host__ __device__ __attribute__((noinline))
#define HM_D_CALL __device__ __attribute__((noinline))
#else
#define HM_HD inline
#define HM_D inline
#define HM_HD_CALL inline
#define HM_D_CALL inline
#endif

namespace hm {

// The reference defines Pi as 3.1415926f and TWO_Pi as the unparenthesised
// `2.f * 3.14159f` (cuda_headers/utils.cuh:10-11); both constants feed the
// environment-map and BSDF arithmetic, so the same values are used here.
constexpr float kPi = 3.1415926f;
constexpr float kTwoPiEnv = 2.f * 3.14159f;  // env wrap constant (utils.cuh:11)
constexpr float kHostPi = 3.14159f;          // host-side constant (scene.cpp:207, :363)

struct V3 {
  float x, y, z;
};

HM_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
HM_HD V3 v3(float s) { return V3{s, s, s}; }
HM_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
HM_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
HM_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
HM_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
HM_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
HM_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
HM_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
HM_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
HM_HD V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
HM_HD V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
HM_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HM_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
HM_HD float length(V3 a) { return sqrtf(dot(a, a)); }
HM_HD V3 normalize(V3 a) {
  float inv = 1.f / sqrtf(dot(a, a));
  return a * inv;
}
HM_HD V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
HM_HD V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
HM_HD bool any_nan(V3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z); }
HM_HD bool any_inf(V3 a) { return isinf(a.x) || isinf(a.y) || isinf(a.z); }
HM_HD bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
HM_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
HM_HD float sqr(float x) { return x * x; }
HM_HD float safe_sqrt(float x) { return sqrtf(fmaxf(0.f, x)); }

// sin and cos for the shading code's angles (|x| below a few hundred; the
// largest are azimuths of at most ~8 pi). Cody and Waite's reduction: k =
// nearest integer to x / (pi/2), r = x - k pi/2 with pi/2 split into three
// FP32 parts (12 + 12 + 24 bits, so k * part is exact for |k| < 2^11), then
// the Taylor polynomials of sin (to r^9) and cos (to r^10) on
// [-pi/4, pi/4], whose truncation errors are below 2e-9; the quadrant k mod 4
// picks and signs the two. This avoids the large-argument reduction that the
// library sinf / cosf carry at every call site (most of their code size).
constexpr float kPio2Hi = 1.5703125f;
constexpr float kPio2Mid = 4.837512970e-04f;
constexpr float kPio2Lo = 7.549790126e-08f;
constexpr float kTwoOverPi = 6.366197467e-01f;
HM_HD void sincos_small(float x, float* s_out, float* c_out) {
  const float k = rintf(x * kTwoOverPi);
  float r = fmaf(-k, kPio2Hi, x);
  r = fmaf(-k, kPio2Mid, r);
  r = fmaf(-k, kPio2Lo, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.755731884e-06f, -1.984127011e-04f);
  ps = fmaf(r2, ps, 8.333333768e-03f);
  ps = fmaf(r2, ps, -1.666666716e-01f);
  const float sn = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.755731998e-07f, 2.480158764e-05f);
  pc = fmaf(r2, pc, -1.388888923e-03f);
  pc = fmaf(r2, pc, 4.166666791e-02f);
  pc = fmaf(r2, pc, -0.5f);
  const float cs = fmaf(r2, pc, 1.f);
  const int q = (int)k & 3;
  const float s = (q & 1) ? cs : sn;
  const float c = (q & 1) ? sn : cs;
  *s_out = (q & 2) ? -s : s;
  *c_out = ((q + 1) & 2) ? -c : c;
}
HM_HD float sin_small(float x) {
  float s, c;
  sincos_small(x, &s, &c);
  return s;
}
HM_HD float cos_small(float x) {
  float s, c;
  sincos_small(x, &s, &c);
  return c;
}

// Rec. 709 luminance (utils.cuh:79-83).
HM_HD float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// Power heuristic with one sample per strategy (utils.cuh:290-294).
HM_HD float power_heuristic(float f, float g) {
  float ff = f * f, gg = g * g;
  return ff / (ff + gg);
}

// Row-vector frame: local = (dot(r0,v), dot(r1,v), dot(r2,v)).
struct Frame {
  V3 r0, r1, r2;
};
HM_HD V3 to_local(const Frame& f, V3 v) { return V3{dot(f.r0, v), dot(f.r1, v), dot(f.r2, v)}; }
// The transpose applied to a local vector (the frames are orthonormal).
HM_HD V3 to_world(const Frame& f, V3 l) { return f.r0 * l.x + f.r1 * l.y + f.r2 * l.z; }

// Tangent frame around a surface normal (utils.cuh:232-259): rows are
// (c1, c2, n). The reference promotes `1./(1.+n.z)` to double; on the GPU
// everything here stays FP32 (the frame is orthonormal either way).
HM_HD Frame basis_from_normal(V3 n) {
  V3 c1, c2;
  if (n.z < -0.999999f) {
    c1 = v3(0.f, -1.f, 0.f);
    c2 = v3(-1.f, 0.f, 0.f);
  } else {
    float a = 1.f / (1.f + n.z);
    float b = -n.x * n.y * a;
    c1 = normalize(v3(1.f - n.x * n.x * a, b, -n.x));
    c2 = normalize(v3(b, 1.f - n.y * n.y * a, -n.y));
  }
  return Frame{c1, c2, n};
}

// BVH leaf child code: ~(first | (count - 1) << kLeafCountShift), a run of
// count consecutive primitive positions from first (a single primitive:
// ~first). Trees hold at most 2^24 primitives.
constexpr int kLeafCountShift = 24;
constexpr int kLeafPosMask = (1 << kLeafCountShift) - 1;
constexpr int kMaxLeafRun = 8;
HM_HD int leaf_code(int first, int count) { return ~(first | ((count - 1) << kLeafCountShift)); }
// The code of the rest of a run after its first position, or 0 (none left).
HM_HD int leaf_rest(int code) {
  const int more = code >> kLeafCountShift;
  return more ? ~(((code & kLeafPosMask) + 1) | ((more - 1) << kLeafCountShift)) : 0;
}

}  // namespace hm
