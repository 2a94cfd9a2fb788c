// Per-path random stream: Murmur3-mixed seed + 32-bit LCG.
//
// Integer arithmetic identical to the reference's generator
// (cuda_headers/lcg_random.cuh:15-66) so every path consumes the same
// uniforms, in the same order, as the reference's raygen programs.
#pragma once

#include "hm_common.h"

namespace hm {

HM_HD uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

// One Murmur3 block-mix step (lcg_random.cuh:15-32).
HM_HD uint32_t murmur_mix(uint32_t h, uint32_t k) {
  k *= 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  h ^= k;
  return rotl32(h, 13) * 5u + 0xe6546b64u;
}

// Murmur3 finaliser (lcg_random.cuh:34-43).
HM_HD uint32_t murmur_fmix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// Seed of pixel (x, y) for frame `frame` (lcg_random.cuh:58-66). The raygen
// programs pass frame = accumId + 10007 (cuda/path_tracing.cu:29).
HM_HD uint32_t path_seed(uint32_t frame, uint32_t x, uint32_t y, uint32_t width) {
  return murmur_fmix(murmur_mix(murmur_mix(0u, x + y * width), frame));
}

// Advance and return the raw 32-bit state (lcg_random.cuh:45-51).
HM_HD uint32_t lcg_next(uint32_t& s) {
  s = s * 1664525u + 1013904223u;
  return s;
}

// Uniform float = float(u32) * 2^-32 (lcg_random.cuh:53-56). The u32->f32
// conversion rounds to nearest, so values >= 2^32-128 return exactly 1.0f.
HM_HD float lcg_uniform(uint32_t& s) { return (float)lcg_next(s) * 2.3283064365386963e-10f; }

}  // namespace hm
