// Device-side helpers and kernel parameter records shared by the HIP kernels.
//
// The library is compiled with -ffp-contract=off: expressions whose rounding
// decides an integer result (unprojection -> cascade and cell, voxel
// coverage) round exactly as written, in the statement order of the GLSL
// they restate.  Throughput-bound loops opt into fused multiply-adds
// explicitly with __builtin_fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/drv.h"

namespace drvk {

constexpr float kPi = 3.14159265358979f;  // shader/utils.glsl:1

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 make3(float x, float y, float z) { return f3{x, y, z}; }

// Saturating float -> int with NaN -> 0 (GLSL int() is undefined outside the
// int range; this fixes one definition shared with the oracle).
__device__ __forceinline__ int to_int(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483647.0f) return 2147483647;
  if (f <= -2147483648.0f) return -2147483647 - 1;
  return static_cast<int>(f);
}

// `vec4(x, y, z, w) * M` with a row-major upload: out[r] = sum_c M[r][c] v[c],
// left to right, unfused.
__device__ __forceinline__ void mat_vec(const float* m, float x, float y, float z, float w, float* out) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float a = x * m[4 * r + 0];
    float b = y * m[4 * r + 1];
    float c = z * m[4 * r + 2];
    float d = w * m[4 * r + 3];
    out[r] = ((a + b) + c) + d;
  }
}

// Pixel centre (x + 0.5, y + 0.5) with reverse-Z depth -> world position
// (cacheGather.comp:113-119, cacheApply.frag:134-135).
__device__ __forceinline__ f3 unproject_pixel(const float* inv_vp, int w, int h, int x, int y, float depth) {
  const float px = static_cast<float>(x) + 0.5f;
  const float py = static_cast<float>(y) + 0.5f;
  const float sx = px / static_cast<float>(w) * 2.0f - 1.0f;
  const float sy = py / static_cast<float>(h) * 2.0f - 1.0f;
  float v[4];
  mat_vec(inv_vp, sx, sy, depth, 1.0f, v);
  return make3(v[0] / v[3], v[1] / v[3], v[2] / v[3]);
}

// ComputeAddressVolumeCascade, shader/lightcache.glsl:109-122.
__device__ __forceinline__ int cascade_of(const drv_volume& vol, int n, f3 wp) {
  int c = 0;
  for (; c < n - 1; ++c) {
    const drv_cascade& k = vol.cascades[c];
    if (wp.x <= k.decision_max[0] && wp.y <= k.decision_max[1] && wp.z <= k.decision_max[2] &&
        wp.x >= k.decision_min[0] && wp.y >= k.decision_min[1] && wp.z >= k.decision_min[2])
      break;
  }
  return c;
}

// ComputeAddressVolumeCascadeTransition, shader/lightcache.glsl:125-134.
__device__ __forceinline__ float cascade_transition(const drv_volume& vol, int c, f3 wp) {
  const drv_cascade& k = vol.cascades[c];
  const float a0 = k.decision_max[0] - wp.x, a1 = k.decision_max[1] - wp.y, a2 = k.decision_max[2] - wp.z;
  const float b0 = wp.x - k.decision_min[0], b1 = wp.y - k.decision_min[1], b2 = wp.z - k.decision_min[2];
  const float m = fminf(fminf(fminf(a0, a1), a2), fminf(fminf(b0, b1), b2));
  const float t = 1.0f - m / (k.world_voxel_size * vol.cav_transition_zone_size);
  return fminf(1.0f, fmaxf(0.0f, t));
}

__device__ __forceinline__ float half_to_float(uint16_t h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

// UnpackNormal16I, shader/utils.glsl:43-60.
__device__ __forceinline__ f3 unpack_normal16i(int16_t a, int16_t b) {
  const float px = static_cast<float>(a) * (kPi / 32768.0f);
  const float py = static_cast<float>(b) * (1.0f / 32768.0f);
  const float sin_phi = sqrtf(1.0f - py * py);
  float s, c;
  sincosf(px, &s, &c);
  const float x = c * sin_phi, y = s * sin_phi, z = py;
  const float l = sqrtf(x * x + y * y + z * z);
  return make3(x / l, y / l, z / l);
}

// Morton_2D_Decode_16bit, shader/cacheLightingRSM.comp:46-62 (x = even bits).
__device__ __forceinline__ void morton_decode(uint32_t m, uint32_t& ox, uint32_t& oy) {
  uint32_t x = m & 0x55555555u, y = (m >> 1) & 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  y = (y | (y >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0f0f0f0fu;
  y = (y | (y >> 2)) & 0x0f0f0f0fu;
  x = (x | (x >> 4)) & 0x00ff00ffu;
  y = (y | (y >> 4)) & 0x00ff00ffu;
  ox = (x | (x >> 8)) & 0x0000ffffu;
  oy = (y | (y >> 8)) & 0x0000ffffu;
}

}  // namespace drvk
