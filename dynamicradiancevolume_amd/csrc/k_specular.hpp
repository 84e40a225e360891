// Device helpers of the indirect specular path (INDIRECT_SPECULAR): the
// R11G11B10F texel format of the per-cache env-map atlas, the hemisphere
// projection that addresses it and the cache-local view frame.
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {

// GL unsigned 11- and 10-bit floats (5-bit exponent, bias 15, 6 or 5
// mantissa bits; GL 4.5 section 2.3.4.3-4): finite values round to the
// nearest representable finite value (ties to even), negative values and
// -inf become 0, finite values above the largest finite value clamp to it,
// +inf stays inf and NaN stays NaN.
__device__ __forceinline__ uint32_t uf_encode(float x, int m) {
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  const uint32_t max_finite = (30u << m) | ((1u << m) - 1u);
  if (x != x) return (31u << m) | 1u;
  if (!(x > 0.0f)) return 0u;
  if (b == 0x7f800000u) return 31u << m;
  if (x < 6.103515625e-05f) {  // below 2^-14: denormal, value = mantissa * 2^(-14 - m)
    return static_cast<uint32_t>(__builtin_rintf(x * __builtin_ldexpf(1.0f, 14 + m)));
  }
  const uint32_t shift = 23u - static_cast<uint32_t>(m);
  const uint32_t r = (b + ((1u << (shift - 1u)) - 1u) + ((b >> shift) & 1u)) >> shift;
  const int e = static_cast<int>(r >> m) - 127 + 15;
  if (e >= 31) return max_finite;
  return (static_cast<uint32_t>(e) << m) | (r & ((1u << m) - 1u));
}

__device__ __forceinline__ float uf_decode(uint32_t c, int m) {
  const uint32_t e = c >> m, f = c & ((1u << m) - 1u);
  if (e == 0u) return __builtin_ldexpf(static_cast<float>(f), -14 - m);
  if (e == 31u) return f ? __builtin_nanf("") : __builtin_inff();
  return __builtin_bit_cast(float, ((e - 15u + 127u) << 23) | (f << (23u - static_cast<uint32_t>(m))));
}

// GL_UNSIGNED_INT_10F_11F_11F_REV: red in bits 0-10, green 11-21, blue 22-31.
__device__ __forceinline__ uint32_t pack_r11g11b10(float r, float g, float b) {
  return uf_encode(r, 6) | (uf_encode(g, 6) << 11) | (uf_encode(b, 5) << 22);
}

__device__ __forceinline__ void unpack_r11g11b10(uint32_t v, float* rgb) {
  rgb[0] = uf_decode(v & 0x7ffu, 6);
  rgb[1] = uf_decode((v >> 11) & 0x7ffu, 6);
  rgb[2] = uf_decode(v >> 22, 5);
}

// atan2(y, x) for the hemisphere projection, by octant reduction: the
// quotient t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] (hardware reciprocal),
// atan(t) = t P(t^2) with P the degree-7 minimax fit of atan(t) / t on
// [0, 1] (relative error 1e-7; a few ulp in FP32), then the octant, the
// sign of x and the sign of y restore the angle.  atan2(0, 0) = 0.
__device__ __forceinline__ float octant_atan2(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  const float s = t * t;
  float p = -0.004693402908742428f;
  p = __builtin_fmaf(p, s, 0.02425288036465645f);
  p = __builtin_fmaf(p, s, -0.05948710814118385f);
  p = __builtin_fmaf(p, s, 0.09914347529411316f);
  p = __builtin_fmaf(p, s, -0.14019502699375153f);
  p = __builtin_fmaf(p, s, 0.19969728589057922f);
  p = __builtin_fmaf(p, s, -0.33331990242004395f);
  p = __builtin_fmaf(p, s, 0.9999998807907104f);
  float a = t * p;
  a = ay > ax ? 1.5707963267948966f - a : a;
  a = x < 0.0f ? 3.141592653589793f - a : a;
  return y < 0.0f ? -a : a;
}

// HemisphericalProjection with HEMIPROJECTION_LAMBERT_CONCENTRICQUAD
// (shader/utils.glsl:102-178, the "better perf" form): z+ hemisphere ->
// [0, 1]^2, equal-area.  The square root is the hardware one (1 ulp).
__device__ __forceinline__ void hemi_projection(float dx, float dy, float dz, float* out) {
  const float r = __builtin_amdgcn_sqrtf(1.0f - dz) * 0.5f;
  float phi = octant_atan2(dy, dx) * (4.0f / kPi);
  if (phi < -1.0f) phi += 8.0f;
  float px, py;
  if (phi < 3.0f) {
    if (phi < 1.0f) {
      px = r;
      py = phi * r;
    } else {
      px = -(phi - 2.0f) * r;
      py = r;
    }
  } else {
    if (phi < 5.0f) {
      px = -r;
      py = -(phi - 4.0f) * r;
    } else {
      px = (phi - 6.0f) * r;
      py = -r;
    }
  }
  out[0] = px + 0.5f;
  out[1] = py + 0.5f;
}

// ComputeLocalViewSpace (shader/lightcache.glsl:98-106): Z towards the
// camera, X = normalize(Z.z, 0, -Z.x), Y = Z x X.  `v * M` of a GLSL mat3
// with these columns is (v.X, v.Y, v.Z).
struct LocalFrame {
  f3 x, y, z;
};

__device__ __forceinline__ f3 normalize3(f3 v) {
  const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return make3(v.x / l, v.y / l, v.z / l);
}

__device__ __forceinline__ LocalFrame local_view_space(const float* cam, f3 wp) {
  LocalFrame f;
  f.z = normalize3(make3(cam[0] - wp.x, cam[1] - wp.y, cam[2] - wp.z));
  f.x = normalize3(make3(f.z.z, 0.0f, -f.z.x));
  f.y = make3(f.z.y * f.x.z - f.z.z * f.x.y, f.z.z * f.x.x - f.z.x * f.x.z, f.z.x * f.x.y - f.z.y * f.x.x);
  return f;
}

__device__ __forceinline__ f3 to_local(const LocalFrame& f, float vx, float vy, float vz) {
  return make3(__builtin_fmaf(vz, f.x.z, __builtin_fmaf(vy, f.x.y, vx * f.x.x)),
               __builtin_fmaf(vz, f.y.z, __builtin_fmaf(vy, f.y.y, vx * f.y.x)),
               __builtin_fmaf(vz, f.z.z, __builtin_fmaf(vy, f.z.y, vx * f.z.x)));
}

// BlinnNormalization(GetBlinnPhongExponentForHemispherePixel(S))
// (lightingfunctions.glsl:38-41, utils.glsl:180-191): the factor every
// env-map texel is scaled by (cacheLightingRSM.comp:97-98).
__device__ __host__ inline float specular_base_normalization(int S) {
  const float e = static_cast<float>(S) * static_cast<float>(S) - 1.0f;
  return (e + 8.0f) / (8.0f * 3.14159265358979f);
}

}  // namespace drvk
