// Trilinear sampling of the corner-expanded voxel mips, shared by the cone
// tracer and the surrogate's feature extraction.
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {

// Per-level constants of the corner-expanded copy, staged in LDS by every
// sampling kernel (a per-lane level index into a kernel-argument table would
// be a memory load on the address path of every tap): byte offset of the
// level's first corner record, corners per row (w + 1) and per slice
// ((w + 1)^2), w = res >> l as float and 2^-l.  Entries past the last level
// repeat the last level, so entry l + 1 can be read beside entry l for any
// level l (the cone trace fetches both with one address).
struct __attribute__((aligned(32))) LevelInfo {  // 32 B: the table index is a shift
  uint32_t off16;
  uint32_t wp;
  uint32_t wp2;
  float w;
  float scale;
  uint32_t pad[3];
};

__device__ __forceinline__ LevelInfo level_info(const VoxelView& vx, int l) {
  LevelInfo li;
  const int lc = l < vx.levels ? l : vx.levels - 1;
  const int w = vx.res >> lc;
  li.off16 = vx.corner_offset[lc] * 16u;
  li.wp = static_cast<uint32_t>(w + 1);
  li.wp2 = li.wp * li.wp;
  li.w = static_cast<float>(w);
  li.scale = __builtin_amdgcn_ldexpf(1.0f, -lc);
  li.pad[0] = li.pad[1] = li.pad[2] = 0u;
  return li;
}

// Stages the level table in LDS; every thread of the workgroup must call it
// (it synchronises the workgroup).
__device__ __forceinline__ void stage_levels(const VoxelView& vx, LevelInfo* lv) {
  if (threadIdx.x < 16u) lv[threadIdx.x] = level_info(vx, static_cast<int>(threadIdx.x));
  __syncthreads();
}

// The corner-expanded copy as a buffer resource: fetches take a 32-bit byte
// offset (no 64-bit address arithmetic per tap) and are bounds-checked.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t corner_rsrc(const VoxelView& vx) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<VoxelCorner*>(vx.corners), static_cast<short>(0),
                                           static_cast<int>(vx.corner_bytes), 0x00020000);
}

typedef uint32_t corner_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ corner_t corner_load(__amdgpu_buffer_rsrc_t r, uint32_t offset) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(offset), 0, 0);
}

// Corner-expanded fetch offset of point p at a mip level and its trilinear
// weights (GL 4.5 section 8.14.2, clamp to edge).  p is in level-0 texel
// units; the lower filter corner floor(p * 2^-l - 0.5) clamped to [-1, w - 1]
// selects the same clamped texels, and is computed one up, as
// floor(p * 2^-l + 0.5) clamped to [0, w] (the corner table's row index).
struct Footprint {
  uint32_t offset;  // byte offset of the corner record
  float ax, ay, az;
};

// L: LevelInfo (off16, wp, wp2, w, scale).
template <class L>
__device__ __forceinline__ Footprint footprint(const L& li, float px, float py, float pz) {
  const float u = __builtin_fmaf(px, li.scale, 0.5f);
  const float v = __builtin_fmaf(py, li.scale, 0.5f);
  const float q = __builtin_fmaf(pz, li.scale, 0.5f);
  // Clamped in float first so the conversion is always in range (NaN -> 0);
  // w is an integer, so truncating the clamped value is floor(u) clamped.
  const uint32_t i = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(u, 0.0f, li.w));
  const uint32_t j = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(v, 0.0f, li.w));
  const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(q, 0.0f, li.w));
  Footprint f;
  // (w + 1)^2 < 2^24 for res <= 512: two 24-bit multiply-adds
  f.offset = ((__umul24(k, li.wp2) + (__umul24(j, li.wp) + i)) << 4) + li.off16;
  // u - floor(u) (v_fract_f32 caps it below 1, which only happens for
  // u in (-2^-24, 0), where both texels of the footprint are the clamped
  // edge texel and the weight has no effect).
  f.ax = __builtin_amdgcn_fractf(u);
  f.ay = __builtin_amdgcn_fractf(v);
  f.az = __builtin_amdgcn_fractf(q);
  return f;
}

// The footprint with a march record's level.  The record index i + j wp +
// k wp^2 is formed in FP32 from the truncated coordinates (exact while it is
// below 2^24, kFloatIndex: a corner copy of at most 2^24 records per level,
// i.e. res <= 128) with one conversion, instead of three float-to-integer
// conversions and two 24-bit multiplies (a conversion or a 24-bit multiply
// costs about two FP32 adds); without kFloatIndex only i + j wp is (res <=
// 512).  The offset equals the integer form's bit for bit.
template <bool kFloatIndex>
__device__ __forceinline__ Footprint footprint(const MarchLevel& li, float px, float py, float pz) {
  const float u = __builtin_fmaf(px, li.scale, 0.5f);
  const float v = __builtin_fmaf(py, li.scale, 0.5f);
  const float q = __builtin_fmaf(pz, li.scale, 0.5f);
  // Clamped first, so the truncation is floor(u) clamped to [0, w].
  const float fi = __builtin_truncf(__builtin_amdgcn_fmed3f(u, 0.0f, li.w));
  const float fj = __builtin_truncf(__builtin_amdgcn_fmed3f(v, 0.0f, li.w));
  const float fk = __builtin_truncf(__builtin_amdgcn_fmed3f(q, 0.0f, li.w));
  const float ij = __builtin_fmaf(fj, li.wp, fi);
  Footprint f;
  if (kFloatIndex) {
    f.offset = (static_cast<uint32_t>(__builtin_fmaf(fk, li.wp2, ij)) << 4) + li.off16;
  } else {
    const uint32_t k = static_cast<uint32_t>(fk), wp2 = static_cast<uint32_t>(li.wp2);
    f.offset = ((__umul24(k, wp2) + static_cast<uint32_t>(ij)) << 4) + li.off16;
  }
  f.ax = __builtin_amdgcn_fractf(u);
  f.ay = __builtin_amdgcn_fractf(v);
  f.az = __builtin_amdgcn_fractf(q);
  return f;
}

__device__ __forceinline__ Footprint footprint(const LevelInfo* lv, int l, float px, float py, float pz) {
  return footprint(lv[l], px, py, pz);
}

// The two halves of a corner record word as floats (exact).
__device__ __forceinline__ float half_lo(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu)));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}

// Trilinear blend of a corner record, in units of 1/255.  The record holds
// the footprint's texels t(x, y, z), x, y, z in {0, 1}, as the coefficients
// of the trilinear polynomial
//   a + b x + y (c + d x) + z (e + f x + y (g + h x)),
// words (a, b), (c, d), (e, f), (g, h), low half first: a = t000,
// b = t100 - t000, c and e the y and z differences of t at x = 0, d, f the
// mixed second differences and g, h the y-z differences of those.  All are
// integers in [-1020, 1020], exact in FP16.  Each word's x term is one
// mixed-precision FMA (FP16 operands, FP32 arithmetic), then one FMA each in
// y and z: 7 operations against 10 for lerps of lerps.
__device__ __forceinline__ float blend_corner(corner_t c, float ax, float ay, float az) {
  const float s0 = __builtin_fmaf(ax, half_hi(c[0]), half_lo(c[0]));
  const float s1 = __builtin_fmaf(ax, half_hi(c[1]), half_lo(c[1]));
  const float s2 = __builtin_fmaf(ax, half_hi(c[2]), half_lo(c[2]));
  const float s3 = __builtin_fmaf(ax, half_hi(c[3]), half_lo(c[3]));
  const float y0 = __builtin_fmaf(ay, s1, s0);
  const float y1 = __builtin_fmaf(ay, s3, s2);
  return __builtin_fmaf(az, y1, y0);
}

// Record n of a shadow block's march plan (see MarchStep): the per-ray
// recurrence of the cone trace with the ray-independent quantities only,
// the same FP32 operations, so the values are those each ray would compute.
__device__ __forceinline__ MarchStep plan_march_step(const VoxelView& vx, float rho, float k, int n) {
  const int q = vx.levels - 1;
  float dist = 0.0f, step = 1.0f, r;
  for (int i = 0;; ++i) {
    dist = dist + step;
    r = dist * rho;
    if (i == n) break;
    step = fmaxf(1.0f, r * k);
  }
  const float lod = __builtin_amdgcn_fmed3f(__builtin_amdgcn_logf(r), 0.0f, static_cast<float>(q));
  const int l = static_cast<int>(lod);
  MarchStep e;
  e.step = step;
  e.dist = dist;
  e.frac = __builtin_amdgcn_fractf(lod);
  e.two = e.frac > 0.0f ? 1u : 0u;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const LevelInfo li = level_info(vx, min(l + h, q));
    e.lv[h] = MarchLevel{li.off16, static_cast<float>(li.wp), static_cast<float>(li.wp2), li.w, li.scale};
  }
  e.pad[0] = e.pad[1] = 0u;
  return e;
}

// Single-level sample (0..1) of the voxel mips at level l, p in level-0 texel units.
__device__ __forceinline__ float sample_level(__amdgpu_buffer_rsrc_t r, const LevelInfo* lv, int l, float px,
                                              float py, float pz) {
  const Footprint f = footprint(lv, l, px, py, pz);
  return blend_corner(corner_load(r, f.offset), f.ax, f.ay, f.az) * (1.0f / 255.0f);
}

}  // namespace drvk
