// Trilinear sampling of the corner-expanded voxel mips, shared by the cone
// tracer and the surrogate's feature extraction.
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {

// Per-level constants of the corner-expanded copy, staged in LDS by every
// sampling kernel (a per-lane level index into a kernel-argument table would
// be a memory load on the address path of every tap): corner offset, corners
// per row (w + 1) and per slice ((w + 1)^2), w = res >> l as float and 2^-l.
struct __attribute__((aligned(32))) LevelInfo {  // 32 B: the table index is a shift
  uint32_t off;
  uint32_t wp;
  uint32_t wp2;
  float w;
  float scale;
  uint32_t pad[3];
};

__device__ __forceinline__ LevelInfo level_info(const VoxelView& vx, int l) {
  LevelInfo li;
  const int w = vx.res >> l;
  li.off = vx.corner_offset[l];
  li.wp = static_cast<uint32_t>(w + 1);
  li.wp2 = li.wp * li.wp;
  li.w = static_cast<float>(w);
  li.scale = __builtin_amdgcn_ldexpf(1.0f, -l);
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

__device__ __forceinline__ corner_t corner_load(__amdgpu_buffer_rsrc_t r, uint32_t index) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(index * 16u), 0, 0);
}

// Corner-expanded fetch index of point p at mip level l and its trilinear
// weights (GL 4.5 section 8.14.2, clamp to edge).  p is in level-0 texel
// units; the lower filter corner floor(p * 2^-l - 0.5) clamped to [-1, w - 1]
// selects the same clamped texels, and is computed one up, as
// floor(p * 2^-l + 0.5) clamped to [0, w] (the corner table's row index).
struct Footprint {
  uint32_t index;
  float ax, ay, az;
};

__device__ __forceinline__ Footprint footprint(const LevelInfo* lv, int l, float px, float py, float pz) {
  const LevelInfo li = lv[l];
  const float u = __builtin_fmaf(px, li.scale, 0.5f);
  const float v = __builtin_fmaf(py, li.scale, 0.5f);
  const float q = __builtin_fmaf(pz, li.scale, 0.5f);
  // Clamped in float first so the conversion is always in range (NaN -> 0);
  // w is an integer, so truncating the clamped value is floor(u) clamped.
  const uint32_t i = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(u, 0.0f, li.w));
  const uint32_t j = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(v, 0.0f, li.w));
  const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(q, 0.0f, li.w));
  Footprint f;
  f.index = li.off + i + __umul24(j, li.wp) + __umul24(k, li.wp2);  // (w + 1)^2 < 2^24 for res <= 512
  // u - floor(u) (v_fract_f32 caps it below 1, which only happens for
  // u in (-2^-24, 0), where both texels of the footprint are the clamped
  // edge texel and the weight has no effect).
  f.ax = __builtin_amdgcn_fractf(u);
  f.ay = __builtin_amdgcn_fractf(v);
  f.az = __builtin_amdgcn_fractf(q);
  return f;
}

// The two halves of a corner row as floats (exact).
__device__ __forceinline__ float half_lo(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu)));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}

// Trilinear blend of the 8 texels of a corner record, in units of 1/255:
// x lerps t0 + ax (t1 - t0), each one mixed-precision FMA of the record's
// halves (the same FP32 value as converting both texels first), then y and z.
__device__ __forceinline__ float blend_corner(corner_t c, float ax, float ay, float az) {
  const float x00 = __builtin_fmaf(ax, half_hi(c[0]), half_lo(c[0]));
  const float x10 = __builtin_fmaf(ax, half_hi(c[1]), half_lo(c[1]));
  const float x01 = __builtin_fmaf(ax, half_hi(c[2]), half_lo(c[2]));
  const float x11 = __builtin_fmaf(ax, half_hi(c[3]), half_lo(c[3]));
  const float y0 = __builtin_fmaf(ay, x10 - x00, x00);
  const float y1 = __builtin_fmaf(ay, x11 - x01, x01);
  return __builtin_fmaf(az, y1 - y0, y0);
}

// Single-level sample (0..1) of the voxel mips at level l, p in level-0 texel units.
__device__ __forceinline__ float sample_level(__amdgpu_buffer_rsrc_t r, const LevelInfo* lv, int l, float px,
                                              float py, float pz) {
  const Footprint f = footprint(lv, l, px, py, pz);
  return blend_corner(corner_load(r, f.index), f.ax, f.ay, f.az) * (1.0f / 255.0f);
}

}  // namespace drvk
