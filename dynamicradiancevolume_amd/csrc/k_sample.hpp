// Trilinear sampling of the corner-expanded voxel mips, shared by the cone
// tracer and the surrogate's feature extraction.
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {

// Corner-expanded fetch address of point p at mip level l and its trilinear
// weights (GL 4.5 section 8.14.2, clamp to edge).  p is in level-0 texel
// units; the lower filter corner floor(p * 2^-l - 0.5) is clamped to
// [-1, w - 1], which selects the same clamped texels.
struct Footprint {
  uint32_t index;
  float ax, ay, az;
};

// `off` is the view's corner_offset table staged in LDS (a per-lane level
// index into the kernel-argument copy would be a memory load per sample).
__device__ __forceinline__ Footprint footprint(const VoxelView& vx, const uint32_t* off, int l, float px, float py,
                                               float pz) {
  const int w = vx.res >> l;
  const float scale = __builtin_amdgcn_ldexpf(1.0f, -l);
  const float u = __builtin_fmaf(px, scale, -0.5f);
  const float v = __builtin_fmaf(py, scale, -0.5f);
  const float q = __builtin_fmaf(pz, scale, -0.5f);
  const float fu = floorf(u), fv = floorf(v), fq = floorf(q);
  const float hi = static_cast<float>(w - 1);
  // Clamped in float first so the conversion is always in range (NaN -> -1).
  const uint32_t i = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fu, -1.0f, hi) + 1.0f);
  const uint32_t j = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fv, -1.0f, hi) + 1.0f);
  const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fq, -1.0f, hi) + 1.0f);
  const uint32_t wp = static_cast<uint32_t>(w + 1);
  Footprint f;
  f.index = off[l] + __umul24(__umul24(k, wp) + j, wp) + i;
  f.ax = u - fu;
  f.ay = v - fv;
  f.az = q - fq;
  return f;
}

// Trilinear blend of the 8 packed unorm8 texels (byte dx + 2 dy + 4 dz),
// in units of 1/255.
__device__ __forceinline__ float blend8(uint64_t c, float ax, float ay, float az) {
  const uint32_t lo = static_cast<uint32_t>(c), hi = static_cast<uint32_t>(c >> 32);
  const float t000 = static_cast<float>(lo & 0xffu), t100 = static_cast<float>((lo >> 8) & 0xffu);
  const float t010 = static_cast<float>((lo >> 16) & 0xffu), t110 = static_cast<float>(lo >> 24);
  const float t001 = static_cast<float>(hi & 0xffu), t101 = static_cast<float>((hi >> 8) & 0xffu);
  const float t011 = static_cast<float>((hi >> 16) & 0xffu), t111 = static_cast<float>(hi >> 24);
  const float x00 = __builtin_fmaf(ax, t100 - t000, t000);
  const float x10 = __builtin_fmaf(ax, t110 - t010, t010);
  const float x01 = __builtin_fmaf(ax, t101 - t001, t001);
  const float x11 = __builtin_fmaf(ax, t111 - t011, t011);
  const float y0 = __builtin_fmaf(ay, x10 - x00, x00);
  const float y1 = __builtin_fmaf(ay, x11 - x01, x01);
  return __builtin_fmaf(az, y1 - y0, y0);
}

// Single-level sample (0..1) of the voxel mips at level l, p in level-0 texel units.
__device__ __forceinline__ float sample_level(const VoxelView& vx, const uint32_t* off, int l, float px, float py,
                                              float pz) {
  const Footprint f = footprint(vx, off, l, px, py, pz);
  return blend8(vx.corners[f.index], f.ax, f.ay, f.az) * (1.0f / 255.0f);
}

// Stages the corner_offset table in LDS; every thread of the workgroup must
// call it (it synchronises the workgroup).
__device__ __forceinline__ void stage_offsets(const VoxelView& vx, uint32_t* off) {
  if (threadIdx.x < 16u) off[threadIdx.x] = vx.corner_offset[threadIdx.x];
  __syncthreads();
}

}  // namespace drvk
