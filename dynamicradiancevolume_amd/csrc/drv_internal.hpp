// Internal interface between the C-ABI host layer (drv_api.cpp) and the HIP
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/drv.h"

namespace drvk {

// std430 LightCacheEntry stride in floats (shader/lightcache.glsl:33-76).
inline int entry_stride(int sh_bands) { return sh_bands >= 2 ? 32 : 16; }

// LightCacheCounter (shader/lightcache.glsl:83-90) followed by the work
// descriptor the gather derives from it on the device.
struct Counter {
  uint32_t groups_x, groups_y, groups_z;  // as written by cachePrepareLighting.comp
  uint32_t total;                         // TotalLightCacheCount (demanded appends)
  uint32_t live;                          // min(total, capacity)
  uint32_t tiles;                         // ceil(live / kGatherTile)
  uint32_t splits;                        // VPL splits of the gather
  uint32_t pad;
};

constexpr int kGatherTile = 256;   // caches per gather work item (one workgroup)
constexpr int kVplChunk = 64;      // VPLs staged in LDS at a time
constexpr int kVisTile = 256;      // caches per visibility work item

// One VPL ("LightInfo", cacheLightingRSM.comp:34-40) packed for LDS staging.
struct Vpl {
  float px, py, pz, disc_area;
  float fr, fg, fb, pad0;
  float nx, ny, nz, pad1;
};

// One shadow block (cacheLightingRSM.comp:171-192): averaged VAL position and
// the cone aperture distToSphereRad.
struct ShadowBlock {
  float x, y, z, rho;
};

// Voxel volume layout: per mip level l (w = res >> l) the R8 texels, x
// fastest, plus the "corner-expanded" copy: for every lower filter corner
// (i, j, k) in [-1, w-1]^3 the 8 clamped texels of its 2x2x2 footprint packed
// into one uint64 (byte dx + 2 dy + 4 dz), so a trilinear fetch is one load.
struct VoxelLayout {
  int res = 0;
  int levels = 0;                  // log2(res) + 1
  std::vector<size_t> mip_offset;  // byte offsets into mips
  std::vector<size_t> corner_offset;  // uint64 offsets into corners
  size_t mip_bytes = 0;
  size_t corner_count = 0;
};

// Device view of the voxel corner layout passed to kernels.
struct VoxelView {
  const uint64_t* corners;
  uint32_t corner_offset[16];
  int res;
  int levels;
};

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s);

void launch_allocate(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const float* depth,
                     uint32_t* cav, float* caches, Counter* counter, hipStream_t s);
void launch_prepare_dispatch(Counter* counter, uint32_t capacity, hipStream_t s);

void launch_voxelize(const drv_volume& vol, int res, const float* positions, const uint32_t* indices, uint32_t n_tri,
                     const float* world, uint8_t* target, hipStream_t s);
void launch_voxel_blend(int res, const uint8_t* target, uint8_t* volume, float adaption, hipStream_t s);
void launch_voxel_mips(const VoxelLayout& layout, uint8_t* mips, hipStream_t s);
void launch_voxel_corners(const VoxelLayout& layout, const uint8_t* mips, uint64_t* corners, hipStream_t s);

void launch_prepare_vpls(const drv_rsm& rsm, const drv_spot_light& light, Vpl* vpls, ShadowBlock* blocks,
                         hipStream_t s);
void launch_visibility(const drv_volume& vol, const VoxelView& vox, const float* caches, int stride,
                       const Counter* counter, const ShadowBlock* blocks, uint32_t n_blocks, float* visibility,
                       uint32_t vis_pitch, uint64_t* step_counter, int grid, hipStream_t s);
void launch_gather(const drv_sh_constants& shc, int sh_bands, const float* caches, int stride,
                   const Counter* counter, const Vpl* vpls, uint32_t n_vpl, const float* visibility,
                   uint32_t vis_pitch, uint32_t sample_interval, int indirect_shadow, float* partial,
                   uint32_t partial_pitch, int grid, hipStream_t s);
void launch_reduce_sh(const drv_sh_constants& shc, int sh_bands, float* caches, int stride, const Counter* counter,
                      const float* partial, uint32_t partial_pitch, int grid, hipStream_t s);

void launch_apply(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const drv_sh_constants& shc,
                  const drv_gbuffer& gb, const uint32_t* cav, const float* caches, const Counter* counter,
                  const float* srgb_lut, float* out_rgb, int accumulate, hipStream_t s);

}  // namespace drvk
