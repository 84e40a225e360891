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
  uint32_t phantoms;                      // appends of corners outside the address volume
};

constexpr int kGatherTile = 64;    // caches per gather / cone-trace tile (one per lane)
constexpr uint32_t kMaxSplits = 256;  // VPL splits of the SH gather (each at least 64 VPLs)

// One VPL ("LightInfo", cacheLightingRSM.comp:34-40) packed for LDS staging.
struct Vpl {
  float px, py, pz, disc_area;
  float fr, fg, fb, pad0;
  float nx, ny, nz, pad1;
};

// One shadow block (cacheLightingRSM.comp:171-192): averaged VAL position and
// the cone aperture distToSphereRad, the march's step factor
// k = 2 / (1 - distToSphereRad) (cacheLightingRSM.comp:227), plus the mean
// VPL normal of the block (an input of the surrogate visibility network only).
struct ShadowBlock {
  float x, y, z, rho;
  float nx, ny, nz, k;
};

// One record of the corner-expanded voxel copy: the 2x2x2 footprint of a
// lower filter corner as the eight coefficients of its trilinear polynomial
// a + b x + y (c + d x) + z (e + f x + y (g + h x)) in four words (a, b),
// (c, d), (e, f), (g, h), each two IEEE halves, low half first (exact: the
// coefficients are integer differences of texels 0..255, in [-1020, 1020]),
// so each word's x term is one mixed-precision FMA.
struct alignas(16) VoxelCorner {
  uint32_t row[4];
};

// One sample of a shadow block's cone march (cacheLightingRSM.comp:211-228).
// The march distance, the step and the mip level pair of every sample depend
// only on the block's aperture (rho and k), not on the ray, so the VPL
// preparation plans the 32 samples once per block and every ray towards the
// block reads them (wave-uniform, scalar loads).  lv[0] is level floor(lod),
// lv[1] level floor(lod) + 1 when two levels are blended (two = 1, frac > 0),
// each as the constants of the corner-expanded copy: byte offset of the
// level's first record, corners per row and per slice (as floats: exact
// integers below 2^24), w and 2^-l.
struct MarchLevel {
  uint32_t off16;
  float wp, wp2, w, scale;
};
struct alignas(64) MarchStep {
  float step;  // advance to this sample
  float dist;  // march distance after the advance
  float frac;  // weight of lv[1]
  uint32_t two;
  MarchLevel lv[2];
  uint32_t pad[2];
};
static_assert(sizeof(MarchStep) == 64, "MarchStep is one 64-byte scalar load");
constexpr int kMarchSteps = 32;

// Voxel volume layout: per mip level l (w = res >> l) the R8 texels, x
// fastest, plus the "corner-expanded" copy: for every lower filter corner
// (i, j, k) in [-1, w-1]^3 the 8 clamped texels of its 2x2x2 footprint in
// one 16-byte VoxelCorner, so a trilinear fetch is one load.
struct VoxelLayout {
  int res = 0;
  int levels = 0;                  // log2(res) + 1
  std::vector<size_t> mip_offset;  // byte offsets into mips
  std::vector<size_t> corner_offset;  // record offsets into corners
  size_t mip_bytes = 0;
  size_t corner_count = 0;
};

// Device view of the voxel corner layout passed to kernels.
struct VoxelView {
  const VoxelCorner* corners;
  uint32_t corner_offset[16];
  uint32_t corner_bytes;  // < 2^32 (voxel resolution <= 512: 2.5e9 bytes)
  int res;
  int levels;
};

// Surrogate MLP parameter layout (FP32 master): W1 [32][64] (row = input),
// b1 [64], W2 [64][64], b2, W3 [64][64], b3, w4 [64], b4.
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = 2048;
constexpr int kMlpW2 = 2112;
constexpr int kMlpB2 = 6208;
constexpr int kMlpW3 = 6272;
constexpr int kMlpB3 = 10368;
constexpr int kMlpW4 = 10432;
constexpr int kMlpB4 = 10496;
constexpr int kMlpParams = 10497;
constexpr int kMlpPackBytes = 1280 * 16;       // inference A operands
constexpr int kMlpTrainW = 64 * 40 + 4 * 64 * 72;  // bf16 LDS images for training

// Inputs of the surrogate's feature extraction.
struct FeatureArgsHost {
  float vmin[3];
  float voxel_size;
  float cam[3];
  VoxelView vox;
};

// One view of a batched light pass (device descriptor).  The plan kernel
// fills tiles, splits and item_begin from the live cache count.
struct BatchView {
  float* caches;
  const Counter* counter;
  const Vpl* vpls;
  const ShadowBlock* blocks;
  VoxelView vox;
  float vmin[3];
  float voxel_size;
  uint32_t n_vpl;
  uint32_t interval;  // VPLs per shadow block
  uint32_t tiles, splits, item_begin, partial_begin;
};

struct BatchQueue {
  uint32_t head;      // next item to hand out
  uint32_t total;     // items of all views
  uint32_t vpl_tile;  // VPL tile length in effect (grown until the partials fit)
  uint32_t pad;
};

// One view of the batched split pipeline, as the host fills it: the view's
// cache list and buffers, its light (for the VPL and shadow-block records)
// and the most VPL splits its partial buffer holds.
struct SplitViewHost {
  float* caches;
  Counter* counter;
  Vpl* vpls;
  ShadowBlock* blocks;
  MarchStep* march;  // kMarchSteps per block (with blocks)
  float* vis;
  uint32_t vis_pitch;
  float* partial;
  uint32_t partial_pitch;
  uint32_t max_splits;
  VoxelView vox;
  float vmin[3];
  float voxel_size;
  uint32_t n_vpl, interval, n_blocks;
  drv_rsm rsm;
  drv_spot_light light;
};

// Item totals of the three queued launches of a batched split pass.
struct SplitTotals {
  uint32_t vis, gather, reduce;
  uint32_t vis_head;  // next visibility item to hand out (reset by the plan)
};

// Launch shapes chosen per context (drv_set_tuning).
// Defaults measured on boxes_1080p (6352 caches x 16384 VPLs).
struct Tuning {
  int visibility_kernel = 0;   // 0: lane per cache, one block at a time; 1: refilling lane pool
  int blocks_per_item = 1;     // shadow blocks per visibility work item (an upper bound)
  int gather_max_splits = 256; // VPL splits of the gather
  int waves_per_cu = 256;      // launch grid of the light pass kernels, in waves per CU
  int refill_min_lanes = 16;   // lane pool: idle lanes that trigger a refill
  int steps_per_batch = 1;     // cone-trace steps planned (and loaded) together: 1..4 or 8
};

// Per-cache specular env-maps (INDIRECT_SPECULAR): the R11G11B10F atlas of
// Renderer::AllocateCacheData (renderer.cpp:253-286) - caches row-major in
// an atlas of `width` texels, `per_row` caches per row, S x S texels each,
// `levels` = log2(S) + 1 mip levels stored back to back - plus the FP32 work
// sums of the current frame and the per-light spill of each map.
constexpr int kSpecMaxLevels = 6;  // S <= 32
struct SpecAtlas {
  uint32_t* atlas = nullptr;
  Vpl* vpls_lanes = nullptr;  // lane-interleaved VPL records of the current light
  size_t vpls_lanes_cap = 0;
  float* work = nullptr;   // [max_caches][S * S][3]
  float* spill = nullptr;  // [max_caches][2S + 1][3]
  uint32_t width = 0, per_row = 0;
  int S = 0, levels = 0, fill = 0;
  uint64_t level_offset[kSpecMaxLevels] = {};
  uint64_t texels = 0;  // all levels
};

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s);

// Clear of the previous claims, pass 1 (claims, phantom appends) and pass 2
// (count, assign: indices, entries, dispatch counter).  `counts` holds
// kAllocChunks + 1 words (the last: claimed cells of the previous call,
// zero with an all-zero CAV), `claimed` one word per CAV cell, `phantoms`
// max_caches positions.
constexpr int kAllocChunks = 1024;
void launch_allocate(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const float* depth,
                     uint32_t* cav, float* caches, Counter* counter, float4* phantoms, uint32_t* counts,
                     uint32_t* claimed, hipStream_t s);

void launch_voxelize(const drv_volume& vol, int res, const float* positions, const uint32_t* indices, uint32_t n_tri,
                     const float* world, uint8_t* target, hipStream_t s);
// Blend and the levels inside each 16^3 brick (one launch); then the levels
// above the brick size and the corner-expanded copy of every level (one
// launch, its first workgroup building the upper levels).
void launch_voxel_blend_mips(const VoxelLayout& layout, const uint8_t* target, uint8_t* mips, float adaption,
                             hipStream_t s);
void launch_voxel_corners(const VoxelLayout& L, uint8_t* mips, VoxelCorner* corners, hipStream_t s);

// Split plan of the gather for a cache counter: enough (tile, split) items
// for the persistent grid (4 waves per workgroup), each split a whole number
// of quanta (shadow blocks, at least 64 VPLs).
struct GatherPlan {
  Counter* counter;
  uint32_t quantum, max_splits, target_items;
};
inline GatherPlan gather_plan(Counter* counter, uint32_t interval, uint32_t max_splits, int grid) {
  return GatherPlan{counter, interval > 64u ? interval : 64u, max_splits, static_cast<uint32_t>(grid) * 4u};
}

// flux_cut: INDIRECT_SPECULAR skips VALs whose flux sums below 0.001
// (cacheLightingRSM.comp:240-243); their flux is stored as zero.
// vpls_lanes (optional): the same records lane-interleaved for the specular
// gather, record m at (m % per_lane) * 64 + m / per_lane, per_lane = n / 64.
// plan (optional): the gather's split plan, written by the first thread.
// march (optional, with blocks): kMarchSteps MarchStep records per shadow
// block, planned for the voxel levels of `vox`.
void launch_prepare_vpls(const drv_rsm& rsm, const drv_spot_light& light, Vpl* vpls, ShadowBlock* blocks,
                         int flux_cut, Vpl* vpls_lanes, const GatherPlan* plan, hipStream_t s,
                         MarchStep* march = nullptr, const VoxelView* vox = nullptr);
// march (optional): the blocks' march plans; without them (or with a tuning
// other than one block walk at one step per batch) the rays plan their own
// samples, with the same values.
void launch_visibility(const drv_volume& vol, const VoxelView& vox, const float* caches, int stride,
                       const Counter* counter, const ShadowBlock* blocks, uint32_t n_blocks, float* visibility,
                       uint32_t vis_pitch, uint64_t* step_counter, const Tuning& tune, int grid, hipStream_t s,
                       const MarchStep* march = nullptr);
void launch_gather(const drv_sh_constants& shc, int sh_bands, const float* caches, int stride,
                   const Counter* counter, const Vpl* vpls, uint32_t n_vpl, const float* visibility,
                   uint32_t vis_pitch, uint32_t sample_interval, int indirect_shadow, float* partial,
                   uint32_t partial_pitch, uint32_t max_splits, int grid, bool plan, hipStream_t s);
void launch_reduce_sh(const drv_sh_constants& shc, int sh_bands, float* caches, int stride, const Counter* counter,
                      const float* partial, uint32_t partial_pitch, int grid, hipStream_t s);

FeatureArgsHost make_feature_args(const drv_volume& vol, const float* cam, const VoxelView& vox);
size_t mlp_train_lds_bytes();
void launch_mlp_pack(const float* params, void* pack, void* train_w, hipStream_t s);
void launch_visibility_mlp(const FeatureArgsHost& fh, const void* pack, const float* params, const float* caches,
                           int stride, const Counter* counter, const Counter* first, const ShadowBlock* blocks,
                           uint32_t n_blocks, float* vis, uint32_t pitch, int grid, hipStream_t s);
void launch_mlp_sample_pairs(const Counter* counter, uint32_t n_blocks, uint32_t n, uint32_t seed, uint32_t* pairs,
                             uint32_t* n_out, hipStream_t s);
void launch_mlp_features(const FeatureArgsHost& fh, const float* caches, int stride, uint32_t n_caches,
                         const ShadowBlock* blocks, uint32_t n_blocks, const uint32_t* pairs, uint32_t n, float* x,
                         const float* vis, uint32_t pitch, float* y, hipStream_t s);
void launch_mlp_forward(const void* pack, const float* params, const float* x, uint32_t n, float* p, int grid,
                        hipStream_t s);
void launch_mlp_batch_indices(uint32_t n, uint32_t batch, uint32_t seed, uint32_t* idx, bool all, hipStream_t s);
uint32_t mlp_train_waves(uint32_t batch, int cu_count);
void launch_mlp_train(const float* x, const float* y, const uint32_t* idx, uint32_t batch, const float* params,
                      const void* train_w, float* partial, uint32_t n_waves, hipStream_t s);
void launch_mlp_adam(const float* partial, uint32_t n_waves, float* params, float* m, float* v, float* grad_out,
                     float lr, float beta1, float beta2, float eps, float bc1, float bc2, hipStream_t s);

void launch_batch(BatchView* views, uint32_t n_views, BatchQueue* q, int stride, int sh_bands, int shadow,
                  uint32_t cpw, uint32_t vpl_tile, float* partial, size_t partial_floats,
                  const drv_sh_constants& shc, int grid, hipStream_t s);

// Allocation-free light pass: a file-scope device array of
// kStaticScratchBytes (VPL and shadow-block records of at most
// kStaticMaxVpls VPLs, the queue descriptor and split partials).
// static_scratch_init (once per device, with that device current) resolves
// its device address; static_scratch returns it (null before init).
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
constexpr uint32_t kStaticMaxVpls = 65536;
hipError_t static_scratch_init(int device);
unsigned char* static_scratch(int device);
void launch_light_static(unsigned char* base, const drv_rsm& rsm, const drv_spot_light& light, BatchView view,
                         int stride, int sh_bands, int shadow, const drv_sh_constants& shc, int grid,
                         hipStream_t s);

// Batched split pass: the device descriptors of n views (split_desc_bytes()
// each, built on the host by build_split_descs and uploaded by the caller),
// then one launch each for the VPL records of every view, the plan (split
// counts and item offsets from the live counts), the cone-traced visibility,
// the SH gather and the split reduction, each over the items of every view.
size_t split_desc_bytes();
void build_split_descs(const SplitViewHost* views, uint32_t n_views, void* host_out);
// The split queue's launches: VPL/block/march records and the plan of every
// view; the queued visibility; the queued SH gather and reduction.
// max_march: the most march records of any view (n_blocks x kMarchSteps,
// 0 without indirect shadows).
void launch_batch_split_prepare(void* descs, uint32_t n_views, SplitTotals* totals, int sh_bands, int shadow,
                                uint32_t max_vpl, uint32_t max_march, uint32_t target_items, hipStream_t s);
void launch_batch_split_visibility(void* descs, uint32_t n_views, SplitTotals* totals, int stride, int grid,
                                   hipStream_t s);
void launch_batch_split_gather(void* descs, uint32_t n_views, SplitTotals* totals, int stride, int sh_bands,
                               int shadow, const drv_sh_constants& shc, int grid, hipStream_t s);

void launch_spot_select(const Counter* counter, uint32_t k, const float* caches, int stride, float* spot,
                        uint32_t spot_cap, Counter* spot_ctr, hipStream_t s);
void launch_spot_error(const float* spot, uint32_t spot_cap, int stride, int sh_bands, const Counter* spot_ctr,
                       float threshold, const Counter* counter, uint32_t* stats, Counter* gate, hipStream_t s);

void launch_render_gbuffer(const float* positions, const float* albedo, const float* normals, uint32_t n_tri,
                           const float* view_proj, int w, int h, unsigned long long* zbits, uint32_t* tri,
                           uint32_t* big, float* depth, int16_t* normal, uint8_t* albedo_srgb, hipStream_t s);
// big: 1 + n_tri uint32 (count, then the list of large triangles)
void launch_render_rsm(const float* positions, const float* albedo, const float* normals, uint32_t n_tri,
                       const drv_spot_light& light, unsigned long long* zbits, uint32_t* tri, uint32_t* big,
                       uint16_t* flux, int16_t* normal, uint16_t* dls, float* shadow, hipStream_t s);
void launch_direct(const drv_config& cfg, const drv_frame& frame, const drv_spot_light& light, const drv_gbuffer& gb,
                   const uint8_t* roughness_metallic, const float* srgb_lut, const float* shadow, int shadow_res,
                   float normal_offset, float bias, float* out_rgb, int accumulate, hipStream_t s);
void launch_tonemap(const float* hdr, size_t n, float exposure, float divider, float* out, hipStream_t s);

void launch_bruteforce(const drv_config& cfg, const drv_frame& frame, const drv_spot_light& light,
                       const drv_gbuffer& gb, const uint8_t* roughness_metallic, const float* srgb_lut,
                       const Vpl* vals, uint32_t n_vals, int diffuse_only, float* out_rgb, int accumulate,
                       hipStream_t s);

// spec: the specular atlas when INDIRECT_SPECULAR is on, else null.
void launch_apply(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const drv_sh_constants& shc,
                  const drv_gbuffer& gb, const uint32_t* cav, const float* caches, const Counter* counter,
                  const float* srgb_lut, const SpecAtlas* spec, float* out_rgb, int accumulate, hipStream_t s);

// Specular gather of one light into the work sums (first != 0: overwrite),
// then quantization, mips and hole filling of every live cache's region.
void launch_specular(const float* cam, const float* caches, int stride, const Counter* counter, const Vpl* vpls,
                     uint32_t n_vpl, const float* vis, uint32_t vis_pitch, uint32_t interval, const SpecAtlas& at,
                     int first, int grid, hipStream_t s);

}  // namespace drvk
