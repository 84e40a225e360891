// Stage C: per-cache VPL gather with cone-traced indirect shadows projected
// onto spherical harmonics.
//
// Restates shader/cacheLightingRSM.comp:80-391 (SH1/SH2 with
// INDIRECT_SHADOW) driven as in Renderer::LightCachesRSM
// (rendering/renderer.cpp:899-933).
//
// MI355X design.  The reference runs one thread per cache and rebuilds the
// 64-VPL LDS tile in every workgroup, then cone-traces inline once per block
// of B^2 Morton-consecutive VPLs.  Here the pass is split into four kernels:
//   1. k_prepare_vpls   - every VPL record (LightInfo) and every shadow block
//                         record (averaged VAL position, cone aperture) is
//                         built once per light into HBM;
//   2. k_visibility     - one lane per (cache, shadow block) cone trace over
//                         the corner-expanded voxel mips (one 16-byte load
//                         per level per step), written block-major so both
//                         writes and the gather's reads coalesce;
//   3. k_gather         - persistent workgroups pull (64-cache tile, group
//                         of 4 VPL splits) items, one split per wave; VPL
//                         records are wave-uniform, so they stream through
//                         the scalar cache into SGPRs and every lane
//                         accumulates 27 SH sums in VGPRs; the four waves'
//                         sums are combined in LDS and one partial per group
//                         is written; SH constants are folded in at write-out;
//   4. k_reduce         - sums the per-group partials in a fixed order and
//                         adds them into the cache entries (the reference's
//                         `Entry.SH += accumulated`).
// Splitting the VPL range fills 256 CUs even when there are only a few
// thousand caches; the split count is derived on the device from the live
// cache count, so no host round trip is needed.
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_sample.hpp"
#include "k_sh.hpp"
#include "k_trace.hpp"
#include "k_light.hpp"
#include "split_plan.hpp"

namespace drvk {
namespace {

// Split planning of the gather: enough (tile, split) items for the grid,
// each split a whole number of quanta (shadow blocks, at least 64 VPLs).
__device__ __forceinline__ void plan_splits(Counter* ctr, uint32_t n_vpl, uint32_t quantum, uint32_t max_splits,
                                            uint32_t target_items) {
  ctr->splits = plan_split_count(ctr->live, n_vpl, quantum, max_splits, target_items);
}

// Workgroups [0, ceil(n / 256)) prepare the VPL and block records; with
// march plans, the workgroups after them write one march record per thread.
__global__ __launch_bounds__(256) void k_prepare_vpls(PrepArgs a) {
  const uint32_t n = static_cast<uint32_t>(a.R) * static_cast<uint32_t>(a.H);
  const uint32_t vpl_groups = (n + 255u) / 256u;
  if (blockIdx.x >= vpl_groups) {
    const uint32_t idx = (blockIdx.x - vpl_groups) * blockDim.x + threadIdx.x;
    if (idx < n / a.interval * kMarchSteps) prepare_march(a, idx);
    return;
  }
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0u && a.plan) plan_splits(a.plan, n, a.quantum, a.max_splits, a.target_items);
  if (m >= n) return;
  prepare_vpl(a, m, n);
}

// ---- cone-traced visibility ---------------------------------------------

struct VisArgs {
  VoxelView vox;
  float vmin[3];
  float voxel_size;
  const float* caches;
  int stride;
  const Counter* counter;
  const ShadowBlock* blocks;
  uint32_t n_blocks;
  float* vis;
  uint32_t pitch;
  uint64_t* steps;  // optional {traces, reference steps, executed steps}
  uint32_t blocks_per_item;
  uint32_t refill_min;  // lane pool: idle lanes that trigger a refill
  const MarchStep* march;  // the blocks' march plans (k_visibility_planned)
};

// Shadow blocks per work item: the requested count, or one when there are
// too few items to occupy the grid (a small spot-check set).
__device__ __forceinline__ uint32_t items_blocks(uint32_t tiles, uint32_t n_blocks, uint32_t want,
                                                 uint32_t n_waves) {
  // The largest count up to `want` that still leaves two items per wave.
  uint32_t bpi = want;
  while (bpi > 1u && tiles * ((n_blocks + bpi - 1) / bpi) < 2u * n_waves) --bpi;
  return bpi;
}

__device__ __forceinline__ void count_steps(uint64_t* steps, uint32_t lane, uint32_t my_traces, uint32_t my_steps,
                                            uint32_t my_exec) {
  // One atomic triple per wave after a butterfly sum.
  uint64_t tr = my_traces, sv = my_steps, ex = my_exec;
  for (int off = 32; off > 0; off >>= 1) {
    tr += __shfl_xor(tr, off);
    sv += __shfl_xor(sv, off);
    ex += __shfl_xor(ex, off);
  }
  if (lane == 0 && (tr | sv)) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&steps[0]), static_cast<unsigned long long>(tr));
    atomicAdd(reinterpret_cast<unsigned long long*>(&steps[1]), static_cast<unsigned long long>(sv));
    atomicAdd(reinterpret_cast<unsigned long long*>(&steps[2]), static_cast<unsigned long long>(ex));
  }
}

// A wave owns 64 caches x blocks_per_item shadow blocks; each lane traces
// its cache's blocks in order (rays from neighbouring caches towards nearby
// VPL blocks stay coherent).  Items are numbered block-group-major (tile
// fastest) and dealt to the persistent waves round robin, so a wave's items
// reach different VPL blocks: rays to some blocks are much longer than to
// others, and a wave whose items all shared one block group (as with the
// tile-major numbering whenever the group count divides the wave count) left
// a tail of slow waves.
template <int kBatch>
__global__ __launch_bounds__(256) void k_visibility(VisArgs a) {
  __shared__ LevelInfo lvi[16];
  stage_levels(a.vox, lvi);
  const __amdgpu_buffer_rsrc_t cr = corner_rsrc(a.vox);
  const uint32_t live = a.counter->live;
  const uint32_t tiles = (live + 63u) / 64u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t bpi = items_blocks(tiles, a.n_blocks, a.blocks_per_item, n_waves);
  const uint32_t groups = (a.n_blocks + bpi - 1) / bpi;
  const uint32_t items = tiles * groups;
  const float fres = static_cast<float>(a.vox.res);
  const float vsz = a.voxel_size * fres;
  uint32_t my_traces = 0, my_steps = 0, my_exec = 0;
  for (uint32_t item = wave; item < items; item += n_waves) {
    const uint32_t group = item / tiles, tile = item - group * tiles;
    const uint32_t cache = tile * 64u + lane;
    if (cache >= live) continue;
    const float* e = a.caches + static_cast<size_t>(cache) * a.stride;
    const f3 wp = make3(e[0], e[1], e[2]);
    // cacheLightingRSM.comp:103-104
    const f3 vpos = make3((wp.x - a.vmin[0]) / vsz, (wp.y - a.vmin[1]) / vsz, (wp.z - a.vmin[2]) / vsz);
    const uint32_t b0 = group * bpi;
    const uint32_t b1 = min(a.n_blocks, b0 + bpi);
    // Each lane walks its blocks on its own: when its ray ends it starts the
    // next block at the next batch (the record was fetched a trace ahead),
    // so lanes idle only once per item instead of behind the longest ray of
    // every block.
    uint32_t b = b0;
    Trace t = trace_setup(a.vox, a.voxel_size, wp, vpos, a.blocks[b]);
    ShadowBlock nb = a.blocks[min(b + 1u, b1 - 1u)];
    while (true) {
      if (trace_batch<kBatch>(a.vox, cr, lvi, t)) {
        a.vis[static_cast<size_t>(b) * a.pitch + cache] = trace_visibility(t);
        my_traces += 1u;
        my_exec += static_cast<uint32_t>(t.s);
        if (a.steps) my_steps += static_cast<uint32_t>(reference_steps(t));
        if (++b >= b1) break;
        t = trace_setup(a.vox, a.voxel_size, wp, vpos, nb);
        nb = a.blocks[min(b + 1u, b1 - 1u)];
      }
    }
  }
  if (a.steps) count_steps(a.steps, lane, my_traces, my_steps, my_exec);
}

// The traces of k_visibility with the blocks' march plans: a wave's 64 rays
// share their block, so the march distance, the step and the level pair of
// every sample (and the block record) are wave-uniform scalar loads, and
// each lane only advances, fetches and blends.  The wave traces its item's
// blocks one after the other, all lanes on the same block.
template <bool kCount, bool kFloatIndex>
__global__ __launch_bounds__(256) void k_visibility_planned(VisArgs a) {
  const __amdgpu_buffer_rsrc_t cr = corner_rsrc(a.vox);
  const uint32_t live = a.counter->live;
  const uint32_t tiles = (live + 63u) / 64u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t bpi = items_blocks(tiles, a.n_blocks, a.blocks_per_item, n_waves);
  const uint32_t groups = (a.n_blocks + bpi - 1) / bpi;
  const uint32_t items = tiles * groups;
  const float fres = static_cast<float>(a.vox.res);
  const float vsz = a.voxel_size * fres;
  uint32_t my_traces = 0, my_steps = 0, my_exec = 0;
  for (uint32_t item = wave; item < items; item += n_waves) {
    const uint32_t group = item / tiles, tile = item - group * tiles;
    const uint32_t cache = tile * 64u + lane;
    const bool valid = cache < live;
    const float* e = a.caches + static_cast<size_t>(valid ? cache : live - 1u) * a.stride;
    const f3 wp = make3(e[0], e[1], e[2]);
    // cacheLightingRSM.comp:103-104
    const f3 vpos = make3((wp.x - a.vmin[0]) / vsz, (wp.y - a.vmin[1]) / vsz, (wp.z - a.vmin[2]) / vsz);
    const uint32_t b0 = group * bpi;
    const uint32_t b1 = min(a.n_blocks, b0 + bpi);
    for (uint32_t b = b0; b < b1; ++b) {
      Trace t = trace_setup(a.vox, a.voxel_size, wp, vpos, block_record(a.blocks, b));
      trace_planned<kCount, kFloatIndex>(march_steps(a.march, b), cr, t, valid);
      if (valid) {
        a.vis[static_cast<size_t>(b) * a.pitch + cache] = trace_visibility(t);
        if (kCount) {
          my_traces += 1u;
          my_exec += static_cast<uint32_t>(t.s);
          my_steps += static_cast<uint32_t>(
              reference_steps_planned(a.march + static_cast<size_t>(b) * kMarchSteps, t));
        }
      }
    }
  }
  if (kCount) count_steps(a.steps, lane, my_traces, my_steps, my_exec);
}

// The same traces with lane refilling: a work item's n_valid x n_blocks
// traces form a pool (block-major, so lanes that refill together mostly
// share a block), every lane holds one trace, and after each batch of steps
// the lanes whose trace ended take the next pool entries (ballot + mbcnt
// ranks) once at least refill_min lanes are idle, so lanes do not idle
// behind the longest ray of the wave.  The tile's cache positions are staged
// in LDS, so a refill costs no global load.
template <int kBatch>
__global__ __launch_bounds__(256) void k_visibility_pool(VisArgs a) {
  __shared__ LevelInfo lvi[16];
  __shared__ float4 cpos[4][64][2];  // per wave: the tile's cache positions (world, voxel units)
  stage_levels(a.vox, lvi);
  const __amdgpu_buffer_rsrc_t cr = corner_rsrc(a.vox);
  const uint32_t live = a.counter->live;
  const uint32_t tiles = (live + 63u) / 64u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t bpi = items_blocks(tiles, a.n_blocks, a.blocks_per_item, n_waves);
  const uint32_t groups = (a.n_blocks + bpi - 1) / bpi;
  const uint32_t items = tiles * groups;
  const float fres = static_cast<float>(a.vox.res);
  const float vsz = a.voxel_size * fres;
  uint32_t my_traces = 0, my_steps = 0, my_exec = 0;
  for (uint32_t item = wave; item < items; item += n_waves) {
    const uint32_t group = item / tiles, tile = item - group * tiles;
    const uint32_t n_valid = min(64u, live - tile * 64u);
    const uint32_t b0 = group * bpi;
    const uint32_t pool = n_valid * (min(a.n_blocks, b0 + bpi) - b0);
    if (lane < n_valid) {
      const float* e = a.caches + static_cast<size_t>(tile * 64u + lane) * a.stride;
      const f3 wp = make3(e[0], e[1], e[2]);
      // cacheLightingRSM.comp:103-104
      cpos[w][lane][0] = make_float4(wp.x, wp.y, wp.z, 0.0f);
      cpos[w][lane][1] = make_float4((wp.x - a.vmin[0]) / vsz, (wp.y - a.vmin[1]) / vsz, (wp.z - a.vmin[2]) / vsz, 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t next = 0u;  // wave-uniform: first entry not yet taken
    bool has = false;
    uint32_t cl = 0u, blk = 0u;
    Trace t{};  // an idle lane: finished, takes no samples
    t.done = true;
    t.s = 32;
    while (true) {
      // Idle lanes take the next entries (block-major: concurrent rays share
      // a block and stay coherent).
      const uint64_t idle = __ballot(!has);
      const uint32_t n_idle = static_cast<uint32_t>(__builtin_popcountll(idle));
      if (next < pool && n_idle != 0u && (n_idle >= a.refill_min || n_idle == 64u)) {
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(idle >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(idle), 0u));
        if (!has && next + rank < pool) {
          const uint32_t i = next + rank;
          const uint32_t bl = n_valid == 64u ? (i >> 6) : i / n_valid;
          cl = i - bl * n_valid;
          blk = b0 + bl;
          const float4 p0 = cpos[w][cl][0], p1 = cpos[w][cl][1];
          t = trace_setup(a.vox, a.voxel_size, make3(p0.x, p0.y, p0.z), make3(p1.x, p1.y, p1.z), a.blocks[blk]);
          has = true;
        }
        next += n_idle;
      }
      if (!__any(has)) break;
      const bool ended = has && trace_batch<kBatch>(a.vox, cr, lvi, t);
      if (ended) {
        a.vis[static_cast<size_t>(blk) * a.pitch + tile * 64u + cl] = trace_visibility(t);
        my_traces += 1u;
        my_exec += static_cast<uint32_t>(t.s);
        if (a.steps) my_steps += static_cast<uint32_t>(reference_steps(t));
        has = false;
        t.done = true;
        t.s = 32;
      }
    }
    __builtin_amdgcn_wave_barrier();  // cpos is restaged by the next item
  }
  if (a.steps) count_steps(a.steps, lane, my_traces, my_steps, my_exec);
}

// ---- SH gather ----------------------------------------------------------

__global__ void k_plan_gather(Counter* ctr, uint32_t n_vpl, uint32_t quantum, uint32_t max_splits,
                              uint32_t target_items) {
  plan_splits(ctr, n_vpl, quantum, max_splits, target_items);
}

template <bool kSh2, bool kShadow>
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ caches, int stride, const Counter* counter,
                                                const Vpl* __restrict__ vpls, uint32_t n_vpl,
                                                const float* __restrict__ vis, uint32_t vis_pitch,
                                                uint32_t interval, float* __restrict__ partial,
                                                uint32_t partial_pitch) {
  __shared__ float part[2][27][64];
  const uint32_t live = counter->live;
  const uint32_t splits = counter->splits;
  const uint32_t groups = (splits + 3u) / 4u;
  const uint32_t tiles = (live + 63u) / 64u;
  const uint32_t items = tiles * groups;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t tile = item / groups, group = item % groups;
    gather_item<kSh2, kShadow>(caches, stride, live, vpls, n_vpl, vis, vis_pitch, interval, partial, partial_pitch,
                               splits, tile, group, part);
    __syncthreads();  // part is rewritten by the next item
  }
}

// Sums the split partials, applies the SH evaluation factors
// (cacheLightingRSM.comp:267-277) and adds into the entry
// (cacheLightingRSM.comp:357-373).  A workgroup owns one (64-cache tile,
// coefficient, channel); its four waves sum interleaved quarters of the
// splits (lane = cache, so every load is a coalesced row) and the quarters
// are combined in a fixed order through LDS.
__global__ __launch_bounds__(256) void k_reduce(float* caches, int stride, const Counter* counter,
                                                const float* partial, uint32_t partial_pitch, drv_sh_constants shc,
                                                int sh2) {
  __shared__ float part[2][4][64];
  const uint32_t live = counter->live;
  const uint32_t groups = (counter->splits + 3u) / 4u;  // the gather wrote one partial per group of 4 splits
  const uint32_t n_jc = sh2 ? 27u : 12u;
  const uint32_t tiles = (live + 63u) / 64u;
  for (uint32_t item = blockIdx.x; item < tiles * n_jc; item += gridDim.x) {
    const uint32_t tile = item % tiles, jc = item / tiles;
    reduce_item(caches, stride, live, partial, partial_pitch, groups, tile, jc, shc, part);
    __syncthreads();
  }
}

}  // namespace

void launch_prepare_vpls(const drv_rsm& rsm, const drv_spot_light& light, Vpl* vpls, ShadowBlock* blocks,
                         int flux_cut, Vpl* vpls_lanes, const GatherPlan* plan, hipStream_t s, MarchStep* march,
                         const VoxelView* vox) {
  PrepArgs a;
  for (int i = 0; i < 16; ++i) a.inv_lvp[i] = light.inverse_light_view_projection[i];
  for (int i = 0; i < 3; ++i) a.light_pos[i] = light.light_position[i];
  a.val_area_factor = light.val_area_factor;
  a.sampling_offset = light.indirect_shadow_sampling_offset;
  a.super_val_width = light.indirect_shadow_super_val_width;
  a.flux = rsm.flux;
  a.normal = rsm.normal;
  a.dls = rsm.depth_lin_sq;
  a.dls_shadow = rsm.depth_lin_sq_shadow;
  a.R = rsm.read_resolution;
  a.R_shadow = rsm.read_resolution >> static_cast<int>(light.indirect_shadow_lod);
  a.H = rsm.read_height > 0 ? rsm.read_height : rsm.read_resolution;
  a.H_shadow = a.H >> static_cast<int>(light.indirect_shadow_lod);
  a.interval = static_cast<uint32_t>(light.indirect_shadow_sample_interval);
  a.block = 1 << static_cast<int>(light.indirect_shadow_lod);
  a.flux_cut = flux_cut;
  a.vpls = vpls;
  a.vpls_lanes = vpls_lanes;
  a.plan = plan ? plan->counter : nullptr;
  a.quantum = plan ? plan->quantum : 0u;
  a.max_splits = plan ? plan->max_splits : 0u;
  a.target_items = plan ? plan->target_items : 0u;
  a.blocks = blocks;
  a.march = blocks && vox ? march : nullptr;
  a.vox = vox ? *vox : VoxelView{};
  const uint32_t n = static_cast<uint32_t>(a.R) * static_cast<uint32_t>(a.H);
  const uint32_t march_groups = a.march ? (n / a.interval * kMarchSteps + 255u) / 256u : 0u;
  hipLaunchKernelGGL(k_prepare_vpls, dim3((n + 255) / 256 + march_groups), dim3(256), 0, s, a);
}

void launch_visibility(const drv_volume& vol, const VoxelView& vox, const float* caches, int stride,
                       const Counter* counter, const ShadowBlock* blocks, uint32_t n_blocks, float* visibility,
                       uint32_t vis_pitch, uint64_t* step_counter, const Tuning& tune, int grid,
                       hipStream_t s, const MarchStep* march) {
  VisArgs a;
  a.vox = vox;
  for (int i = 0; i < 3; ++i) a.vmin[i] = vol.volume_world_min[i];
  a.voxel_size = vol.voxel_size_in_world;
  a.caches = caches;
  a.stride = stride;
  a.counter = counter;
  a.blocks = blocks;
  a.n_blocks = n_blocks;
  a.vis = visibility;
  a.pitch = vis_pitch;
  a.steps = step_counter;
  a.blocks_per_item = static_cast<uint32_t>(tune.blocks_per_item);
  a.refill_min = static_cast<uint32_t>(tune.refill_min_lanes);
  a.march = march;
  const dim3 g(grid), b(256);
  if (march && tune.visibility_kernel == 0 && tune.steps_per_batch == 1) {
    // Record indices in FP32 while every level's corner copy has fewer than
    // 2^24 records ((res + 1)^3 < 2^24, res <= 128).
    const uint64_t r1 = static_cast<uint64_t>(vox.res) + 1u;
    const bool fidx = r1 * r1 * r1 < (1ull << 24);
    if (step_counter) {
      if (fidx) {
        hipLaunchKernelGGL((k_visibility_planned<true, true>), g, b, 0, s, a);
      } else {
        hipLaunchKernelGGL((k_visibility_planned<true, false>), g, b, 0, s, a);
      }
    } else if (fidx) {
      hipLaunchKernelGGL((k_visibility_planned<false, true>), g, b, 0, s, a);
    } else {
      hipLaunchKernelGGL((k_visibility_planned<false, false>), g, b, 0, s, a);
    }
  } else if (tune.visibility_kernel == 1) {
    switch (tune.steps_per_batch) {
      case 1: hipLaunchKernelGGL(k_visibility_pool<1>, g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL(k_visibility_pool<3>, g, b, 0, s, a); break;
      case 4: hipLaunchKernelGGL(k_visibility_pool<4>, g, b, 0, s, a); break;
      case 8: hipLaunchKernelGGL(k_visibility_pool<8>, g, b, 0, s, a); break;
      default: hipLaunchKernelGGL(k_visibility_pool<2>, g, b, 0, s, a); break;
    }
  } else {
    switch (tune.steps_per_batch) {
      case 1: hipLaunchKernelGGL(k_visibility<1>, g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL(k_visibility<3>, g, b, 0, s, a); break;
      case 4: hipLaunchKernelGGL(k_visibility<4>, g, b, 0, s, a); break;
      case 8: hipLaunchKernelGGL(k_visibility<8>, g, b, 0, s, a); break;
      default: hipLaunchKernelGGL(k_visibility<2>, g, b, 0, s, a); break;
    }
  }
}

void launch_gather(const drv_sh_constants& shc, int sh_bands, const float* caches, int stride,
                   const Counter* counter, const Vpl* vpls, uint32_t n_vpl, const float* visibility,
                   uint32_t vis_pitch, uint32_t interval, int indirect_shadow, float* partial,
                   uint32_t partial_pitch, uint32_t max_splits, int grid, bool plan, hipStream_t s) {
  (void)shc;
  if (plan) {
    const GatherPlan p = gather_plan(const_cast<Counter*>(counter), interval, max_splits, grid);
    hipLaunchKernelGGL(k_plan_gather, dim3(1), dim3(1), 0, s, p.counter, n_vpl, p.quantum, p.max_splits,
                       p.target_items);
  }
  const bool sh2 = sh_bands >= 2;
  if (sh2 && indirect_shadow)
    hipLaunchKernelGGL((k_gather<true, true>), dim3(grid), dim3(256), 0, s, caches, stride, counter, vpls, n_vpl,
                       visibility, vis_pitch, interval, partial, partial_pitch);
  else if (sh2)
    hipLaunchKernelGGL((k_gather<true, false>), dim3(grid), dim3(256), 0, s, caches, stride, counter, vpls, n_vpl,
                       visibility, vis_pitch, interval, partial, partial_pitch);
  else if (indirect_shadow)
    hipLaunchKernelGGL((k_gather<false, true>), dim3(grid), dim3(256), 0, s, caches, stride, counter, vpls, n_vpl,
                       visibility, vis_pitch, interval, partial, partial_pitch);
  else
    hipLaunchKernelGGL((k_gather<false, false>), dim3(grid), dim3(256), 0, s, caches, stride, counter, vpls, n_vpl,
                       visibility, vis_pitch, interval, partial, partial_pitch);
}

void launch_reduce_sh(const drv_sh_constants& shc, int sh_bands, float* caches, int stride, const Counter* counter,
                      const float* partial, uint32_t partial_pitch, int grid, hipStream_t s) {
  // grid: persistent workgroups over (tile, coefficient, channel) items
  hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(256), 0, s, caches, stride, counter, partial, partial_pitch, shc,
                     sh_bands >= 2 ? 1 : 0);
}

}  // namespace drvk
