// Batched light pass: the caches of many views (each with its own cache
// list, VPLs, shadow blocks and voxel volume) lit by ONE persistent
// work-queue kernel, cone tracing fused into the VPL sum.
//
// Restates the same per-cache computation as the single-view path
// (shader/cacheLightingRSM.comp:80-374: cone-traced shadowing per shadow
// block, then the SH projection of every VPL of the block), organised for a
// batch of uneven views on MI355X:
//
// * k_batch_plan (one thread) turns each view's live cache count into work
//   items (view, tile of `cpw` caches, VPL split of `vpl_tile` VPLs) and lays
//   the views out back to back in one queue;
// * k_batch_gather: persistent waves pop items from an atomic queue head, so
//   views of any size share the GPU without a launch per view.  Each lane
//   owns one cache of the tile; with cpw < 64 the 64 / cpw lane groups of a
//   wave take interleaved shadow blocks of the item and their sums are
//   combined at the end.  Each (cache, block) pair is traced exactly once;
//   with one split per view the sums go straight into the entries, otherwise
//   into per-item partials;
// * k_batch_reduce sums the split partials of the views that have them.
//
// The split queue (caches per wave DRV_BATCH_SPLIT_QUEUE) runs the
// single-view kernels' work items of every view in one launch per stage
// instead: VPL records of every view, a one-thread plan (split counts and
// item offsets from the live counts), then the cone-traced visibility, the SH
// gather and the reduction, each persistent over the concatenated items of
// all views.  Small views no longer leave the GPU idle behind their own
// launches, and no view waits for another's tail.
#include <algorithm>
#include <cstring>

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_light.hpp"
#include "k_sh.hpp"
#include "k_trace.hpp"
#include "split_plan.hpp"

namespace drvk {
namespace {

__global__ void k_batch_plan(BatchView* views, uint32_t n_views, uint32_t cpw, uint32_t vpl_tile,
                             uint64_t partial_floats, BatchQueue* q) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t max_vpl = 1u;
  for (uint32_t v = 0; v < n_views; ++v) max_vpl = max(max_vpl, views[v].n_vpl);
  // The VPL tile doubles until the split partials fit the buffer (one split
  // per view always fits: it needs none).
  uint32_t tile = vpl_tile;
  while (true) {
    uint32_t offset = 0;
    uint64_t slots = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
      BatchView& b = views[v];
      const uint32_t live = b.counter->live;
      b.tiles = (live + cpw - 1u) / cpw;
      b.splits = b.n_vpl / min(tile, b.n_vpl);
      b.item_begin = offset;
      b.partial_begin = static_cast<uint32_t>(slots);
      offset += b.tiles * b.splits;
      if (b.splits > 1u) slots += static_cast<uint64_t>(b.tiles) * b.splits;
    }
    if (slots * 27u * cpw <= partial_floats || tile >= max_vpl) {
      q->total = offset;
      break;
    }
    tile *= 2u;
  }
  q->head = 0u;
  q->vpl_tile = tile;
}

// View of a queue item (views are laid out back to back; n_views is small).
__device__ __forceinline__ uint32_t find_view(const BatchView* views, uint32_t n_views, uint32_t item) {
  uint32_t lo = 0, hi = n_views;  // largest v with item_begin <= item
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].item_begin <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct BatchArgs {
  BatchView* views;
  uint32_t n_views;
  BatchQueue* q;
  int stride;
  uint32_t cpw;
  float* partial;  // [partial slot][27][cpw]
  drv_sh_constants shc;
};

template <bool kSh2, bool kShadow>
__global__ __launch_bounds__(256) void k_batch_gather(BatchArgs a) {
  __shared__ LevelInfo lvs[4][16];  // per wave: voxel level table of the current item's view
  const uint32_t lane = threadIdx.x & 63u;
  LevelInfo* lvi = lvs[threadIdx.x >> 6];
  uint32_t staged = 0xffffffffu;  // view whose table is in `lvi`
  const uint32_t cpw = a.cpw;
  const uint32_t nsub = 64u / cpw;
  const uint32_t sub = lane / cpw, cl = lane % cpw;
  const uint32_t total = a.q->total;
  while (true) {
    uint32_t item = 0u;
    if (lane == 0u) item = atomicAdd(&a.q->head, 1u);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
    if (item >= total) break;
    const uint32_t v = find_view(a.views, a.n_views, item);
    const BatchView& bv = a.views[v];
    if (v != staged) {  // wave-local table: store, then read after a wavefront fence
      __builtin_amdgcn_wave_barrier();
      if (lane < 16u) lvi[lane] = level_info(bv.vox, static_cast<int>(lane));
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      staged = v;
    }
    const uint32_t local = item - bv.item_begin;
    const uint32_t tile = local / bv.splits, split = local % bv.splits;
    const uint32_t live = bv.counter->live;
    const uint32_t cache = tile * cpw + cl;
    const bool valid = cache < live;
    const float* e = bv.caches + static_cast<size_t>(valid ? cache : 0u) * a.stride;
    const f3 wp = make3(e[0], e[1], e[2]);
    const float fres = static_cast<float>(bv.vox.res);
    const float vsz = bv.voxel_size * fres;
    // cacheLightingRSM.comp:103-104
    const f3 vpos = make3((wp.x - bv.vmin[0]) / vsz, (wp.y - bv.vmin[1]) / vsz, (wp.z - bv.vmin[2]) / vsz);
    const uint32_t chunk = bv.n_vpl / bv.splits;
    const uint32_t interval = bv.interval;
    const uint32_t b_first = split * chunk / interval, n_blk = chunk / interval;
    ShSums acc;
    sh_zero(acc);
    for (uint32_t bi = sub; bi < n_blk; bi += nsub) {
      const uint32_t b = b_first + bi;
      float shadowing = 1.0f;
      if (kShadow) {
        // cacheLightingRSM.comp:167-231, the same trace as k_visibility.
        Trace t = trace_setup(bv.vox, bv.voxel_size, wp, vpos, bv.blocks[b]);
        if (!valid) {
          t.done = true;
          t.s = 32;
        }
        bool finished = !valid;
        const __amdgpu_buffer_rsrc_t cr = corner_rsrc(bv.vox);
        while (!finished) finished = trace_batch(bv.vox, cr, lvi, t);
        shadowing = trace_visibility(t);
      }
      const uint32_t m0 = b * interval;
#pragma unroll 4
      for (uint32_t m = m0; m < m0 + interval; ++m) sh_accumulate<kSh2>(acc, bv.vpls[m], wp.x, wp.y, wp.z, shadowing);
    }
    // Lane groups of the same cache: fixed-order combine.
    for (uint32_t off = cpw; off < 64u; off <<= 1) {
#pragma unroll
      for (int k = 0; k < 27; ++k) acc.s[k] += __shfl_xor(acc.s[k], static_cast<int>(off));
    }
    if (!valid || sub != 0u) continue;
    const int n_out = kSh2 ? 27 : 12;
    if (bv.splits == 1u) {
      float* entry = bv.caches + static_cast<size_t>(cache) * a.stride;
#pragma unroll
      for (int k = 0; k < 27; ++k)
        if (k < n_out) sh_add_coefficient(entry, static_cast<uint32_t>(k) / 3u, static_cast<uint32_t>(k) % 3u, acc.s[k],
                                          acc.s[k % 3], a.shc);
    } else {
      float* p = a.partial + (static_cast<size_t>(bv.partial_begin) + local) * 27u * cpw + cl;
#pragma unroll
      for (int k = 0; k < 27; ++k)
        if (k < n_out) p[static_cast<size_t>(k) * cpw] = acc.s[k];
    }
  }
}

// Split partials of view blockIdx.y, summed in split order into its entries.
__global__ __launch_bounds__(256) void k_batch_reduce(BatchArgs a, int sh2) {
  const BatchView& bv = a.views[blockIdx.y];
  if (bv.splits <= 1u) return;
  const uint32_t live = bv.counter->live;
  const uint32_t n_coef = sh2 ? 9u : 4u;
  const uint32_t total = live * n_coef * 3u;
  const uint32_t cpw = a.cpw;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const uint32_t c = t % live;
    const uint32_t jc = t / live;
    const uint32_t j = jc / 3u, ch = jc % 3u;
    const uint32_t tile = c / cpw, cl = c % cpw;
    const size_t first = static_cast<size_t>(bv.partial_begin) + static_cast<size_t>(tile) * bv.splits;
    float v = 0.0f, v0 = 0.0f;
    for (uint32_t sp = 0; sp < bv.splits; ++sp) {
      const float* p = a.partial + (first + sp) * 27u * cpw + cl;
      v += p[static_cast<size_t>(jc) * cpw];
      if (j == 6u) v0 += p[static_cast<size_t>(ch) * cpw];
    }
    sh_add_coefficient(bv.caches + static_cast<size_t>(c) * a.stride, j, ch, v, v0, a.shc);
  }
}

// ---- split queue ---------------------------------------------------------------

struct SplitDev {
  PrepArgs prep;  // VPL and shadow-block records of the view's light
  float* caches;
  const Counter* counter;
  float* vis;
  float* partial;
  VoxelView vox;
  float vmin[3];
  float voxel_size;
  uint32_t vis_pitch, partial_pitch, max_splits;
  uint32_t n_vpl, interval, n_blocks;
  // written by the plan
  uint32_t live, tiles, splits;
  uint32_t vis_begin, gather_begin, reduce_begin;
};

// Workgroups [0, vpl_groups) of row y prepare view y's VPL and block
// records, the ones after them its march records (one per thread).
__global__ __launch_bounds__(256) void k_split_prepare(const SplitDev* views, uint32_t vpl_groups) {
  const PrepArgs& a = views[blockIdx.y].prep;
  const uint32_t n = static_cast<uint32_t>(a.R) * static_cast<uint32_t>(a.H);
  if (blockIdx.x >= vpl_groups) {
    const uint32_t idx = (blockIdx.x - vpl_groups) * blockDim.x + threadIdx.x;
    if (a.march && idx < n / a.interval * kMarchSteps) prepare_march(a, idx);
    return;
  }
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  prepare_vpl(a, m, n);
}

// Split counts and item offsets of every view: the smallest power of two s
// of splits with sum(tiles) x s >= target_items (a view takes fewer when its
// VPLs or its partial buffer allow fewer), so the gather fills the grid over
// all views together instead of per view.  One workgroup: the views' live
// counts are loaded in parallel (a chain of dependent loads per view in one
// thread took ~1 us per view), then one thread walks the LDS copy.
constexpr uint32_t kPlanChunk = 1024;
// Shadow blocks of one visibility item of the split queue (a lane traces its
// cache towards them one after the other).
constexpr uint32_t kSplitBpi = 4;
__global__ __launch_bounds__(256) void k_split_plan(SplitDev* views, uint32_t n_views, uint32_t target_items,
                                                    int shadow, int sh2, SplitTotals* totals) {
  __shared__ uint32_t live_s[kPlanChunk];
  __shared__ uint32_t all_tiles_s;
  __shared__ uint32_t s_s;
  __shared__ uint32_t run[3];  // running vis, gather, reduce offsets
  if (threadIdx.x == 0) {
    all_tiles_s = 0u;
    run[0] = run[1] = run[2] = 0u;
  }
  __syncthreads();
  // pass 1: tiles of every view, summed
  for (uint32_t base = 0; base < n_views; base += kPlanChunk) {
    const uint32_t m = min(kPlanChunk, n_views - base);
    uint32_t sum = 0;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      SplitDev& d = views[base + i];
      d.live = d.counter->live;
      d.tiles = (d.live + 63u) / 64u;
      sum += d.tiles;
    }
    atomicAdd(&all_tiles_s, sum);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 1u;
    while (s < kMaxSplits && static_cast<uint64_t>(all_tiles_s) * s < target_items) s *= 2u;
    s_s = s;
  }
  __syncthreads();
  const uint32_t n_jc = sh2 ? 27u : 12u;
  // pass 2: per-view splits in parallel, offsets by one thread over LDS
  for (uint32_t base = 0; base < n_views; base += kPlanChunk) {
    const uint32_t m = min(kPlanChunk, n_views - base);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      SplitDev& d = views[base + i];
      const uint32_t quantum = d.interval > 64u ? d.interval : 64u;
      d.splits = plan_split_count(d.live, d.n_vpl, quantum, min(s_s, d.max_splits), 0xffffffffu);
      live_s[i] = (d.tiles << 9) | ((d.splits + 3u) / 4u);  // tiles < 2^23, groups <= 64
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t i = 0; i < m; ++i) {
        SplitDev& d = views[base + i];
        const uint32_t tiles = live_s[i] >> 9, groups = live_s[i] & 511u;
        d.vis_begin = run[0];
        d.gather_begin = run[1];
        d.reduce_begin = run[2];
        if (shadow) run[0] += tiles * ((d.n_blocks + kSplitBpi - 1u) / kSplitBpi);
        run[1] += tiles * groups;
        run[2] += tiles * n_jc;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals->vis = run[0];
    totals->gather = run[1];
    totals->reduce = run[2];
    totals->vis_head = 0u;
  }
}

// The view an item of a stage belongs to: the largest v with begin(v) <= item
// (views in order, offsets non-decreasing; a view without items of the stage
// shares its offset with the next and is never returned for a valid item).
template <int kStage>
__device__ __forceinline__ uint32_t split_view(const SplitDev* views, uint32_t n_views, uint32_t item) {
  uint32_t lo = 0, hi = n_views;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t b = kStage == 0 ? views[mid].vis_begin : kStage == 1 ? views[mid].gather_begin
                                                                          : views[mid].reduce_begin;
    if (b <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// What the visibility stage reads of a view, loaded when a wave's items
// move into it (wave-uniform scalar loads).
struct VisView {
  uint32_t begin, end, tiles, live, pitch, n_blocks;
  bool fidx;  // record indices in FP32 (every level below 2^24 records)
  const float* caches;
  float* vis;
  const ShadowBlock* blocks;
  const MarchStep* march;
  VoxelView vox;
  float vmin[3];
  float voxel_size;
};

__device__ __forceinline__ VisView vis_view(const SplitDev* views, uint32_t v) {
  typedef const __attribute__((address_space(4))) SplitDev* Desc;
  const Desc d = (Desc)(views + v);  // NOLINT(old-style-cast)
  VisView w;
  w.begin = d->vis_begin;
  w.tiles = d->tiles;
  w.n_blocks = d->n_blocks;
  w.end = w.begin + w.tiles * ((w.n_blocks + kSplitBpi - 1u) / kSplitBpi);
  w.live = d->live;
  w.pitch = d->vis_pitch;
  w.caches = d->caches;
  w.vis = d->vis;
  w.blocks = d->prep.blocks;
  w.march = d->prep.march;
  w.vox.corners = d->vox.corners;
  w.vox.corner_bytes = d->vox.corner_bytes;
  w.vox.res = d->vox.res;
  w.vox.levels = d->vox.levels;
  for (int i = 0; i < 3; ++i) w.vmin[i] = d->vmin[i];
  w.voxel_size = d->voxel_size;
  const uint64_t r1 = static_cast<uint64_t>(w.vox.res) + 1u;
  w.fidx = r1 * r1 * r1 < (1ull << 24);
  return w;
}

// Items a wave takes from the visibility queue at a time: 1 or 2 contend
// for the queue head, 8 and more put the items of more views in flight at
// once (each view traces its own voxel copy).
constexpr uint32_t kVisChunk = 4;

// Cone-traced visibility of every view: items (view, group of kSplitBpi
// shadow blocks, 64-cache tile), group-major within a view, handed out in
// chunks of kVisChunk consecutive items by an atomic queue head, so a wave's
// items mostly share a view and a block group (its descriptor, block records
// and march records stay in SGPRs and the scalar cache) and the waves that
// finish first take more.  Each lane loads its cache once per item and traces
// it towards the group's blocks with the code of k_visibility_planned (the
// view's march plan), so the values are those of the single-view kernels.
__global__ __launch_bounds__(256) void k_split_visibility(const SplitDev* views, uint32_t n_views, int stride,
                                                          SplitTotals* totals) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t items = totals->vis;
  VisView w;
  w.begin = w.end = 0u;
  while (true) {
    uint32_t base = 0u;
    if (lane == 0u) base = atomicAdd(&totals->vis_head, kVisChunk);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
    if (base >= items) break;
    const uint32_t stop = min(items, base + kVisChunk);
    for (uint32_t item = base; item < stop; ++item) {
      if (item < w.begin || item >= w.end) w = vis_view(views, split_view<0>(views, n_views, item));
      const uint32_t local = item - w.begin;
      const uint32_t group = local / w.tiles, tile = local - group * w.tiles;
      const uint32_t cache = tile * 64u + lane;
      const bool valid = cache < w.live;
      const float* e = w.caches + static_cast<size_t>(valid ? cache : w.live - 1u) * stride;
      const f3 wp = make3(e[0], e[1], e[2]);
      const float vsz = w.voxel_size * static_cast<float>(w.vox.res);
      // cacheLightingRSM.comp:103-104
      const f3 vpos = make3((wp.x - w.vmin[0]) / vsz, (wp.y - w.vmin[1]) / vsz, (wp.z - w.vmin[2]) / vsz);
      const __amdgpu_buffer_rsrc_t cr = corner_rsrc(w.vox);
      const uint32_t b0 = group * kSplitBpi, b1 = min(w.n_blocks, b0 + kSplitBpi);
      for (uint32_t b = b0; b < b1; ++b) {
        Trace t = trace_setup(w.vox, w.voxel_size, wp, vpos, block_record(w.blocks, b));
        if (w.fidx) trace_planned<false, true>(march_steps(w.march, b), cr, t, valid);
        else trace_planned<false, false>(march_steps(w.march, b), cr, t, valid);
        if (valid) w.vis[static_cast<size_t>(b) * w.pitch + cache] = trace_visibility(t);
      }
    }
  }
}

// SH gather of every view: workgroup items (view, 64-cache tile, group of 4
// VPL splits), the single-view gather's item body.
template <bool kSh2, bool kShadow>
__global__ __launch_bounds__(256) void k_split_gather(const SplitDev* views, uint32_t n_views, int stride,
                                                      const SplitTotals* totals) {
  __shared__ float part[2][27][64];
  const uint32_t items = totals->gather;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t v = split_view<1>(views, n_views, item);
    const SplitDev& d = views[v];
    const uint32_t groups = (d.splits + 3u) / 4u;
    const uint32_t local = item - d.gather_begin;
    const uint32_t tile = local / groups, group = local - tile * groups;
    gather_item<kSh2, kShadow>(d.caches, stride, d.live, d.prep.vpls, d.n_vpl, d.vis, d.vis_pitch, d.interval,
                               d.partial, d.partial_pitch, d.splits, tile, group, part);
    __syncthreads();  // part is rewritten by the next item
  }
}

// Split reduction of every view: workgroup items (view, coefficient x
// channel, 64-cache tile), the single-view reduction's item body.
__global__ __launch_bounds__(256) void k_split_reduce(const SplitDev* views, uint32_t n_views, int stride,
                                                      const SplitTotals* totals, drv_sh_constants shc, int sh2) {
  __shared__ float part[2][4][64];
  const uint32_t items = totals->reduce;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t v = split_view<2>(views, n_views, item);
    const SplitDev& d = views[v];
    const uint32_t local = item - d.reduce_begin;
    const uint32_t jc = local / d.tiles, tile = local - jc * d.tiles;
    reduce_item(d.caches, stride, d.live, d.partial, d.partial_pitch, (d.splits + 3u) / 4u, tile, jc, shc, part);
    __syncthreads();
  }
}

}  // namespace

size_t split_desc_bytes() { return sizeof(SplitDev); }

void build_split_descs(const SplitViewHost* views, uint32_t n_views, void* host_out) {
  SplitDev* out = static_cast<SplitDev*>(host_out);
  for (uint32_t v = 0; v < n_views; ++v) {
    const SplitViewHost& h = views[v];
    SplitDev d;
    std::memset(&d, 0, sizeof(d));
    PrepArgs& a = d.prep;
    for (int i = 0; i < 16; ++i) a.inv_lvp[i] = h.light.inverse_light_view_projection[i];
    for (int i = 0; i < 3; ++i) a.light_pos[i] = h.light.light_position[i];
    a.val_area_factor = h.light.val_area_factor;
    a.sampling_offset = h.light.indirect_shadow_sampling_offset;
    a.super_val_width = h.light.indirect_shadow_super_val_width;
    a.flux = h.rsm.flux;
    a.normal = h.rsm.normal;
    a.dls = h.rsm.depth_lin_sq;
    a.dls_shadow = h.rsm.depth_lin_sq_shadow;
    a.R = h.rsm.read_resolution;
    a.R_shadow = h.rsm.read_resolution >> static_cast<int>(h.light.indirect_shadow_lod);
    a.H = h.rsm.read_height > 0 ? h.rsm.read_height : h.rsm.read_resolution;
    a.H_shadow = a.H >> static_cast<int>(h.light.indirect_shadow_lod);
    a.interval = static_cast<uint32_t>(h.light.indirect_shadow_sample_interval);
    a.block = 1 << static_cast<int>(h.light.indirect_shadow_lod);
    a.flux_cut = 0;
    a.vpls = h.vpls;
    a.vpls_lanes = nullptr;
    a.plan = nullptr;
    a.blocks = h.blocks;
    a.march = h.blocks ? h.march : nullptr;
    a.vox = h.vox;
    d.caches = h.caches;
    d.counter = h.counter;
    d.vis = h.vis;
    d.partial = h.partial;
    d.vox = h.vox;
    for (int i = 0; i < 3; ++i) d.vmin[i] = h.vmin[i];
    d.voxel_size = h.voxel_size;
    d.vis_pitch = h.vis_pitch;
    d.partial_pitch = h.partial_pitch;
    d.max_splits = h.max_splits;
    d.n_vpl = h.n_vpl;
    d.interval = h.interval;
    d.n_blocks = h.n_blocks;
    out[v] = d;
  }
}

void launch_batch_split_prepare(void* descs, uint32_t n_views, SplitTotals* totals, int sh_bands, int shadow,
                                uint32_t max_vpl, uint32_t max_march, uint32_t target_items, hipStream_t s) {
  SplitDev* d = static_cast<SplitDev*>(descs);
  const uint32_t vpl_groups = (max_vpl + 255u) / 256u;
  hipLaunchKernelGGL(k_split_prepare, dim3(vpl_groups + (max_march + 255u) / 256u, n_views), dim3(256), 0, s, d,
                     vpl_groups);
  hipLaunchKernelGGL(k_split_plan, dim3(1), dim3(256), 0, s, d, n_views, target_items, shadow, sh_bands >= 2 ? 1 : 0,
                     totals);
}

void launch_batch_split_visibility(void* descs, uint32_t n_views, SplitTotals* totals, int stride, int grid,
                                   hipStream_t s) {
  hipLaunchKernelGGL(k_split_visibility, dim3(grid), dim3(256), 0, s, static_cast<SplitDev*>(descs), n_views, stride,
                     totals);
}

void launch_batch_split_gather(void* descs, uint32_t n_views, SplitTotals* totals, int stride, int sh_bands,
                               int shadow, const drv_sh_constants& shc, int grid, hipStream_t s) {
  SplitDev* d = static_cast<SplitDev*>(descs);
  const bool sh2 = sh_bands >= 2;
  if (sh2 && shadow) hipLaunchKernelGGL((k_split_gather<true, true>), dim3(grid), dim3(256), 0, s, d, n_views, stride, totals);
  else if (sh2) hipLaunchKernelGGL((k_split_gather<true, false>), dim3(grid), dim3(256), 0, s, d, n_views, stride, totals);
  else if (shadow) hipLaunchKernelGGL((k_split_gather<false, true>), dim3(grid), dim3(256), 0, s, d, n_views, stride, totals);
  else hipLaunchKernelGGL((k_split_gather<false, false>), dim3(grid), dim3(256), 0, s, d, n_views, stride, totals);
  hipLaunchKernelGGL(k_split_reduce, dim3(grid), dim3(256), 0, s, d, n_views, stride, totals, shc, sh2 ? 1 : 0);
}

void launch_batch(BatchView* views, uint32_t n_views, BatchQueue* q, int stride, int sh_bands, int shadow,
                  uint32_t cpw, uint32_t vpl_tile, float* partial, size_t partial_floats,
                  const drv_sh_constants& shc, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, s, views, n_views, cpw, vpl_tile,
                     static_cast<uint64_t>(partial_floats), q);
  BatchArgs a;
  a.views = views;
  a.n_views = n_views;
  a.q = q;
  a.stride = stride;
  a.cpw = cpw;
  a.partial = partial;
  a.shc = shc;
  const bool sh2 = sh_bands >= 2;
  if (sh2 && shadow) hipLaunchKernelGGL((k_batch_gather<true, true>), dim3(grid), dim3(256), 0, s, a);
  else if (sh2) hipLaunchKernelGGL((k_batch_gather<true, false>), dim3(grid), dim3(256), 0, s, a);
  else if (shadow) hipLaunchKernelGGL((k_batch_gather<false, true>), dim3(grid), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_batch_gather<false, false>), dim3(grid), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_batch_reduce, dim3(64, n_views), dim3(256), 0, s, a, sh2 ? 1 : 0);
}

}  // namespace drvk
