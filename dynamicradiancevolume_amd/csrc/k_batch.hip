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
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_sh.hpp"
#include "k_trace.hpp"

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
        bool finished = false;
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

}  // namespace

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
