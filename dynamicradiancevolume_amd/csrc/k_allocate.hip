// Stage A: G-buffer-driven cache allocation into the clipmap address volume.
//
// Restates shader/cacheGather.comp:93-164 (per-pixel cell, 16x16 tile
// dedupe with the leader rule, CAS-locked claim of the 8 corner cells) and
// shader/cachePrepareLighting.comp:8-14, driven as in
// Renderer::AllocateCaches (rendering/renderer.cpp:951-992).
//
// MI355X design: one 256-thread workgroup (4 wave64) per 16x16 pixel tile,
// the two tile lists in LDS, and a wave-aggregated append: per corner the
// lanes that won their CAS are counted with one ballot, one lane reserves the
// whole run of cache indices with a single atomicAdd, and each lane takes its
// slot by a prefix popcount.  The set of claimed cells and every cache
// position is independent of the append order, which is the only thing the
// reference leaves nondeterministic.
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Two buffers cleared by one launch (the allocation's counter and CAV).
__global__ void k_fill_u32_pair(uint32_t* p, size_t n, uint32_t* q, size_t m, uint32_t v) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t j = i; j < n; j += stride) p[j] = v;
  for (size_t j = i; j < m; j += stride) q[j] = v;
}

__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  // 16 bytes per thread where aligned, bytes at the tail.
  const size_t n16 = (reinterpret_cast<uintptr_t>(p) & 15u) ? 0 : n / 16;
  const uint32_t w = 0x01010101u * v;
  uint4 q = make_uint4(w, w, w, w);
  for (size_t j = i; j < n16; j += stride) reinterpret_cast<uint4*>(p)[j] = q;
  for (size_t j = n16 * 16 + i; j < n; j += stride) p[j] = v;
}

struct AllocArgs {
  drv_volume vol;
  float inv_vp[16];
  const float* depth;
  uint32_t* cav;
  float* caches;
  Counter* counter;
  int width, height, R, C, stride;
  uint32_t capacity;
  int transitions;
};

// GetCache1DCoord, cacheGather.comp:20-30.
__device__ __forceinline__ int cache_1d_coord(const drv_volume& vol, int R, f3 wp, int c) {
  const drv_cascade& k = vol.cascades[c];
  int gx = to_int((wp.x - k.min[0]) / k.world_voxel_size);
  int gy = to_int((wp.y - k.min[1]) / k.world_voxel_size);
  int gz = to_int((wp.z - k.min[2]) / k.world_voxel_size);
  gx = min(R - 1, max(0, gx));
  gy = min(R - 1, max(0, gy));
  gz = min(R - 1, max(0, gz));
  return gx + gy * R + gz * R * R + c * R * R * R;
}

// AllocateCaches, cacheGather.comp:32-91, called by every lane of the wave
// (inactive lanes pass active = false) so the appends can be aggregated.
// A lane issues its 8 corner CASes back to back and waits once; then one
// ballot per corner numbers the winners (corner-major, lane-minor) and a
// single atomicAdd per wave reserves the whole run of cache indices.
__device__ void allocate_corners(const AllocArgs& a, bool active, int coord, int c) {
  const int R = a.R, R2 = R * R, R3 = R2 * R;
  int bx = 0, by = 0, bz = 0;
  if (active) {
    coord -= R3 * c;
    bz = coord / R2;
    by = (coord - bz * R2) / R;
    bx = coord % R;
  }
  uint32_t cell[8];
  uint32_t old[8];
  uint32_t inside_mask = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    // offsets (0,0,0),(0,1,0),(0,0,1),(0,1,1),(1,0,0),(1,1,0),(1,0,1),(1,1,1)
    const int lx = bx + (i >> 2), ly = by + (i & 1), lz = bz + ((i >> 1) & 1);
    const int gx = lx + c * R;
    const bool inside = active && gx >= 0 && gx < a.C * R && ly >= 0 && ly < R && lz >= 0 && lz < R;
    cell[i] = inside ? static_cast<uint32_t>((lz * R + ly) * (a.C * R) + gx) : 0u;
    old[i] = 1u;
    if (inside) {
      inside_mask |= 1u << i;
      old[i] = atomicCAS(&a.cav[cell[i]], 0u, 0xFFFFFFFFu);
    }
  }
  // A lock on a cell outside the image "succeeds" (image atomics outside the
  // image return 0) and appends a cache that no cell points to.
  uint32_t won = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool inside = (inside_mask >> i) & 1u;
    if (active && (!inside || old[i] == 0u)) won |= 1u << i;
  }
  const unsigned lane = __lane_id();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t masks[8];
  uint32_t total = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    masks[i] = __ballot((won >> i) & 1u);
    total += static_cast<uint32_t>(__popcll(masks[i]));
  }
  if (total == 0) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&a.counter->total, total);
  base = __shfl(base, 0);
  const drv_cascade& k = a.vol.cascades[c < 0 ? 0 : c];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if ((won >> i) & 1u) {
      const uint32_t idx = base + static_cast<uint32_t>(__popcll(masks[i] & lt_mask));
      if (idx < a.capacity) {
        const int lx = bx + (i >> 2), ly = by + (i & 1), lz = bz + ((i >> 1) & 1);
        const float px = static_cast<float>(lx) * k.world_voxel_size + k.min[0];
        const float py = static_cast<float>(ly) * k.world_voxel_size + k.min[1];
        const float pz = static_cast<float>(lz) * k.world_voxel_size + k.min[2];
        float4* e = reinterpret_cast<float4*>(a.caches + static_cast<size_t>(idx) * a.stride);
        e[0] = make_float4(px, py, pz, 0.0f);
        for (int j = 1; j < a.stride / 4; ++j) e[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      if ((inside_mask >> i) & 1u) a.cav[cell[i]] = idx + 1u;  // the cell stays locked for others
    }
    base += static_cast<uint32_t>(__popcll(masks[i]));
  }
}

__global__ __launch_bounds__(256) void k_allocate(AllocArgs a) {
  __shared__ int list[256];
  __shared__ int list2[256];
  const int t = threadIdx.x;
  const int lx = t & 15, ly = t >> 4;
  const int x = blockIdx.x * 16 + lx, y = blockIdx.y * 16 + ly;
  int own = -1, own2 = -1, c = -1;
  if (x < a.width && y < a.height) {
    const float d = a.depth[static_cast<size_t>(y) * a.width + x];
    if (d > 0.0001f) {
      const f3 wp = unproject_pixel(a.inv_vp, a.width, a.height, x, y, d);
      c = cascade_of(a.vol, a.C, wp);
      own = cache_1d_coord(a.vol, a.R, wp, c);
      if (a.transitions) {
        const float tr = cascade_transition(a.vol, c, wp);
        if (tr > 0.0f && c < a.C - 1) own2 = cache_1d_coord(a.vol, a.R, wp, c + 1);
      }
    }
  }
  list[t] = own;
  list2[t] = own2;
  __syncthreads();
  // Leader rule, cacheGather.comp:142-150: no left / up / up-left neighbour
  // of the tile holds the same cell, or this is thread (0,0).  Threads on the
  // left column and top row compare against themselves and never lead.
  const int ux = max(0, lx - 1), uy = max(0, ly - 1);
  const bool lead = (((list[uy * 16 + lx] != own) && (list[ly * 16 + ux] != own) && (list[uy * 16 + ux] != own)) ||
                     (ux == lx && uy == ly)) &&
                    own != -1;
  // Transition list, cacheGather.comp:154-162: mirrored towards right/down;
  // min(15, l + 1) == (15, 15) holds for the four threads with lx, ly >= 14.
  const int vx = min(15, lx + 1), vy = min(15, ly + 1);
  const bool lead2 = a.transitions &&
                     (((list2[vy * 16 + lx] != own2) && (list2[ly * 16 + vx] != own2) &&
                       (list2[vy * 16 + vx] != own2)) ||
                      (vx == 15 && vy == 15)) &&
                     own2 != -1;
  allocate_corners(a, lead, own, c);
  if (a.transitions) allocate_corners(a, lead2, own2, c + 1);
}

// cachePrepareLighting.comp:8-14 plus the live count and cache tiling used by
// the gather (capacity clamp = robust buffer access).
__global__ void k_prepare_dispatch(Counter* ctr, uint32_t capacity) {
  const uint32_t total = ctr->total;
  ctr->groups_x = (total + 63u) / 64u;
  ctr->groups_y = 1u;
  ctr->groups_z = 1u;
  const uint32_t live = min(total, capacity);
  ctr->live = live;
  ctr->tiles = (live + kGatherTile - 1) / kGatherTile;
  ctr->splits = 1u;
}

}  // namespace

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n, v);
}

void launch_fill_u32_pair(uint32_t* p, size_t n, uint32_t* q, size_t m, uint32_t v, hipStream_t s) {
  const size_t blocks = std::min<size_t>((std::max(n, m) + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_u32_pair, dim3(static_cast<unsigned>(std::max<size_t>(blocks, 1))), dim3(256), 0, s, p, n,
                     q, m, v);
}

void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n / 16 + 255) / 256 + 1, 8192);
  hipLaunchKernelGGL(k_fill_u8, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n, v);
}

void launch_allocate(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const float* depth,
                     uint32_t* cav, float* caches, Counter* counter, hipStream_t s) {
  AllocArgs a;
  a.vol = vol;
  for (int i = 0; i < 16; ++i) a.inv_vp[i] = frame.inverse_view_projection[i];
  a.depth = depth;
  a.cav = cav;
  a.caches = caches;
  a.counter = counter;
  a.width = cfg.width;
  a.height = cfg.height;
  a.R = cfg.cav_resolution;
  a.C = cfg.cav_cascades;
  a.stride = entry_stride(cfg.sh_bands);
  a.capacity = static_cast<uint32_t>(cfg.max_caches);
  a.transitions = cfg.cav_transition_size > 0.0f ? 1 : 0;
  const dim3 grid((cfg.width + 15) / 16, (cfg.height + 15) / 16);
  hipLaunchKernelGGL(k_allocate, grid, dim3(256), 0, s, a);
}

void launch_prepare_dispatch(Counter* counter, uint32_t capacity, hipStream_t s) {
  hipLaunchKernelGGL(k_prepare_dispatch, dim3(1), dim3(1), 0, s, counter, capacity);
}

}  // namespace drvk
