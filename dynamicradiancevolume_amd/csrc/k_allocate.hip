// Stage A: G-buffer-driven cache allocation into the clipmap address volume.
//
// Restates shader/cacheGather.comp:93-164 (per-pixel cell, 16x16 tile
// dedupe with the leader rule, claim of the 8 corner cells) and
// shader/cachePrepareLighting.comp:8-14, driven as in
// Renderer::AllocateCaches (rendering/renderer.cpp:951-992).
//
// MI355X design: the reference's per-corner CAS lock and append form a
// chain of dependent atomics on a few hot cells.  Here a claim is a plain
// store of a mark (pass 1: one 256-thread workgroup per 16x16 pixel tile,
// the two tile lists in LDS), and the claimed cells are numbered afterwards
// by a two-launch count / assign over the address volume in a fixed cell
// order (pass 2).  The set of claimed cells, every cache position and the
// demanded count are the reference's; the append order, which the reference
// leaves to its atomics, becomes deterministic for the claimed cells and
// groups spatial neighbours into the same gather tile.
#include <algorithm>
#include <cmath>

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
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

// Marks pass 1 leaves in a claimed cell until pass 2 gives it its cache
// index.  A cell of slab c + 1 at x = 0 is also the x = R corner of cascade c
// (the reference's corner offsets run into the next slab of the atlas);
// which cascade's position it gets depends on whose claim lands, in the
// reference (CAS order) as here (store order).
constexpr uint32_t kClaimed = 0xFFFFFFFFu;      // claimed by a corner of the cell's own cascade
constexpr uint32_t kClaimedLeft = 0xFFFFFFFEu;  // claimed by the x = R corner of the cascade before

struct AllocArgs {
  drv_volume vol;
  float inv_vp[16];
  float inv_size[DRV_MAX_CASCADES];  // 1 / world_voxel_size where that is a power of two (exact)
  const float* depth;
  uint32_t* cav;
  Counter* counter;
  float4* phantoms;  // positions of appended corners outside the CAV, [capacity]
  int width, height, R, C;
  uint32_t capacity;
  int transitions;
};

// A cell of cascade c: local coordinates and the flat index of
// GetCache1DCoord (cacheGather.comp:20-30), which the leader rule compares.
struct Cell {
  int x, y, z, flat;
};

// GetCache1DCoord, cacheGather.comp:20-30.  With power-of-two voxel sizes
// (kPow2) the divisions are multiplications by the exact reciprocal: the
// same correctly rounded quotient.
template <bool kPow2>
__device__ __forceinline__ Cell cell_of(const AllocArgs& a, f3 wp, int c) {
  const drv_cascade& k = a.vol.cascades[c];
  float qx, qy, qz;
  if (kPow2) {
    const float s = a.inv_size[c];
    qx = (wp.x - k.min[0]) * s;
    qy = (wp.y - k.min[1]) * s;
    qz = (wp.z - k.min[2]) * s;
  } else {
    qx = (wp.x - k.min[0]) / k.world_voxel_size;
    qy = (wp.y - k.min[1]) / k.world_voxel_size;
    qz = (wp.z - k.min[2]) / k.world_voxel_size;
  }
  const int R = a.R;
  Cell e;
  e.x = min(R - 1, max(0, to_int(qx)));
  e.y = min(R - 1, max(0, to_int(qy)));
  e.z = min(R - 1, max(0, to_int(qz)));
  e.flat = e.x + e.y * R + e.z * R * R + c * R * R * R;
  return e;
}

// Entry.Position of corner (lx, ly, lz) of cascade c, cacheGather.comp:75-77.
__device__ __forceinline__ float4 corner_position(const drv_volume& vol, int c, int lx, int ly, int lz) {
  const drv_cascade& k = vol.cascades[c];
  return make_float4(static_cast<float>(lx) * k.world_voxel_size + k.min[0],
                     static_cast<float>(ly) * k.world_voxel_size + k.min[1],
                     static_cast<float>(lz) * k.world_voxel_size + k.min[2], 0.0f);
}

// AllocateCaches, cacheGather.comp:32-91, first half: the claims.  The
// reference locks each of the 8 corner cells with a CAS and the winner
// appends; the set of claimed cells and each cell's position do not depend
// on which claim wins, so a claim here is a plain store of a mark (no round
// trip) and the appends of claimed cells happen in pass 2.  A lock on a cell
// outside the image "succeeds" (image atomics outside the image return 0)
// and appends a cache no cell points to: those appends are made here, one
// atomicAdd per wave, into the phantom list.  Called by every lane of the
// wave (inactive lanes pass active = false) with the cell (bx, by, bz) of
// cascade c.
__device__ void mark_corners(const AllocArgs& a, bool active, int bx, int by, int bz, int c) {
  const int R = a.R;
  uint32_t outside = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    // offsets (0,0,0),(0,1,0),(0,0,1),(0,1,1),(1,0,0),(1,1,0),(1,0,1),(1,1,1)
    const int lx = bx + (i >> 2), ly = by + (i & 1), lz = bz + ((i >> 1) & 1);
    const int gx = lx + c * R;
    const bool inside = gx < a.C * R && ly < R && lz < R;  // the offsets are >= 0
    if (active && inside) a.cav[(lz * R + ly) * (a.C * R) + gx] = lx == R ? kClaimedLeft : kClaimed;
    if (active && !inside) outside |= 1u << i;
  }
  uint64_t masks[8];
  uint32_t total = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    masks[i] = __ballot((outside >> i) & 1u);
    total += static_cast<uint32_t>(__popcll(masks[i]));
  }
  if (total == 0) return;
  const unsigned lane = __lane_id();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&a.counter->phantoms, total);
  base = __shfl(base, 0);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if ((outside >> i) & 1u) {
      const uint32_t p = base + static_cast<uint32_t>(__popcll(masks[i] & lt_mask));
      if (p < a.capacity) a.phantoms[p] = corner_position(a.vol, c, bx + (i >> 2), by + (i & 1), bz + ((i >> 1) & 1));
    }
    base += static_cast<uint32_t>(__popcll(masks[i]));
  }
}

// 16x16 pixel tiles per pass-1 workgroup.
constexpr int kAllocTiles = 2;

template <bool kPow2>
__global__ __launch_bounds__(256) void k_allocate(AllocArgs a) {
  // Two horizontally neighbouring 16x16 tiles per workgroup: both tiles'
  // depths are loaded before either is unprojected, and both tiles' cell
  // lists are built before one barrier.  The leader rule stays per tile.
  __shared__ int list[kAllocTiles][256];
  __shared__ int list2[kAllocTiles][256];
  const int t = threadIdx.x;
  const int lx = t & 15, ly = t >> 4;
  const int y = blockIdx.y * 16 + ly;
  float d[kAllocTiles];
#pragma unroll
  for (int h = 0; h < kAllocTiles; ++h) {
    const int x = (blockIdx.x * kAllocTiles + h) * 16 + lx;
    d[h] = (x < a.width && y < a.height) ? a.depth[static_cast<size_t>(y) * a.width + x] : 0.0f;
  }
  int own[kAllocTiles], own2[kAllocTiles], c[kAllocTiles];
  Cell e[kAllocTiles], e2[kAllocTiles];
#pragma unroll
  for (int h = 0; h < kAllocTiles; ++h) {
    const int x = (blockIdx.x * kAllocTiles + h) * 16 + lx;
    own[h] = -1;
    own2[h] = -1;
    c[h] = -1;
    e[h] = Cell{0, 0, 0, -1};
    e2[h] = Cell{0, 0, 0, -1};
    // Pixels outside the image read depth 0: no geometry, as in the reference.
    if (d[h] > 0.0001f) {
      const f3 wp = unproject_pixel(a.inv_vp, a.width, a.height, x, y, d[h]);
      c[h] = cascade_of(a.vol, a.C, wp);
      e[h] = cell_of<kPow2>(a, wp, c[h]);
      own[h] = e[h].flat;
      if (a.transitions) {
        const float tr = cascade_transition(a.vol, c[h], wp);
        if (tr > 0.0f && c[h] < a.C - 1) {
          e2[h] = cell_of<kPow2>(a, wp, c[h] + 1);
          own2[h] = e2[h].flat;
        }
      }
    }
    list[h][t] = own[h];
    list2[h][t] = own2[h];
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < kAllocTiles; ++h) {
    // Leader rule, cacheGather.comp:142-150: no left / up / up-left neighbour
    // of the tile holds the same cell, or this is thread (0,0).  Threads on
    // the left column and top row compare against themselves and never lead.
    const int ux = max(0, lx - 1), uy = max(0, ly - 1);
    const int* l1 = list[h];
    const bool lead = (((l1[uy * 16 + lx] != own[h]) && (l1[ly * 16 + ux] != own[h]) &&
                        (l1[uy * 16 + ux] != own[h])) ||
                       (ux == lx && uy == ly)) &&
                      own[h] != -1;
    // Transition list, cacheGather.comp:154-162: mirrored towards right/down;
    // min(15, l + 1) == (15, 15) holds for the four threads with lx, ly >= 14.
    const int vx = min(15, lx + 1), vy = min(15, ly + 1);
    const int* l2 = list2[h];
    const bool lead2 = a.transitions &&
                       (((l2[vy * 16 + lx] != own2[h]) && (l2[ly * 16 + vx] != own2[h]) &&
                         (l2[vy * 16 + vx] != own2[h])) ||
                        (vx == 15 && vy == 15)) &&
                       own2[h] != -1;
    mark_corners(a, lead, e[h].x, e[h].y, e[h].z, c[h]);
    if (a.transitions) mark_corners(a, lead2, e2[h].x, e2[h].y, e2[h].z, c[h] + 1);
  }
}

// ---- pass 2: cache indices in cell order -------------------------------
//
// Claimed cells are numbered in cascade-major order of 4x4x4 bricks (cells
// of one brick are consecutive), so the caches a gather tile holds are
// neighbours in space; the phantoms follow them.  Each workgroup numbers a
// fixed chunk of that order: one launch counts the claims per chunk, the
// next adds the counts of the chunks before (at most kAllocChunks of them)
// and writes the indices, the entries and the dispatch counter.

struct AssignArgs {
  drv_volume vol;
  uint32_t* cav;
  float* caches;
  Counter* counter;
  const float4* phantoms;
  uint32_t* counts;  // [gridDim.x] claims per chunk; [kAllocChunks] claimed cells of the last allocation
  uint32_t* claimed;  // flat CAV index of every claimed cell, by cache index
  int R, C, stride;
  uint32_t n_cells, chunk, capacity;
};

// Logical cell L (cascade c, brick, cell in brick) -> cascade and local
// coordinates; bricks need R % 4 == 0, else rows of x.
struct CellCoord {
  int c, x, y, z;
};

// Every third bit of v, packed (the inverse of a 3D Morton interleave).
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v ^ (v >> 2)) & 0x030C30C3u;
  v = (v ^ (v >> 4)) & 0x0300F00Fu;
  v = (v ^ (v >> 8)) & 0x030000FFu;
  v = (v ^ (v >> 16)) & 0x000003FFu;
  return v;
}

__device__ __forceinline__ CellCoord cell_coord(const AssignArgs& a, uint32_t L) {
  const uint32_t R = static_cast<uint32_t>(a.R), R3 = R * R * R;
  const uint32_t c = L / R3, r = L - c * R3;
  uint32_t x, y, z;
  if ((R & 3u) == 0u) {
    const uint32_t nb = R >> 2, b = r >> 6, w = r & 63u;
    uint32_t bx, by, bz;
    if ((nb & (nb - 1u)) == 0u) {
      bx = compact3(b);
      by = compact3(b >> 1);
      bz = compact3(b >> 2);
    } else {
      bz = b / (nb * nb);
      const uint32_t rem = b - bz * nb * nb;
      by = rem / nb;
      bx = rem - by * nb;
    }
    x = (bx << 2) | (w & 3u);
    y = (by << 2) | ((w >> 2) & 3u);
    z = (bz << 2) | (w >> 4);
  } else {
    z = r / (R * R);
    const uint32_t q = r - z * R * R;
    y = q / R;
    x = q - y * R;
  }
  return CellCoord{static_cast<int>(c), static_cast<int>(x), static_cast<int>(y), static_cast<int>(z)};
}

__device__ __forceinline__ uint32_t cell_flat(const AssignArgs& a, const CellCoord& k) {
  return static_cast<uint32_t>((k.z * a.R + k.y) * (a.C * a.R) + k.c * a.R + k.x);
}

__device__ __forceinline__ bool is_claim(uint32_t v) { return v >= kClaimedLeft; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// renderer.cpp:969-970 clears the whole address volume before each
// allocation; only the cells the previous allocation claimed are non-zero
// (the volume is cleared once at creation), so only those are cleared, with
// the counter.
__global__ __launch_bounds__(256) void k_allocate_clear(uint32_t* cav, const uint32_t* claimed,
                                                       const uint32_t* counts, Counter* counter) {
  const uint32_t n = counts[kAllocChunks];
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) cav[claimed[i]] = 0u;
  if (blockIdx.x == 0 && threadIdx.x < sizeof(Counter) / 4) reinterpret_cast<uint32_t*>(counter)[threadIdx.x] = 0u;
}

// Steps (256 cells each, 1024 with bricks) a workgroup loads ahead in the
// count / assign scans.
constexpr int kScanAhead = 4;

// With bricks (R % 4 == 0) the 4 consecutive cells L .. L + 3 (L % 4 == 0)
// are one brick row, 4 consecutive x at a 16-byte aligned CAV address: a
// thread takes them with one load.  Chunks are multiples of 256 cells.
__device__ __forceinline__ uint4 cav_row(const AssignArgs& a, uint32_t L, uint32_t& flat, CellCoord& k) {
  k = cell_coord(a, L);
  flat = cell_flat(a, k);
  return *reinterpret_cast<const uint4*>(a.cav + flat);
}

__device__ __forceinline__ uint32_t claims4(const uint4& v) {
  return static_cast<uint32_t>(is_claim(v.x)) + is_claim(v.y) + is_claim(v.z) + is_claim(v.w);
}

__global__ __launch_bounds__(256) void k_allocate_count(AssignArgs a) {
  __shared__ uint32_t part[4];
  const uint32_t begin = blockIdx.x * a.chunk;
  const uint32_t end = min(a.n_cells, begin + a.chunk);
  uint32_t n = 0;
  if ((a.R & 3) == 0) {
    for (uint32_t L0 = begin + 4u * threadIdx.x; L0 < end; L0 += 1024u * kScanAhead) {
      uint4 v[kScanAhead];
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) {
        const uint32_t L = L0 + 1024u * u;
        uint32_t flat;
        CellCoord k;
        v[u] = L < end ? cav_row(a, L, flat, k) : make_uint4(0u, 0u, 0u, 0u);
      }
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) n += claims4(v[u]);
    }
  } else {
    for (uint32_t L0 = begin + threadIdx.x; L0 < end; L0 += 256u * kScanAhead) {
      uint32_t v[kScanAhead];
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) {
        const uint32_t L = L0 + 256u * u;
        v[u] = L < end ? a.cav[cell_flat(a, cell_coord(a, L))] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) n += is_claim(v[u]);
    }
  }
  n = wave_sum(n);
  if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) a.counts[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Cache idx for the claimed cell k (CAV value v, flat index flat): its entry
// (Entry.Position from the claiming cascade, cacheGather.comp:75-77, and
// zero SH) and its place in the list of claimed cells.
__device__ __forceinline__ void assign_cell(const AssignArgs& a, const CellCoord& k, uint32_t v, uint32_t flat,
                                            uint32_t idx) {
  if (idx < a.capacity) {
    const bool left = v == kClaimedLeft;
    const float4 p = corner_position(a.vol, left ? k.c - 1 : k.c, left ? a.R : k.x, k.y, k.z);
    float4* e = reinterpret_cast<float4*>(a.caches + static_cast<size_t>(idx) * a.stride);
    e[0] = p;
    for (int j = 1; j < a.stride / 4; ++j) e[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  a.claimed[idx] = flat;
}

__global__ __launch_bounds__(256) void k_allocate_assign(AssignArgs a) {
  __shared__ uint32_t part[2][4];
  __shared__ uint32_t wcnt[2][4];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  // Claims before this chunk and in all chunks.
  uint32_t before = 0, all = 0;
  for (uint32_t i = t; i < gridDim.x; i += 256u) {
    const uint32_t v = a.counts[i];
    all += v;
    before += i < blockIdx.x ? v : 0u;
  }
  before = wave_sum(before);
  all = wave_sum(all);
  if (lane == 0) {
    part[0][w] = before;
    part[1][w] = all;
  }
  __syncthreads();
  uint32_t base = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
  const uint32_t n_claimed = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
  const uint32_t n_phantoms = a.counter->phantoms;
  const uint32_t total = n_claimed + n_phantoms;
  if (blockIdx.x == 0 && t == 0) {
    // cachePrepareLighting.comp:8-14 plus the live count and cache tiling used
    // by the gather (capacity clamp = robust buffer access).
    Counter* ctr = a.counter;
    ctr->total = total;
    ctr->groups_x = (total + 63u) / 64u;
    ctr->groups_y = 1u;
    ctr->groups_z = 1u;
    const uint32_t live = min(total, a.capacity);
    ctr->live = live;
    ctr->tiles = (live + kGatherTile - 1) / kGatherTile;
    ctr->splits = 1u;
    a.counts[kAllocChunks] = n_claimed;  // the cells the next allocation clears
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t begin = blockIdx.x * a.chunk;
  const uint32_t end = min(a.n_cells, begin + a.chunk);
  uint32_t step = 0;
  if ((a.R & 3) == 0) {
    // Bricks: a thread takes a brick row of 4 cells (one 16-byte load and,
    // when it holds a claim, one 16-byte store of the new CAV values); cells
    // are numbered in thread order, 1024 per step of the workgroup.
    for (uint32_t L0 = begin; L0 < end; L0 += 1024u * kScanAhead) {
      CellCoord k[kScanAhead];
      uint32_t flat[kScanAhead];
      uint4 v[kScanAhead];
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) {
        const uint32_t L = L0 + 1024u * u + 4u * t;
        k[u] = CellCoord{};
        flat[u] = 0u;
        v[u] = make_uint4(0u, 0u, 0u, 0u);
        if (L < end) v[u] = cav_row(a, L, flat[u], k[u]);
      }
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u, step ^= 1u) {
        const uint32_t cv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
        uint64_t m[4];
        uint32_t wave_n = 0, before_lane = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          m[i] = __ballot(is_claim(cv[i]));
          wave_n += static_cast<uint32_t>(__popcll(m[i]));
          before_lane += static_cast<uint32_t>(__popcll(m[i] & lt_mask));
        }
        if (lane == 0) wcnt[step][w] = wave_n;
        __syncthreads();  // wcnt[step] is rewritten two steps on, after the next barrier
        uint32_t off = base + before_lane;
        for (uint32_t j = 0; j < w; ++j) off += wcnt[step][j];
        uint32_t nv[4] = {cv[0], cv[1], cv[2], cv[3]};
        bool any = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (!is_claim(cv[i])) continue;
          any = true;
          CellCoord ki = k[u];
          ki.x += i;
          assign_cell(a, ki, cv[i], flat[u] + static_cast<uint32_t>(i), off);
          nv[i] = off + 1u;  // cacheGather.comp:79
          ++off;
        }
        if (any) *reinterpret_cast<uint4*>(a.cav + flat[u]) = make_uint4(nv[0], nv[1], nv[2], nv[3]);
        base += (wcnt[step][0] + wcnt[step][1]) + (wcnt[step][2] + wcnt[step][3]);
      }
    }
  } else {
    for (uint32_t L0 = begin; L0 < end; L0 += 256u * kScanAhead) {
      CellCoord k[kScanAhead];
      uint32_t flat[kScanAhead], v[kScanAhead];
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u) {
        const uint32_t L = L0 + 256u * u + t;
        k[u] = CellCoord{};
        flat[u] = 0u;
        v[u] = 0u;
        if (L < end) {
          k[u] = cell_coord(a, L);
          flat[u] = cell_flat(a, k[u]);
          v[u] = a.cav[flat[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < kScanAhead; ++u, step ^= 1u) {
        const bool claim = is_claim(v[u]);  // 0 past the end
        const uint64_t m = __ballot(claim);
        if (lane == 0) wcnt[step][w] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();  // wcnt[step] is rewritten two steps on, after the next barrier
        uint32_t off = base;
        for (uint32_t j = 0; j < w; ++j) off += wcnt[step][j];
        if (claim) {
          const uint32_t idx = off + static_cast<uint32_t>(__popcll(m & lt_mask));
          assign_cell(a, k[u], v[u], flat[u], idx);
          a.cav[flat[u]] = idx + 1u;  // cacheGather.comp:79
        }
        base += (wcnt[step][0] + wcnt[step][1]) + (wcnt[step][2] + wcnt[step][3]);
      }
    }
  }
  // The phantoms, after the claimed cells.
  for (uint32_t p = blockIdx.x * 256u + t; p < n_phantoms; p += gridDim.x * 256u) {
    const uint32_t idx = n_claimed + p;
    if (idx >= a.capacity) break;
    float4* e = reinterpret_cast<float4*>(a.caches + static_cast<size_t>(idx) * a.stride);
    e[0] = a.phantoms[p];
    for (int j = 1; j < a.stride / 4; ++j) e[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

}  // namespace

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n, v);
}

void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n / 16 + 255) / 256 + 1, 8192);
  hipLaunchKernelGGL(k_fill_u8, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n, v);
}

void launch_allocate(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const float* depth,
                     uint32_t* cav, float* caches, Counter* counter, float4* phantoms, uint32_t* counts,
                     uint32_t* claimed, hipStream_t s) {
  hipLaunchKernelGGL(k_allocate_clear, dim3(64), dim3(256), 0, s, cav, claimed, counts, counter);
  AllocArgs a;
  a.vol = vol;
  for (int i = 0; i < 16; ++i) a.inv_vp[i] = frame.inverse_view_projection[i];
  a.depth = depth;
  a.cav = cav;
  a.counter = counter;
  a.phantoms = phantoms;
  a.width = cfg.width;
  a.height = cfg.height;
  a.R = cfg.cav_resolution;
  a.C = cfg.cav_cascades;
  a.capacity = static_cast<uint32_t>(cfg.max_caches);
  a.transitions = cfg.cav_transition_size > 0.0f ? 1 : 0;
  bool pow2 = true;
  for (int c = 0; c < DRV_MAX_CASCADES; ++c) {
    a.inv_size[c] = 0.0f;
    if (c >= a.C) continue;
    int e = 0;
    const float m = std::frexp(vol.cascades[c].world_voxel_size, &e);  // size = m 2^e, m in [0.5, 1)
    if (m == 0.5f && e > -120 && e < 120) {
      a.inv_size[c] = std::ldexp(1.0f, 1 - e);
    } else {
      pow2 = false;
    }
  }
  const int tiles_x = (cfg.width + 15) / 16;
  const dim3 grid((tiles_x + kAllocTiles - 1) / kAllocTiles, (cfg.height + 15) / 16);
  if (pow2) {
    hipLaunchKernelGGL(k_allocate<true>, grid, dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_allocate<false>, grid, dim3(256), 0, s, a);
  }

  AssignArgs b;
  b.vol = vol;
  b.cav = cav;
  b.caches = caches;
  b.counter = counter;
  b.phantoms = phantoms;
  b.counts = counts;
  b.claimed = claimed;
  b.R = cfg.cav_resolution;
  b.C = cfg.cav_cascades;
  b.stride = entry_stride(cfg.sh_bands);
  b.n_cells = static_cast<uint32_t>(static_cast<uint64_t>(cfg.cav_cascades) * cfg.cav_resolution *
                                    cfg.cav_resolution * cfg.cav_resolution);
  b.capacity = static_cast<uint32_t>(cfg.max_caches);
  const uint32_t chunks = std::min<uint32_t>((b.n_cells + 1023u) / 1024u, static_cast<uint32_t>(kAllocChunks));
  b.chunk = ((b.n_cells + chunks - 1u) / chunks + 255u) & ~255u;
  const uint32_t groups = (b.n_cells + b.chunk - 1u) / b.chunk;
  hipLaunchKernelGGL(k_allocate_count, dim3(groups), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_allocate_assign, dim3(groups), dim3(256), 0, s, b);
}


}  // namespace drvk
