// Stage B: scene voxelization, temporal blend and the voxel mip chain.
//
// Restates Voxelization::VoxelizeScene (rendering/voxelization.cpp:90-176):
// the conservative raster pass shader/voxelize.{vert,geom,frag}, the blend
// shader/voxelblend.comp:8-18 and the mip loop shader/voxelmipmap.comp:8-12.
//
// MI355X design: the geometry-shader raster becomes one thread per
// triangle that evaluates the geometry shader's pushed-out edge planes and
// depth plane directly and walks the covered pixels of its dominant-axis
// projection; the blend is a flat byte stream; each mip level is an exact
// integer 2x2x2 mean.  After the mips a "corner-expanded" copy of every level
// packs the 8 texels of each trilinear footprint into one uint64 so the cone
// tracer issues one 8-byte load per level instead of eight byte gathers.
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

struct VoxelizeArgs {
  float vmin[3], vmax[3];
  float world[16];
  const float* positions;
  const uint32_t* indices;
  uint32_t n_tri;
  int res;
  uint8_t* target;
};

__device__ __forceinline__ void store_voxel(uint8_t* target, int res, int side, int sx, int sy, int sz) {
  // UnswizzlePos, voxelize.frag:16-19
  int x, y, z;
  if (side == 0) {
    x = sz; y = sy; z = sx;
  } else if (side == 1) {
    x = sx; y = sz; z = sy;
  } else {
    x = sx; y = sy; z = sz;
  }
  // imageStore outside the image has no effect.
  if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return;
  target[(static_cast<size_t>(z) * res + y) * res + x] = 255;
}

__global__ __launch_bounds__(256) void k_voxelize(VoxelizeArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tri) return;
  const int res = a.res;
  const float fres = static_cast<float>(res);
  const float h = 1.0f / fres;
  // voxelize.vert:15-23
  float P[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const uint32_t vi = a.indices ? a.indices[3 * t + v] : 3 * t + v;
    const float* p = a.positions + 3 * static_cast<size_t>(vi);
    float w4[4];
    mat_vec(a.world, p[0], p[1], p[2], 1.0f, w4);
#pragma unroll
    for (int i = 0; i < 3; ++i) P[v][i] = (w4[i] - a.vmin[i]) / (a.vmax[i] - a.vmin[i]) * 2.0f - 1.0f;
  }
  // voxelize.geom:43-48: dominant axis of the (normalized) face normal
  const float v0x = P[1][0] - P[0][0], v0y = P[1][1] - P[0][1], v0z = P[1][2] - P[0][2];
  const float v1x = P[2][0] - P[0][0], v1y = P[2][1] - P[0][1], v1z = P[2][2] - P[0][2];
  const float c0 = v0y * v1z - v0z * v1y, c1 = v0z * v1x - v0x * v1z, c2 = v0x * v1y - v0y * v1x;
  const float cl = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
  const float an0 = fabsf(c0 / cl), an1 = fabsf(c1 / cl), an2 = fabsf(c2 / cl);
  int side = an0 > an1 ? 0 : 1;
  side = (side == 0 ? an0 : an1) > an2 ? side : 2;
  // voxelize.geom:53-75: swizzle into the raster plane
  float r[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    if (side == 0) {
      r[v][0] = P[v][2]; r[v][1] = P[v][1]; r[v][2] = P[v][0];
    } else if (side == 1) {
      r[v][0] = P[v][0]; r[v][1] = P[v][2]; r[v][2] = P[v][1];
    } else {
      r[v][0] = P[v][0]; r[v][1] = P[v][1]; r[v][2] = P[v][2];
    }
  }
  // voxelize.geom:81-83: raster AABB grown by half a voxel, in pixels
  float bb0 = fminf(fminf(r[0][0], r[1][0]), r[2][0]) - h;
  float bb1 = fminf(fminf(r[0][1], r[1][1]), r[2][1]) - h;
  float bb2 = fmaxf(fmaxf(r[0][0], r[1][0]), r[2][0]) + h;
  float bb3 = fmaxf(fmaxf(r[0][1], r[1][1]), r[2][1]) + h;
  bb0 = (bb0 * 0.5f + 0.5f) * fres;
  bb1 = (bb1 * 0.5f + 0.5f) * fres;
  bb2 = (bb2 * 0.5f + 0.5f) * fres;
  bb3 = (bb3 * 0.5f + 0.5f) * fres;
  // voxelize.geom:86-102: edge lines cross((d,0),(p,1)) = (d.y, -d.x, d.x p.y - d.y p.x),
  // oriented by the winding and pushed out by h * (|a| + |b|).
  const float ax = r[0][0] - r[2][0], ay = r[0][1] - r[2][1];
  const float bx = r[1][0] - r[0][0], by = r[1][1] - r[0][1];
  const float ex = r[2][0] - r[1][0], ey = r[2][1] - r[1][1];
  float pl[3][3] = {{ay, -ax, ax * r[2][1] - ay * r[2][0]},
                    {by, -bx, bx * r[0][1] - by * r[0][0]},
                    {ey, -ex, ex * r[1][1] - ey * r[1][0]}};
  const float wind_raw = ax * by - bx * ay;
  const float wind = wind_raw > 0.0f ? 1.0f : (wind_raw < 0.0f ? -1.0f : 0.0f);
  if (wind == 0.0f) return;  // zero-area primitive
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pl[i][0] = pl[i][0] * wind;
    pl[i][1] = pl[i][1] * wind;
    pl[i][2] = pl[i][2] * wind;
    pl[i][2] = pl[i][2] - (h * fabsf(pl[i][0]) + h * fabsf(pl[i][1]));
  }
  // voxelize.geom:118-128: expanded vertices = pairwise line intersections,
  // each keeping its original depth.
  float V[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const float* p = pl[v == 0 ? 0 : (v == 1 ? 1 : 2)];
    const float* q = pl[v == 0 ? 1 : (v == 1 ? 2 : 0)];
    const float cx = p[1] * q[2] - p[2] * q[1];
    const float cy = p[2] * q[0] - p[0] * q[2];
    const float cw = p[0] * q[1] - p[1] * q[0];
    V[v][0] = cx / cw;
    V[v][1] = cy / cw;
    V[v][2] = r[v][2];
  }
  const float ex1 = V[1][0] - V[0][0], ey1 = V[1][1] - V[0][1];
  const float ex2 = V[2][0] - V[0][0], ey2 = V[2][1] - V[0][1];
  const float area = ex1 * ey2 - ex2 * ey1;
  if (!(area != 0.0f) || !isfinite(area)) return;
  const float dz1 = V[1][2] - V[0][2], dz2 = V[2][2] - V[0][2];
  // dFdxCoarse / dFdyCoarse of gl_FragCoord.z * res == NDC depth gradient
  const float gx = (dz1 * ey2 - dz2 * ey1) / area;
  const float gy = (dz2 * ex1 - dz1 * ex2) / area;
  const float max_change = sqrtf(gx * gx + gy * gy) * 1.414f;
  // voxelize.frag:26-30 keeps the pixel centres c = p + 0.5 with
  // bb_lo <= c <= bb_hi, i.e. ceil(bb_lo - 0.5) <= p <= floor(bb_hi - 0.5).
  // bb - 0.5 is exact for bb >= 0.25 (Sterbenz up to 1, on bb's grid above);
  // a bound below 0.5 only ever yields p <= 0 (lower) or p < 0 (upper), as
  // the exact value does, so the ranges hold exactly the passing centres.
  const int px0 = max(0, to_int(ceilf(bb0 - 0.5f)));
  const int py0 = max(0, to_int(ceilf(bb1 - 0.5f)));
  const int px1 = min(res - 1, to_int(floorf(bb2 - 0.5f)));
  const int py1 = min(res - 1, to_int(floorf(bb3 - 0.5f)));
  // res is a power of two, so x / res == x * h exactly.
  for (int py = py0; py <= py1; ++py) {
    const float fy = static_cast<float>(py) + 0.5f;
    const float ny = fy * h * 2.0f - 1.0f;
    for (int px = px0; px <= px1; ++px) {
      const float fx = static_cast<float>(px) + 0.5f;
      const float nx = fx * h * 2.0f - 1.0f;
      const bool in = (pl[0][0] * nx + pl[0][1] * ny + pl[0][2] <= 0.0f) &&
                      (pl[1][0] * nx + pl[1][1] * ny + pl[1][2] <= 0.0f) &&
                      (pl[2][0] * nx + pl[2][1] * ny + pl[2][2] <= 0.0f);
      if (!in) continue;
      const float zn = V[0][2] + gx * (nx - V[0][0]) + gy * (ny - V[0][1]);
      const float zw = zn * 0.5f + 0.5f;
      // voxelize.frag:33-56
      const float vz = zw * fres;
      const int iz = to_int(vz);
      store_voxel(a.target, res, side, px, py, iz);
      if (iz != to_int(vz - max_change)) store_voxel(a.target, res, side, px, py, iz - 1);
      if (iz != to_int(vz + max_change)) store_voxel(a.target, res, side, px, py, iz + 1);
    }
  }
}

// voxelblend.comp:8-18 (old + sign(target - old) * AdaptionFactor, stored as
// unorm8) and voxelmipmap.comp:8-12 (the linear sample at the centre of each
// 2x2x2 child block is the 8-texel mean, computed exactly and rounded half
// up), fused: one workgroup per 16^3 brick of level 0 blends its texels and
// builds the brick's levels 1..4 in LDS; the first workgroup of
// k_voxel_corners then builds the levels above from the bricks' top texels.
constexpr int kBrick = 16;

struct MipArgs {
  const uint8_t* target;
  uint8_t* mips;
  uint32_t mip_offset[16];
  int res, levels;
  int brick, brick_levels;  // B = min(16, res), log2(B)
  float f;
};

// voxelblend.comp:8-18: old + sign(target - old) * AdaptionFactor, unorm8.
__device__ __forceinline__ uint8_t blend_voxel(uint32_t target, uint32_t old, float f) {
  const float tv = static_cast<float>(target) / 255.0f;
  const float o = static_cast<float>(old) / 255.0f;
  const float d = tv - o;
  const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
  float v = o + sg * f;
  v = fminf(1.0f, fmaxf(0.0f, v));
  return static_cast<uint8_t>(floorf(v * 255.0f + 0.5f));
}

__device__ __forceinline__ uint8_t mean8(const uint8_t* s, int w, int x, int y, int z) {
  const uint8_t* s0 = s + (static_cast<size_t>(2 * z) * w + 2 * y) * w + 2 * x;
  const uint8_t* s1 = s0 + static_cast<size_t>(w) * w;
  const int sum = s0[0] + s0[1] + s0[w] + s0[w + 1] + s1[0] + s1[1] + s1[w] + s1[w + 1];
  return static_cast<uint8_t>((sum + 4) >> 3);
}

// Dynamic LDS: the brick's levels, B^3 + (B/2)^3 + ... bytes.
__global__ __launch_bounds__(256) void k_voxel_blend_mips(MipArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int B = a.brick, res = a.res;
  const int nb = res / B;
  const int bx = static_cast<int>(blockIdx.x) % nb, by = (static_cast<int>(blockIdx.x) / nb) % nb,
            bz = static_cast<int>(blockIdx.x) / (nb * nb);
  const int tid = static_cast<int>(threadIdx.x);
  // Level 0: voxelblend.comp:8-18.  With 16^3 bricks a thread blends one
  // 16-byte row (one load of target and volume each).
  if (B == kBrick) {
    const int y = tid % B, z = tid / B;
    const size_t g = (static_cast<size_t>(bz * B + z) * res + (by * B + y)) * res + bx * B;
    const uint4 tr = *reinterpret_cast<const uint4*>(a.target + g);
    const uint4 orow = *reinterpret_cast<const uint4*>(a.mips + g);
    const uint32_t tw[4] = {tr.x, tr.y, tr.z, tr.w}, ow[4] = {orow.x, orow.y, orow.z, orow.w};
    uint32_t nw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t packed = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        packed |= static_cast<uint32_t>(blend_voxel((tw[k] >> (8 * b)) & 0xffu, (ow[k] >> (8 * b)) & 0xffu, a.f))
                  << (8 * b);
      nw[k] = packed;
    }
    const uint4 nr = make_uint4(nw[0], nw[1], nw[2], nw[3]);
    *reinterpret_cast<uint4*>(a.mips + g) = nr;
    *reinterpret_cast<uint4*>(lds + tid * 16) = nr;
  } else {
    for (int t = tid; t < B * B * B; t += 256) {
      const int x = t % B, y = (t / B) % B, z = t / (B * B);
      const size_t g = (static_cast<size_t>(bz * B + z) * res + (by * B + y)) * res + (bx * B + x);
      const uint8_t q = blend_voxel(a.target[g], a.mips[g], a.f);
      a.mips[g] = q;
      lds[t] = q;
    }
  }
  __syncthreads();
  // Levels 1..log2(B) inside the brick: voxelmipmap.comp:8-12.
  int src = 0, dst = B * B * B;
  for (int k = 1; k <= a.brick_levels; ++k) {
    const int sp = B >> (k - 1), sk = B >> k, wk = res >> k;
    uint8_t* out = a.mips + a.mip_offset[k];
    for (int t = tid; t < sk * sk * sk; t += 256) {
      const int x = t % sk, y = (t / sk) % sk, z = t / (sk * sk);
      const uint8_t q = mean8(lds + src, sp, x, y, z);
      lds[dst + t] = q;
      out[(static_cast<size_t>(bz * sk + z) * wk + (by * sk + y)) * wk + (bx * sk + x)] = q;
    }
    __syncthreads();
    src = dst;
    dst += sk * sk * sk;
  }
}

struct CornerArgs {
  uint8_t* mips;
  VoxelCorner* corners;
  uint32_t mip_offset[16];
  uint32_t corner_offset[17];
  uint32_t quad_offset[17];  // threads (four corners of a row each) before level l
  int res;
  int levels;
  int brick, brick_levels;   // levels 0..brick_levels come from the brick kernel
};

// Four consecutive corners of one row of level l (quad q of the level): their
// 2x2x2 footprints share 5 texels in each of 4 rows, which are loaded once
// and all before the first use (the work is bound by load latency).
__device__ __forceinline__ void corner_quad(const CornerArgs& a, const uint8_t* lvl, int l, uint32_t q) {
  const int w = a.res >> l;
  const uint32_t wp = static_cast<uint32_t>(w) + 1u, qpr = (wp + 3u) / 4u;
  const uint32_t row = q / qpr, qx = q - row * qpr;  // row = (cz + 1) * wp + (cy + 1)
  const uint32_t rz = row / wp, ry = row - rz * wp;
  const int cx0 = static_cast<int>(4u * qx) - 1, cy = static_cast<int>(ry) - 1, cz = static_cast<int>(rz) - 1;
  uint32_t t[4][5];  // [y bit | z bit << 1][x - cx0]
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int y = min(w - 1, max(0, cy + (r & 1)));
    const int z = min(w - 1, max(0, cz + (r >> 1)));
    const uint8_t* src = lvl + (static_cast<size_t>(z) * w + y) * w;
#pragma unroll
    for (int c = 0; c < 5; ++c) t[r][c] = src[min(w - 1, max(0, cx0 + c))];
  }
  VoxelCorner* dst = a.corners + a.corner_offset[l] + static_cast<size_t>(row) * wp;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int cx = cx0 + k;
    if (cx > w - 1) break;
    // Per row r = dy + 2 dz: A_r = t at x, B_r = t at x + 1 minus t at x.
    // The record's words are the trilinear polynomial's coefficient pairs
    // (A, B) of row 0, then their differences in y, in z and in y and z.
    int A[4], B[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      A[r] = static_cast<int>(t[r][k]);
      B[r] = static_cast<int>(t[r][k + 1]) - A[r];
    }
    const int lo[4] = {A[0], A[1] - A[0], A[2] - A[0], A[3] - A[2] - A[1] + A[0]};
    const int hi[4] = {B[0], B[1] - B[0], B[2] - B[0], B[3] - B[2] - B[1] + B[0]};
    VoxelCorner c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t h0 = __builtin_bit_cast(uint16_t, static_cast<_Float16>(static_cast<float>(lo[r])));
      const uint32_t h1 = __builtin_bit_cast(uint16_t, static_cast<_Float16>(static_cast<float>(hi[r])));
      c.row[r] = h0 | (h1 << 16);
    }
    dst[cx + 1] = c;
  }
}

// The corner-expanded copy of every level, and the levels above the brick
// size (more than one brick per axis).  Workgroup 0 builds those upper
// levels from level brick_levels in LDS (voxelmipmap.comp:8-12), stores
// them and expands their corners from LDS; the other workgroups expand the
// levels the brick kernel wrote.  (The kernel boundary after the brick
// kernel makes every brick's writes visible; an in-kernel device-scope fence
// per brick costs an L2 write-back on this multi-XCD part.)
// Dynamic LDS: (res / B)^3 + (res / 2B)^3 + ... bytes when upper levels exist.
__global__ __launch_bounds__(256) void k_voxel_corners(CornerArgs a) {
  extern __shared__ uint8_t up[];
  const int top = a.brick_levels + 1;  // first upper level
  if (blockIdx.x == 0) {
    if (top >= a.levels) return;
    const int nb = a.res / a.brick;
    const int tid = static_cast<int>(threadIdx.x);
    const uint8_t* lb = a.mips + a.mip_offset[a.brick_levels];
    for (int t = tid; t < nb * nb * nb; t += 256) up[t] = lb[t];
    __syncthreads();
    int usrc = 0, udst = nb * nb * nb;
    for (int k = top; k < a.levels; ++k) {
      const int wp = a.res >> (k - 1), wk = a.res >> k;
      uint8_t* out = a.mips + a.mip_offset[k];
      for (int t = tid; t < wk * wk * wk; t += 256) {
        const int x = t % wk, y = (t / wk) % wk, z = t / (wk * wk);
        const uint8_t q = mean8(up + usrc, wp, x, y, z);
        up[udst + t] = q;
        out[t] = q;
      }
      __syncthreads();
      const uint32_t n = a.quad_offset[k + 1] - a.quad_offset[k];
      for (uint32_t q = tid; q < n; q += 256u) corner_quad(a, up + udst, k, q);
      usrc = udst;
      udst += wk * wk * wk;
    }
    return;
  }
  const uint32_t total = a.quad_offset[min(top, a.levels)];
  for (uint32_t i = (blockIdx.x - 1u) * blockDim.x + threadIdx.x; i < total; i += (gridDim.x - 1u) * blockDim.x) {
    int l = 0;
    while (l + 1 < a.levels && i >= a.quad_offset[l + 1]) ++l;
    corner_quad(a, a.mips + a.mip_offset[l], l, i - a.quad_offset[l]);
  }
}

}  // namespace

void launch_voxelize(const drv_volume& vol, int res, const float* positions, const uint32_t* indices, uint32_t n_tri,
                     const float* world, uint8_t* target, hipStream_t s) {
  if (n_tri == 0) return;
  VoxelizeArgs a;
  for (int i = 0; i < 3; ++i) {
    a.vmin[i] = vol.volume_world_min[i];
    a.vmax[i] = vol.volume_world_max[i];
  }
  for (int i = 0; i < 16; ++i) a.world[i] = world[i];
  a.positions = positions;
  a.indices = indices;
  a.n_tri = n_tri;
  a.res = res;
  a.target = target;
  hipLaunchKernelGGL(k_voxelize, dim3((n_tri + 255) / 256), dim3(256), 0, s, a);
}

void launch_voxel_blend_mips(const VoxelLayout& L, const uint8_t* target, uint8_t* mips, float adaption,
                             hipStream_t s) {
  MipArgs a;
  a.target = target;
  a.mips = mips;
  for (int l = 0; l < L.levels && l < 16; ++l) a.mip_offset[l] = static_cast<uint32_t>(L.mip_offset[l]);
  a.res = L.res;
  a.levels = L.levels;
  a.brick = std::min(kBrick, L.res);
  int bl = 0;
  while ((1 << (bl + 1)) <= a.brick) ++bl;
  a.brick_levels = bl;
  a.f = adaption;
  const int nb = L.res / a.brick;
  size_t lds = 0;
  for (int k = 0; k <= bl; ++k) lds += static_cast<size_t>(a.brick >> k) * (a.brick >> k) * (a.brick >> k);
  hipLaunchKernelGGL(k_voxel_blend_mips, dim3(static_cast<unsigned>(nb * nb * nb)), dim3(256), lds, s, a);
}

void launch_voxel_corners(const VoxelLayout& L, uint8_t* mips, VoxelCorner* corners, hipStream_t s) {
  CornerArgs a;
  a.mips = mips;
  a.corners = corners;
  a.res = L.res;
  a.levels = L.levels;
  a.brick = std::min(kBrick, L.res);
  int bl = 0;
  while ((1 << (bl + 1)) <= a.brick) ++bl;
  a.brick_levels = bl;
  for (int l = 0; l < L.levels; ++l) {
    a.mip_offset[l] = static_cast<uint32_t>(L.mip_offset[l]);
    a.corner_offset[l] = static_cast<uint32_t>(L.corner_offset[l]);
  }
  a.corner_offset[L.levels] = static_cast<uint32_t>(L.corner_count);
  uint32_t quads = 0;
  for (int l = 0; l < L.levels; ++l) {
    a.quad_offset[l] = quads;
    const uint32_t wp = static_cast<uint32_t>(L.res >> l) + 1u;
    quads += wp * wp * ((wp + 3u) / 4u);
  }
  a.quad_offset[L.levels] = quads;
  const uint32_t brick_quads = a.quad_offset[std::min(bl + 1, L.levels)];
  size_t upper = 0;
  if (bl + 1 < L.levels)
    for (int w = L.res / a.brick; w >= 1; w /= 2) upper += static_cast<size_t>(w) * w * w;
  // Workgroup 0 for the upper levels, the rest for the brick levels.
  const unsigned blocks = 1u + static_cast<unsigned>(std::min<size_t>((brick_quads + 255) / 256, 8192));
  hipLaunchKernelGGL(k_voxel_corners, dim3(blocks), dim3(256), upper, s, a);
}

}  // namespace drvk
