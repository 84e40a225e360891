// Device bodies of the light pass shared by the single-view kernels
// (k_gather.hip) and the batched kernels (k_batch.hip): the VPL / shadow-
// block records of one RSM texel, one (tile, split group) item of the SH
// gather and one (tile, coefficient) item of the split reduction
// (shader/cacheLightingRSM.comp:80-374).
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_sample.hpp"
#include "k_sh.hpp"

namespace drvk {

struct PrepArgs {
  float inv_lvp[16];
  float light_pos[3];
  float val_area_factor;
  float sampling_offset;
  float super_val_width;
  const uint16_t* flux;
  const int16_t* normal;
  const uint16_t* dls;
  const uint16_t* dls_shadow;
  int R;         // read level width (the Morton order's x range)
  int R_shadow;  // width of the shadow level
  int H;         // read level height: R, or R / 2 for a 2:1 RSM
  int H_shadow;
  uint32_t interval;
  int block;  // B = 2^lod texels per block side
  int flux_cut;
  Vpl* vpls;
  Vpl* vpls_lanes;  // optional lane-interleaved copy: [m % per_lane][m / per_lane], per_lane = n / 64
  // Optional split plan of the gather (k_plan_gather), done by the first thread.
  Counter* plan;
  uint32_t quantum, max_splits, target_items;
  ShadowBlock* blocks;
  // Optional march plans of the blocks (kMarchSteps records each) for the
  // voxel levels of vox.
  MarchStep* march;
  VoxelView vox;
};

// World position of an RSM sample at distance d from the light
// (cacheLightingRSM.comp:154-155, 175-176).
__device__ __forceinline__ f3 rsm_position(const PrepArgs& a, float s, float t, float d) {
  float clip[4];
  mat_vec(a.inv_lvp, s * 2.0f - 1.0f, t * 2.0f - 1.0f, 0.0f, 1.0f, clip);
  const float px = clip[0] / clip[3] - a.light_pos[0];
  const float py = clip[1] / clip[3] - a.light_pos[1];
  const float pz = clip[2] / clip[3] - a.light_pos[2];
  const float l = sqrtf(px * px + py * py + pz * pz);
  return make3(a.light_pos[0] + px / l * d, a.light_pos[1] + py / l * d, a.light_pos[2] + pz / l * d);
}

__device__ __forceinline__ float rg16f_texel(const uint16_t* img, int w, int h, int i, int j, int comp) {
  i = min(w - 1, max(0, i));
  j = min(h - 1, max(0, j));
  return half_to_float(img[2 * (static_cast<size_t>(j) * w + i) + comp]);
}

// Bilinear (d, d^2) of the shadow level at a block's sampling point, the
// block's first texel (x, y) plus the sampling offset
// (cacheLightingRSM.comp:171-174).
__device__ __forceinline__ void block_depth_moments(const PrepArgs& a, uint32_t x, uint32_t y, float dd[2]) {
  const float fR = static_cast<float>(a.R), fH = static_cast<float>(a.H);
  const float ss = (static_cast<float>(x) + a.sampling_offset) / fR;
  const float st = (static_cast<float>(y) + a.sampling_offset) / fH;
  const int w = a.R_shadow, h = a.H_shadow;
  const float u = ss * static_cast<float>(w) - 0.5f, vv = st * static_cast<float>(h) - 0.5f;
  const float fu = floorf(u), fv = floorf(vv);
  const int i0 = to_int(fu), j0 = to_int(fv);
  const float al = u - fu, be = vv - fv;
#pragma unroll
  for (int c = 0; c < 2; ++c)
    dd[c] = (1.0f - al) * (1.0f - be) * rg16f_texel(a.dls_shadow, w, h, i0, j0, c) +
            al * (1.0f - be) * rg16f_texel(a.dls_shadow, w, h, i0 + 1, j0, c) +
            (1.0f - al) * be * rg16f_texel(a.dls_shadow, w, h, i0, j0 + 1, c) +
            al * be * rg16f_texel(a.dls_shadow, w, h, i0 + 1, j0 + 1, c);
}

// Cone aperture of a block from its depth moments (cacheLightingRSM.comp:180-193).
__device__ __forceinline__ float block_aperture(const PrepArgs& a, const float dd[2]) {
  const float variance = dd[1] - dd[0] * dd[0];
  const float cand = sqrtf(variance) * 2.0f / dd[0];
  // GLSL max(x, y) = (x < y) ? y : x, so a NaN candidate keeps SuperValWidth.
  return (a.super_val_width < cand) ? cand : a.super_val_width;
}

// The VPL record m (and the shadow block record when m starts a block) of
// the light in `a`, n = R^2 VPLs; every lane of a wave must call it with
// m < n (n is a multiple of 64, so whole waves are in or out).
__device__ __forceinline__ void prepare_vpl(const PrepArgs& a, uint32_t m, uint32_t n) {
  uint32_t x, y;
  morton_decode(m, x, y);
  // cacheLightingRSM.comp:141-158
  const float fR = static_cast<float>(a.R), fH = static_cast<float>(a.H);
  const float s = (static_cast<float>(x) + 0.5f) / fR;
  const float t = (static_cast<float>(y) + 0.5f) / fH;
  const size_t i = static_cast<size_t>(y) * a.R + x;
  const float d = half_to_float(a.dls[2 * i]);
  const f3 p = rsm_position(a, s, t, d);
  const f3 nrm = unpack_normal16i(a.normal[2 * i], a.normal[2 * i + 1]);
  Vpl v;
  v.px = p.x;
  v.py = p.y;
  v.pz = p.z;
  v.disc_area = d * d * a.val_area_factor;
  v.fr = half_to_float(a.flux[3 * i]);
  v.fg = half_to_float(a.flux[3 * i + 1]);
  v.fb = half_to_float(a.flux[3 * i + 2]);
  // cacheLightingRSM.comp:240-243: with INDIRECT_SPECULAR a VAL whose flux
  // sums below 0.001 is skipped (`continue`); a zero flux adds nothing.
  if (a.flux_cut && (v.fr + v.fg) + v.fb < 0.001f) v.fr = v.fg = v.fb = 0.0f;
  v.pad0 = 0.0f;
  v.nx = nrm.x;
  v.ny = nrm.y;
  v.nz = nrm.z;
  v.pad1 = 0.0f;
  a.vpls[m] = v;
  if (a.vpls_lanes) {
    const uint32_t per_lane = n / 64u;
    a.vpls_lanes[static_cast<size_t>(m % per_lane) * 64u + m / per_lane] = v;
  }
  // Mean normal of a block's VPLs: the B x B texel square of a block is the
  // `interval` Morton-consecutive VPLs m .. m + interval - 1, i.e. adjacent
  // lanes of one wave when interval <= 64 (sum by butterfly); larger blocks
  // are summed by their first lane.
  float bsx = nrm.x, bsy = nrm.y, bsz = nrm.z;
  if (a.blocks && a.interval <= 64u) {
    for (uint32_t o = 1; o < a.interval; o <<= 1) {
      bsx += __shfl_xor(bsx, static_cast<int>(o));
      bsy += __shfl_xor(bsy, static_cast<int>(o));
      bsz += __shfl_xor(bsz, static_cast<int>(o));
    }
  }
  if (a.blocks && (m % a.interval) == 0u) {
    float dd[2];
    block_depth_moments(a, x, y, dd);
    const f3 ap = rsm_position(a, (static_cast<float>(x) + a.sampling_offset) / fR,
                               (static_cast<float>(y) + a.sampling_offset) / fH, dd[0]);
    const float rho = block_aperture(a, dd);
    float sx = bsx, sy = bsy, sz = bsz;
    if (a.interval > 64u) {
      sx = sy = sz = 0.0f;
      const int B = a.block;
      for (int dy = 0; dy < B; ++dy)
        for (int dx = 0; dx < B; ++dx) {
          const size_t t = static_cast<size_t>(y + dy) * a.R + (x + dx);
          const f3 vn = unpack_normal16i(a.normal[2 * t], a.normal[2 * t + 1]);
          sx += vn.x;
          sy += vn.y;
          sz += vn.z;
        }
    }
    const float nl = sqrtf(sx * sx + sy * sy + sz * sz);
    const float inv = nl > 0.0f ? 1.0f / nl : 0.0f;
    a.blocks[m / a.interval] = ShadowBlock{ap.x, ap.y, ap.z, rho, sx * inv, sy * inv, sz * inv, 2.0f / (1.0f - rho)};
  }
}

// Record n of the march plan of shadow block b (one thread per record,
// n_blocks x kMarchSteps of them).  The aperture is recomputed from the
// shadow level as the block's own thread computes it (the same value).
__device__ __forceinline__ void prepare_march(const PrepArgs& a, uint32_t idx) {
  const uint32_t b = idx / kMarchSteps, n = idx % kMarchSteps;
  uint32_t x, y;
  morton_decode(b * a.interval, x, y);
  float dd[2];
  block_depth_moments(a, x, y, dd);
  const float rho = block_aperture(a, dd);
  a.march[idx] = plan_march_step(a.vox, rho, 2.0f / (1.0f - rho), static_cast<int>(n));
}

// One (64-cache tile, group of 4 VPL splits) item of the SH gather, one
// split per wave of the 256-thread workgroup (all four waves call it with
// the same item): raw sums of the split's VPLs, combined over the four waves
// in LDS in a fixed order, one partial per group written coefficient-major,
// partial[(group * 27 + k) * partial_pitch + cache].
template <bool kSh2, bool kShadow>
__device__ __forceinline__ void gather_item(const float* __restrict__ caches, int stride, uint32_t live,
                                            const Vpl* __restrict__ vpls, uint32_t n_vpl,
                                            const float* __restrict__ vis, uint32_t vis_pitch, uint32_t interval,
                                            float* __restrict__ partial, uint32_t partial_pitch, uint32_t splits,
                                            uint32_t tile, uint32_t group, float (*part)[27][64]) {
  const uint32_t chunk = n_vpl / splits;
  const uint32_t lane = threadIdx.x & 63u;
  // Wave-uniform split index (readfirstlane lets the compiler keep the item
  // loop and the VPL / block addresses in SGPRs).
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n_out = kSh2 ? 27 : 12;
  const uint32_t split = group * 4u + w;
  const uint32_t cache = tile * 64u + lane;
  const bool valid = cache < live;
  const uint32_t c = valid ? cache : 0u;
  const float* e = caches + static_cast<size_t>(c) * stride;
  const float wx = e[0], wy = e[1], wz = e[2];
  // Raw sums; the SH evaluation factors are applied at write-out.
  ShSums acc;
  sh_zero(acc);
  if (split < splits) {
    const uint32_t m0 = split * chunk, m1 = m0 + chunk;
    uint32_t blk = m0 / interval;  // the split starts on a block boundary
    // The next block's shadowing is loaded a block ahead, so the skip test
    // below does not wait for it.
    float next = kShadow ? vis[static_cast<size_t>(blk) * vis_pitch + c] : 1.0f;
    for (uint32_t mb = m0; mb < m1; mb += interval, ++blk) {
      const float shadowing = next;
      if (kShadow) {
        if (mb + interval < m1) next = vis[static_cast<size_t>(blk + 1u) * vis_pitch + c];
        // A block that no lane of the wave sees would add only exact
        // zeros (finite flux and distances): its VPLs are skipped.
        if (__ballot(valid && shadowing != 0.0f) == 0ull) continue;
      }
#pragma unroll 4
      for (uint32_t m = mb; m < mb + interval; ++m) sh_accumulate<kSh2>(acc, vpls[m], wx, wy, wz, shadowing);
    }
  }
  // The workgroup's four splits of the tile are combined in LDS in a fixed
  // order, (s0 + s1) + (s2 + s3), so one partial per group goes to HBM.
  if (w & 1u) {
#pragma unroll
    for (int k = 0; k < 27; ++k)
      if (k < n_out) part[w >> 1][k][lane] = acc.s[k];
  }
  __syncthreads();
  if (!(w & 1u)) {
#pragma unroll
    for (int k = 0; k < 27; ++k)
      if (k < n_out) acc.s[k] += part[w >> 1][k][lane];
  }
  __syncthreads();
  if (w == 2u) {
#pragma unroll
    for (int k = 0; k < 27; ++k)
      if (k < n_out) part[0][k][lane] = acc.s[k];
  }
  __syncthreads();
  if (w == 0u && valid) {
    // Raw sums, coefficient-major: partial[(group * 27 + k) * pitch + cache].
    float* p = partial + static_cast<size_t>(group) * 27u * partial_pitch + cache;
#pragma unroll
    for (int k = 0; k < 27; ++k)
      if (k < n_out) p[static_cast<size_t>(k) * partial_pitch] = acc.s[k] + part[0][k][lane];
  }
}

// One (64-cache tile, coefficient x channel) item of the split reduction:
// the four waves sum interleaved quarters of the `groups` partials (lane =
// cache, every load a coalesced row), the quarters are combined in a fixed
// order through LDS, and the sum (with the SH factors) is added into the
// entry (cacheLightingRSM.comp:357-373).
__device__ __forceinline__ void reduce_item(float* caches, int stride, uint32_t live, const float* partial,
                                            uint32_t partial_pitch, uint32_t groups, uint32_t tile, uint32_t jc,
                                            const drv_sh_constants& shc, float (*part)[4][64]) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const size_t split_stride = static_cast<size_t>(27u) * partial_pitch;
  const uint32_t j = jc / 3u, ch = jc % 3u;
  const uint32_t c = tile * 64u + lane;
  const uint32_t cc = c < live ? c : 0u;
  const float* p = partial + static_cast<size_t>(jc) * partial_pitch + cc;
  const float* p0 = partial + static_cast<size_t>(ch) * partial_pitch + cc;  // raw sum of rad
  float v = 0.0f, v0 = 0.0f;
#pragma unroll 8
  for (uint32_t sp = w; sp < groups; sp += 4u) {
    v += p[sp * split_stride];
    if (j == 6u) v0 += p0[sp * split_stride];
  }
  part[0][w][lane] = v;
  part[1][w][lane] = v0;
  __syncthreads();
  if (w == 0u && c < live) {
    const float s = ((part[0][0][lane] + part[0][1][lane]) + part[0][2][lane]) + part[0][3][lane];
    const float s0 = ((part[1][0][lane] + part[1][1][lane]) + part[1][2][lane]) + part[1][3][lane];
    sh_add_coefficient(caches + static_cast<size_t>(c) * stride, j, ch, s, s0, shc);
  }
}

}  // namespace drvk
