// Stage D: cache-to-pixel interpolation.
//
// Restates shader/cacheApply.frag:28-195 with SampleCacheIrradiance
// (shader/lightcache.glsl:137-183), drawn by Renderer::ApplyCaches
// (rendering/renderer.cpp:1047-1079) as a full-screen pass with GL_ONE/GL_ONE
// blending into the HDR target.
//
// MI355X design: one lane per pixel over 16x16 pixel tiles (a 256-thread
// workgroup), so the 8 corner cells of neighbouring pixels share CAV lines and
// cache entries in L2.  Addresses that are unclaimed (CAV 0 -> 0xFFFFFFFF) or
// past the live count read as zero, as robust buffer access does in GL.
//
// With INDIRECT_SPECULAR (k_apply<true>) each of the 8 caches also
// contributes a trilinear lookup in its env-map of the specular atlas
// (cacheApply.frag:100-106, 139-154), and the diffuse colour is the
// metallic-weighted one of ComputeMaterialColors.
#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_specular.hpp"

namespace drvk {
namespace {

struct ApplyArgs {
  drv_volume vol;
  float inv_vp[16];
  drv_sh_constants shc;
  const float* depth;
  const int16_t* normal;
  const uint8_t* albedo;
  const uint32_t* cav;
  const float* caches;
  const Counter* counter;
  const float* srgb_lut;
  float* out;
  int width, height, R, C, stride, sh2, transitions, accumulate;
  // INDIRECT_SPECULAR
  const uint8_t* roughness_metallic;  // RG8 or null (roughness 1, metallic 0)
  const uint32_t* atlas;
  uint64_t level_offset[kSpecMaxLevels];
  uint32_t atlas_width, per_row, per_row_log2;  // caches per atlas row: a power of two
  int S, levels;
  float cam[3];
};

// One level's bilinear footprint inside a cache's S_l x S_l region: the
// lookup is clamped half a texel of the coarser level away from the region's
// border (cacheApply.frag:151-152), so the 2x2 footprint, its weights and
// its offset within the region are the same for all 8 caches.  The texture
// coordinate (lookup + cache offset) * S_l - 0.5 is split into the in-region
// part lookup * S_l - 0.5 (exact in FP32 for lookups >= half a texel) and
// the integer cache offset, which is the exact value of GL's coordinate.
struct Tap {
  int l;          // level
  int i0, j0;     // footprint origin within the region
  float w[4];     // GL 4.5 section 8.14.2 weights: (i0,j0), (i0+1,j0), (i0,j0+1), (i0+1,j0+1)
};

__device__ __forceinline__ Tap make_tap(const ApplyArgs& a, int l, const float* lookup) {
  Tap t;
  t.l = l;
  const float sl = static_cast<float>(a.S >> l);
  const float u = lookup[0] * sl - 0.5f, v = lookup[1] * sl - 0.5f;
  const float fu = floorf(u), fv = floorf(v);
  t.i0 = to_int(fu);
  t.j0 = to_int(fv);
  const float al = u - fu, be = v - fv;
  t.w[0] = (1.0f - al) * (1.0f - be);
  t.w[1] = al * (1.0f - be);
  t.w[2] = (1.0f - al) * be;
  t.w[3] = al * be;
  return t;
}

// R11G11B10F texel -> RGB.  GL's unsigned 11- and 10-bit floats have the
// 5-bit exponent and the bias (15) of IEEE half precision, so a channel's
// bits shifted to the top of a half's mantissa are exactly that half
// (normals, denormals, inf and NaN alike), and the hardware half -> float
// conversion decodes it: a shift, a mask and a conversion per channel.
__device__ __forceinline__ float half_bits(uint32_t h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
}

__device__ __forceinline__ void decode_texel(uint32_t x, float* rgb) {
  rgb[0] = half_bits((x << 4) & 0x7FF0u);   // bits 0-10: 6-bit mantissa, 5-bit exponent
  rgb[1] = half_bits((x >> 7) & 0x7FF0u);   // bits 11-21
  rgb[2] = half_bits((x >> 17) & 0x7FE0u);  // bits 22-31: 5-bit mantissa
}

// Bilinear tap of cache (cx, cy) at tap.l, clamp to the atlas edge.
__device__ __forceinline__ void tap_cache(const ApplyArgs& a, const Tap& t, uint32_t cx, uint32_t cy, float* out) {
  const int sl = a.S >> t.l;
  const int wl = static_cast<int>(a.atlas_width >> t.l);
  const int x0 = static_cast<int>(cx) * sl + t.i0, y0 = static_cast<int>(cy) * sl + t.j0;
  const int ia = min(max(x0, 0), wl - 1), ib = min(max(x0 + 1, 0), wl - 1);
  const int ja = min(max(y0, 0), wl - 1), jb = min(max(y0 + 1, 0), wl - 1);
  const uint32_t* lv = a.atlas + a.level_offset[t.l];
  const uint32_t ra = static_cast<uint32_t>(ja) * static_cast<uint32_t>(wl), rb = static_cast<uint32_t>(jb) * static_cast<uint32_t>(wl);
  const uint32_t q00 = lv[ra + ia], q10 = lv[ra + ib], q01 = lv[rb + ia], q11 = lv[rb + ib];
  float t00[3], t10[3], t01[3], t11[3];
  decode_texel(q00, t00);
  decode_texel(q10, t10);
  decode_texel(q01, t01);
  decode_texel(q11, t11);
  for (int c = 0; c < 3; ++c)
    out[c] = __builtin_fmaf(t.w[3], t11[c], __builtin_fmaf(t.w[2], t01[c], __builtin_fmaf(t.w[1], t10[c], t.w[0] * t00[c])));
}

// The per-pixel inputs of the specular lookup (cacheApply.frag:139-154) and
// the textureLod level selection of LINEAR_MIPMAP_LINEAR (GL 4.5 section
// 8.14.3, the sampler bound at renderer.cpp:1070-1071): level 0 at lod <= 0,
// the last level at lod >= q, otherwise floor(lod) and floor(lod) + 1
// blended by frac(lod).
struct SpecPixel {
  float lookup[2];  // specEnvLookupCoord
  float lod;        // specularEnvmapLod
  float color[3];   // specularColor
  Tap t1, t2;       // the two levels (t2 unused when blend = 0)
  float blend;
  bool two;
};

constexpr float kInvPi = 1.0f / kPi;

// UnpackNormal16I (utils.glsl:43-60) for shading: the angle in revolutions
// (a / 65536) through the hardware sine and cosine, normalized with one
// reciprocal square root.  Within ~1e-6 of the exactly rounded unpack, far
// below the apply's tolerance; the exact unpack stays where normals are
// repacked bit-exactly (the RSM mip chain).
__device__ __forceinline__ f3 shading_normal(int16_t a, int16_t b) {
  const float rev = static_cast<float>(a) * (1.0f / 65536.0f);
  const float py = static_cast<float>(b) * (1.0f / 32768.0f);
  const float sin_phi = __builtin_sqrtf(fmaxf(0.0f, 1.0f - py * py));
  const float x = __builtin_amdgcn_cosf(rev) * sin_phi, y = __builtin_amdgcn_sinf(rev) * sin_phi, z = py;
  const float inv = __builtin_amdgcn_rsqf(__builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x)));
  return make3(x * inv, y * inv, z * inv);
}

// SampleCacheIrradiance, lightcache.glsl:137-183, on an entry already in
// registers (q[k] = float4 k of the std430 LightCacheEntry).
__device__ __forceinline__ void cache_irradiance(const ApplyArgs& a, const float4* q, f3 n, float* irr) {
  // {SH1neg1, SH00_r}, {SH10, SH00_g}, {SH1pos1, SH00_b}
  const float CL0 = a.shc.cos_lobe0, CL1 = a.shc.cos_lobe1;
  // Terms accumulate in the reference's order, each as one fused multiply-add.
  const float ny = -(CL1 * n.y), nz = CL1 * n.z, nx = -(CL1 * n.x);
  float v0 = q[1].w * CL0, v1 = q[2].w * CL0, v2 = q[3].w * CL0;
  v0 = __builtin_fmaf(q[1].x, ny, v0);
  v1 = __builtin_fmaf(q[1].y, ny, v1);
  v2 = __builtin_fmaf(q[1].z, ny, v2);
  v0 = __builtin_fmaf(q[2].x, nz, v0);
  v1 = __builtin_fmaf(q[2].y, nz, v1);
  v2 = __builtin_fmaf(q[2].z, nz, v2);
  v0 = __builtin_fmaf(q[3].x, nx, v0);
  v1 = __builtin_fmaf(q[3].y, nx, v1);
  v2 = __builtin_fmaf(q[3].z, nx, v2);
  if (a.sh2) {
    // {SH2neg2, SH20_r}, {SH2neg1, SH20_g}, {SH2pos1, SH20_b}, {SH2pos2, -}
    const float CL2 = a.shc.cos_lobe2n2_p1_n1, CL20 = a.shc.cos_lobe20, CL22 = a.shc.cos_lobe2p2;
    const float fxy = -(CL2 * n.x * n.y), fyz = CL2 * n.y * n.z, fxz = CL2 * n.x * n.z;
    const float f20 = CL20 * (n.z * n.z * 3.0f - 1.0f), f22 = CL22 * (n.x * n.x - n.y * n.y);
    v0 = __builtin_fmaf(q[4].x, fxy, v0);
    v1 = __builtin_fmaf(q[4].y, fxy, v1);
    v2 = __builtin_fmaf(q[4].z, fxy, v2);
    v0 = __builtin_fmaf(q[5].x, fyz, v0);
    v1 = __builtin_fmaf(q[5].y, fyz, v1);
    v2 = __builtin_fmaf(q[5].z, fyz, v2);
    v0 = __builtin_fmaf(q[4].w, f20, v0);
    v1 = __builtin_fmaf(q[5].w, f20, v1);
    v2 = __builtin_fmaf(q[6].w, f20, v2);
    v0 = __builtin_fmaf(q[6].x, fxz, v0);
    v1 = __builtin_fmaf(q[6].y, fxz, v1);
    v2 = __builtin_fmaf(q[6].z, fxz, v2);
    v0 = __builtin_fmaf(q[7].x, f22, v0);
    v1 = __builtin_fmaf(q[7].y, f22, v1);
    v2 = __builtin_fmaf(q[7].z, f22, v2);
  }
  irr[0] = fmaxf(v0, 0.0f);
  irr[1] = fmaxf(v1, 0.0f);
  irr[2] = fmaxf(v2, 0.0f);
}

// ComputeLightingFromCaches, cacheApply.frag:28-118.  The lanes of a wave
// (16x4 pixels) mostly share their 8 caches (2.1 distinct addresses per
// corner on a 1080p frame), and 8 x 7 per-lane 16-byte loads keep the
// texture address and data units busy.  So per corner the wave's first two
// distinct entries are fetched once - lanes 0-27 and 32-59 load one float
// each, the 8 corners' loads in flight together - and staged in LDS, where
// their lanes read them; lanes with a third or later address load their own
// entry.  Per lane the arithmetic and its order are those of one fetch per
// pixel.  Every lane of the wave must call this (the staging is
// cooperative); `active` = false lanes add nothing.
// Two neighbouring CAV cells (dword alignment suffices for the 8-byte load).
struct __attribute__((aligned(4))) CellPair {
  uint32_t a, b;
};

template <bool kSpec>
__device__ __forceinline__ void lighting(const ApplyArgs& a, bool active, f3 wp, f3 n, int c, const float* diffuse,
                                         uint32_t live, const SpecPixel& sp, float* stage, float* out) {
  const drv_cascade& k = a.vol.cascades[c];
  const float acx = (wp.x - k.min[0]) / k.world_voxel_size;
  const float acy = (wp.y - k.min[1]) / k.world_voxel_size;
  const float acz = (wp.z - k.min[2]) / k.world_voxel_size;
  const int ax = to_int(acx), ay = to_int(acy), az = to_int(acz);
  const float ix = acx - static_cast<float>(ax), iy = acy - static_cast<float>(ay), iz = acz - static_cast<float>(az);
  const float jx = 1.0f - ix, jy = 1.0f - iy, jz = 1.0f - iz;
  const float w[8] = {jx * jy * jz, ix * jy * jz, jx * iy * jz, ix * iy * jz,
                      jx * jy * iz, ix * jy * iz, jx * iy * iz, ix * iy * iz};
  const int R = a.R, CR = a.C * a.R;
  uint32_t addr[8];
  // Offsets x fastest: (0,0,0),(1,0,0),(0,1,0),(1,1,0),(0,0,1),(1,0,1),
  // (0,1,1),(1,1,1).  The two cells of a pair (x, x + 1) are neighbours in
  // the CAV row, so each pair is one 8-byte load (4-byte aligned) when both
  // lie in the row; texelFetch outside the CAV reads 0 -> address 0xFFFFFFFF.
  const int x0 = ax + R * c;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int y = ay + (p & 1), z = az + (p >> 1);
    const bool row = active && y >= 0 && z >= 0 && y < R && z < R;
    uint32_t v0 = 0u, v1 = 0u;
    if (row) {
      const uint32_t* r = a.cav + (static_cast<uint32_t>(z) * R + y) * CR;
      if (x0 >= 0 && x0 + 1 < CR) {
        const CellPair q = *reinterpret_cast<const CellPair*>(r + x0);
        v0 = q.a;
        v1 = q.b;
      } else {
        if (x0 >= 0 && x0 < CR) v0 = r[x0];
        if (x0 + 1 >= 0 && x0 + 1 < CR) v1 = r[x0 + 1];
      }
    }
    addr[2 * p] = v0 - 1u;
    addr[2 * p + 1] = v1 - 1u;
  }
  const int nq = a.sh2 ? 8 : 4;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  float spec[3] = {0.0f, 0.0f, 0.0f};
  if (kSpec) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (addr[i] < live) {
        // cacheApply.frag:102-106
        const uint32_t cx = addr[i] & (a.per_row - 1u), cy = addr[i] >> a.per_row_log2;
        float smp[3];
        tap_cache(a, sp.t1, cx, cy, smp);
        if (sp.two) {
          float s2[3];
          tap_cache(a, sp.t2, cx, cy, s2);
          for (int k = 0; k < 3; ++k) smp[k] = __builtin_fmaf(sp.blend, s2[k], (1.0f - sp.blend) * smp[k]);
        }
        for (int k = 0; k < 3; ++k) spec[k] = __builtin_fmaf(smp[k], w[i], spec[k]);
      }
    }
  }
  // An address past the live caches (unclaimed cell, overflow) adds zero.
  const uint32_t lane = __lane_id();
  const uint32_t sub = lane & 31u;
  const uint32_t nf = static_cast<uint32_t>(4 * (nq - 1));  // entry floats 4 .. 4 + nf - 1 are read
  uint32_t from[8];  // 0: first address, 1: second, 2: own load, 3: none
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool ok = addr[i] < live;
    uint64_t pend = __ballot(ok);
    uint32_t ua = 0u, ub = 0u;
    bool in_a = false, in_b = false, have = false;
    if (pend != 0ull) {
      ua = static_cast<uint32_t>(
          __builtin_amdgcn_readlane(static_cast<int>(addr[i]), static_cast<int>(__builtin_ctzll(pend))));
      in_a = ok && addr[i] == ua;
      pend &= ~__ballot(in_a);
      have = lane < 32u;
    }
    if (pend != 0ull) {
      ub = static_cast<uint32_t>(
          __builtin_amdgcn_readlane(static_cast<int>(addr[i]), static_cast<int>(__builtin_ctzll(pend))));
      in_b = ok && !in_a && addr[i] == ub;
      have = true;
    }
    from[i] = in_a ? 0u : in_b ? 1u : ok ? 2u : 3u;
    const uint32_t u = lane < 32u ? ua : ub;
    const bool load = have && sub < nf;
    stage[i * 64 + lane] = load ? a.caches[static_cast<size_t>(u) * a.stride + 4u + sub] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (from[i] == 3u) continue;
    // q[4..7] (the L2 bands) are loaded and read only with sh2.
    float4 q[8];
    if (from[i] == 2u) {
      const float4* e4 = reinterpret_cast<const float4*>(a.caches + static_cast<size_t>(addr[i]) * a.stride);
#pragma unroll
      for (int j = 1; j < 4; ++j) q[j] = e4[j];
      if (a.sh2) {
#pragma unroll
        for (int j = 4; j < 8; ++j) q[j] = e4[j];
      }
    } else {
      const float4* st = reinterpret_cast<const float4*>(stage + i * 64 + 32 * from[i]);
#pragma unroll
      for (int j = 1; j < 4; ++j) q[j] = st[j - 1];
      if (a.sh2) {
#pragma unroll
        for (int j = 4; j < 8; ++j) q[j] = st[j - 1];
      }
    }
    float irr[3];
    cache_irradiance(a, q, n, irr);
    acc[0] = __builtin_fmaf(irr[0], w[i], acc[0]);
    acc[1] = __builtin_fmaf(irr[1], w[i], acc[1]);
    acc[2] = __builtin_fmaf(irr[2], w[i], acc[2]);
  }
  __builtin_amdgcn_wave_barrier();  // the stage is rewritten by the next call
  out[0] = acc[0] * diffuse[0] * kInvPi;
  out[1] = acc[1] * diffuse[1] * kInvPi;
  out[2] = acc[2] * diffuse[2] * kInvPi;
  if (kSpec) {  // cacheApply.frag:116
    out[0] = out[0] + spec[0] * sp.color[0];
    out[1] = out[1] + spec[1] * sp.color[1];
    out[2] = out[2] + spec[2] * sp.color[2];
  }
}

// The pixel's G-buffer texels are all loaded before the discard test and
// the sRGB table sits in LDS, so a wave waits for one G-buffer round trip.
template <bool kSpec>
__global__ __launch_bounds__(256) void k_apply(ApplyArgs a) {
  __shared__ float lut[256];
  __shared__ __align__(16) float stage[4][8 * 64];  // per wave: the shared entries of each corner
  lut[threadIdx.x] = a.srgb_lut[threadIdx.x];
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const bool inside = x < a.width && y < a.height;
  const size_t p = inside ? static_cast<size_t>(y) * a.width + x : 0;
  const float d = a.depth[p];
  const uint32_t nrm = reinterpret_cast<const uint32_t*>(a.normal)[p];  // RG16I texel
  const uint32_t al0 = a.albedo[3 * p], al1 = a.albedo[3 * p + 1], al2 = a.albedo[3 * p + 2];
  uint32_t rm0 = 255u, rm1 = 0u;
  if (kSpec && a.roughness_metallic) {
    rm0 = a.roughness_metallic[2 * p];
    rm1 = a.roughness_metallic[2 * p + 1];
  }
  __syncthreads();
  // cacheApply.frag:127-131: discard.  Discarded and outside lanes still run
  // the (cooperative) cache fetches, inactive.
  const bool valid = inside && !(d < 0.00001f);
  const uint32_t live = a.counter->live;
  const f3 wp = unproject_pixel(a.inv_vp, a.width, a.height, x, y, d);
  const int c = cascade_of(a.vol, a.C, wp);
  const f3 n = shading_normal(static_cast<int16_t>(nrm & 0xFFFFu), static_cast<int16_t>(nrm >> 16));
  float diffuse[3] = {lut[al0], lut[al1], lut[al2]};
  SpecPixel sp{};
  if (kSpec) {
    // cacheApply.frag:139-154 with lightingfunctions.glsl:26-44 and
    // utils.glsl:193-201 (no roughness / metallic texture: roughness 1,
    // metallic 0).
    const float rough = static_cast<float>(rm0) / 255.0f;
    const float metal = static_cast<float>(rm1) / 255.0f;
    for (int k = 0; k < 3; ++k) {
      const float base = diffuse[k];
      diffuse[k] = base * (1.0f - metal) + 0.02f * metal;
      sp.color[k] = 0.04f * (1.0f - metal) + base * metal;
    }
    const float rsq = rough * rough;
    const float blinn = 2.0f / (rsq * rsq + 0.0005f);
    const float fs = static_cast<float>(a.S);
    sp.lod = fmaxf(0.0f, log2f(fs * fs / (1.0f + blinn)) * 0.5f);
    const float max_half = 0.5f / (exp2f(-ceilf(sp.lod)) * fs);
    const LocalFrame fr = local_view_space(a.cam, wp);
    const f3 cn = to_local(fr, n.x, n.y, n.z);
    hemi_projection(cn.x, cn.y, cn.z, sp.lookup);
    for (int k = 0; k < 2; ++k) sp.lookup[k] = fminf(fmaxf(sp.lookup[k], max_half), 1.0f - max_half);
    const int q = a.levels - 1;
    sp.two = false;
    sp.blend = 0.0f;
    if (sp.lod <= 0.0f) {
      sp.t1 = make_tap(a, 0, sp.lookup);
    } else if (sp.lod >= static_cast<float>(q)) {
      sp.t1 = make_tap(a, q, sp.lookup);
    } else {
      const float d1 = floorf(sp.lod);
      sp.blend = sp.lod - d1;
      sp.t1 = make_tap(a, to_int(d1), sp.lookup);
      sp.t2 = make_tap(a, to_int(d1) + 1, sp.lookup);
      sp.two = true;
    }
  }
  float* st = stage[threadIdx.x >> 6];
  float col[3];
  lighting<kSpec>(a, valid, wp, n, c, diffuse, live, sp, st, col);
  if (a.transitions) {  // cacheApply.frag:172-185
    const float t = valid ? cascade_transition(a.vol, c, wp) : 0.0f;
    const bool two = t > 0.0f && c < a.C - 1;
    if (__ballot(two) != 0ull) {
      float col2[3];
      lighting<kSpec>(a, two, wp, n, two ? c + 1 : c, diffuse, live, sp, st, col2);
      if (two) {
        col[0] = col[0] * (1.0f - t) + col2[0] * t;
        col[1] = col[1] * (1.0f - t) + col2[1] * t;
        col[2] = col[2] * (1.0f - t) + col2[2] * t;
      }
    }
  }
  if (!inside) return;
  float* o = a.out + 3 * p;
  if (!valid) {
    if (!a.accumulate) o[0] = o[1] = o[2] = 0.0f;
    return;
  }
  if (a.accumulate) {
    o[0] += col[0];
    o[1] += col[1];
    o[2] += col[2];
  } else {
    o[0] = col[0];
    o[1] = col[1];
    o[2] = col[2];
  }
}

}  // namespace

void launch_apply(const drv_config& cfg, const drv_volume& vol, const drv_frame& frame, const drv_sh_constants& shc,
                  const drv_gbuffer& gb, const uint32_t* cav, const float* caches, const Counter* counter,
                  const float* srgb_lut, const SpecAtlas* spec, float* out_rgb, int accumulate, hipStream_t s) {
  ApplyArgs a;
  a.vol = vol;
  for (int i = 0; i < 16; ++i) a.inv_vp[i] = frame.inverse_view_projection[i];
  a.shc = shc;
  a.depth = gb.depth;
  a.normal = gb.normal;
  a.albedo = gb.albedo;
  a.cav = cav;
  a.caches = caches;
  a.counter = counter;
  a.srgb_lut = srgb_lut;
  a.out = out_rgb;
  a.width = cfg.width;
  a.height = cfg.height;
  a.R = cfg.cav_resolution;
  a.C = cfg.cav_cascades;
  a.stride = entry_stride(cfg.sh_bands);
  a.sh2 = cfg.sh_bands >= 2 ? 1 : 0;
  a.transitions = cfg.cav_transition_size > 0.0f ? 1 : 0;
  a.accumulate = accumulate;
  a.roughness_metallic = gb.roughness_metallic;
  a.atlas = nullptr;
  for (int l = 0; l < kSpecMaxLevels; ++l) a.level_offset[l] = 0;
  a.atlas_width = a.per_row = a.per_row_log2 = 0;
  a.S = a.levels = 0;
  for (int i = 0; i < 3; ++i) a.cam[i] = frame.camera_position[i];
  if (spec) {
    a.atlas = spec->atlas;
    for (int l = 0; l < kSpecMaxLevels; ++l) a.level_offset[l] = spec->level_offset[l];
    a.atlas_width = spec->width;
    a.per_row = spec->per_row;
    a.per_row_log2 = static_cast<uint32_t>(__builtin_ctz(spec->per_row));
    a.S = spec->S;
    a.levels = spec->levels;
  }
  const dim3 grid((cfg.width + 15) / 16, (cfg.height + 15) / 16);
  if (spec)
    hipLaunchKernelGGL(k_apply<true>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_apply<false>, grid, dim3(256), 0, s, a);
}

}  // namespace drvk
