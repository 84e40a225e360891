// Indirect specular (INDIRECT_SPECULAR with DIRECT_SPECULAR_MAP_WRITE, the
// reference default, rendering/renderer.cpp:45-49): every cache keeps an
// S x S hemisphere env-map of the VAL radiance it reflects, in an
// R11G11B10F atlas with a mip chain.
//
//   k_specular_gather    cacheLightingRSM.comp:87-99, 240-243, 281-335: per
//                        (cache, VAL) the cache-local half vector picks a
//                        texel of the cache's map and the reflected radiance
//                        is added there;
//   k_specular_finalize  renderer.cpp:994-1045 (PrepareSpecularEnvmaps):
//                        shader/specularenvmap_mipmap.frag for every level,
//                        then shader/specularenvmap_fillholes.frag from the
//                        hole-fill level down to level 1.
//
// MI355X design.  A wave owns one cache at a time and keeps its map in LDS
// in FP32 (the reference read-modify-writes the R11G11B10F image once per
// VAL).  Lane l takes the Morton-consecutive VALs l*n/64 ... (l+1)*n/64 - 1
// of one RSM block, read from a lane-interleaved copy so every step's loads
// coalesce.  Neighbouring VALs of a block mostly reflect into the same
// texel, so a lane sums runs in registers and adds a run into the LDS map
// only when its texel changes: a quarter of the steps add, and the 64
// blocks of a step rarely hit the same texel (LDS float atomics to one
// address serialize, which made per-VAL adds four times slower).  The
// light's FP32 sums go to a
// per-cache work buffer that accumulates over the lights of a frame; the
// finalize kernel (one workgroup per cache) quantizes, builds the mips and
// fills holes in LDS and writes every level of the cache's atlas region
// once.
//
// Texel index S: ivec2(HemisphericalProjection(h) * S + 0.5) reaches S for
// half vectors near the horizon, so the reference's image store lands in
// the first column / row of the next cache in the atlas (or outside the
// image, where it is dropped).  The map here is (S+1) x (S+1): the extra
// column / row / corner are routed to the atlas neighbours by the finalize
// kernel, which adds them after the neighbour's own sum (the reference's
// unsynchronised read-modify-write of that texel has no defined order).
#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_specular.hpp"

namespace drvk {
namespace {

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SpecGatherArgs {
  float cam[3];
  const float* caches;
  int stride;
  const Counter* counter;
  const Vpl* vpls;   // lane-interleaved: lane l, step i reads record i * 64 + l = VPL l * n / 64 + i
  uint32_t n_vpl;
  const float* vis;  // [block][vis_pitch] or null (unshadowed)
  uint32_t vis_pitch;
  uint32_t interval;        // VPLs per shadow block, a power of two
  uint32_t interval_log2;
  int S;
  float* work;   // [cache][S * S][3]
  float* spill;  // [cache][2S + 1][3]: column x = S (y = 0..S-1), row y = S (x = 0..S-1), corner
  int first;     // 1: the first light of the cache list overwrites the work sums
};

// ivec2(p * S + 0.5) clamped to [0, S] (NaN -> 0): the clamp runs on the
// float, so the conversion is always in range.
__device__ __forceinline__ int texel_index(float p, int S) {
  const float fs = static_cast<float>(S);
  return static_cast<int>(fminf(fmaxf(__builtin_fmaf(p, fs, 0.5f), 0.0f), fs));
}

// Dynamic LDS: 4 maps of (S+1)^2 RGB floats.  Two waves share a cache:
// lane l of wave half h walks the second index range h of the lane's VALs
// (steps h * (per_lane / 2) ... of the lane-interleaved copy) into its own map,
// and the pair writes map 0 + map 1 out (a fixed order), so twice as many
// waves, each half as long, balance the SIMDs.
__global__ __launch_bounds__(256) void k_specular_gather(SpecGatherArgs a) {
  extern __shared__ float lds[];
  const int S = a.S, S1 = S + 1;
  const uint32_t map_floats = static_cast<uint32_t>(S1 * S1 * 3);
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t pair = w >> 1, half = w & 1u;
  float* map = lds + w * map_floats;
  const float* map0 = lds + (2u * pair) * map_floats;
  const float* map1 = map0 + map_floats;
  const uint32_t live = a.counter->live;
  const uint32_t per_lane = a.n_vpl / 64u;  // n_vpl = R^2 with R >= 8
  const uint32_t mid = per_lane / 2u;
  const uint32_t i0 = half ? mid : 0u, i1 = half ? per_lane : mid;
  for (uint32_t base = blockIdx.x * 2u; base < live; base += gridDim.x * 2u) {  // workgroup-uniform
    const uint32_t c = base + pair;
    const bool active = c < live;
    for (uint32_t i = lane; i < map_floats; i += 64u) map[i] = 0.0f;
    wave_lds_sync();
    if (active) {
      const float* e = a.caches + static_cast<size_t>(c) * a.stride;
      const f3 wp = make3(e[0], e[1], e[2]);
      const LocalFrame fr = local_view_space(a.cam, wp);
      const uint32_t m0 = lane * per_lane;
      float shadowing = 1.0f;
      int run_tex = -1;
      float run_r = 0.0f, run_g = 0.0f, run_b = 0.0f;
      for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t m = m0 + i;
        const Vpl v = a.vpls[static_cast<size_t>(i) * 64u + lane];
        // The block's shadowing changes every `interval` VPLs.
        if (a.vis && (i == i0 || (m & (a.interval - 1u)) == 0u))
          shadowing = a.vis[static_cast<size_t>(m >> a.interval_log2) * a.vis_pitch + c];
        // radianceIn, cacheLightingRSM.comp:246-264 (the SH gather's expressions)
        const float tx0 = v.px - wp.x, ty0 = v.py - wp.y, tz0 = v.pz - wp.z;
        const float d2 = __builtin_fmaf(tz0, tz0, __builtin_fmaf(ty0, ty0, tx0 * tx0));
        const float inv = __builtin_amdgcn_rsqf(d2);
        const float tx = tx0 * inv, ty = ty0 * inv, tz = tz0 * inv;
        const float nd = -__builtin_fmaf(v.nz, tz, __builtin_fmaf(v.ny, ty, v.nx * tx));
        const float f = fminf(1.0f, fmaxf(0.0f, nd)) * shadowing * __builtin_amdgcn_rcpf(d2 + v.disc_area);
        // cacheLightingRSM.comp:281-335
        const f3 l = to_local(fr, tx, ty, tz);
        const float hz0 = l.z + 1.0f;
        const float hi = __builtin_amdgcn_rsqf(__builtin_fmaf(hz0, hz0, __builtin_fmaf(l.y, l.y, l.x * l.x)));
        const float hx = l.x * hi, hy = l.y * hi, hz = hz0 * hi;
        float proj[2];
        hemi_projection(hx, hy, hz, proj);
        const int ix = texel_index(proj[0], S), iy = texel_index(proj[1], S);
        const float k = f * fminf(1.0f, fmaxf(0.0f, __builtin_fmaf(hz, l.z, __builtin_fmaf(hy, l.y, hx * l.x))));
        // Run accumulation: consecutive VALs of a lane mostly hit the same
        // texel, so the lane sums in registers and adds the run into the map
        // when the texel changes (zero contributions change nothing).
        const float cr = v.fr * k, cg = v.fg * k, cb = v.fb * k;
        if (!(cr == 0.0f && cg == 0.0f && cb == 0.0f)) {
          const int tex = iy * S1 + ix;
          if (tex != run_tex) {
            if (run_tex >= 0) {
              float* t = map + run_tex * 3;
              atomicAdd(t, run_r);
              atomicAdd(t + 1, run_g);
              atomicAdd(t + 2, run_b);
            }
            run_tex = tex;
            run_r = cr;
            run_g = cg;
            run_b = cb;
          } else {
            run_r += cr;
            run_g += cg;
            run_b += cb;
          }
        }
      }
      if (run_tex >= 0) {
        float* t = map + run_tex * 3;
        atomicAdd(t, run_r);
        atomicAdd(t + 1, run_g);
        atomicAdd(t + 2, run_b);
      }
    }
    __syncthreads();  // both maps of every pair are complete
    if (active) {
      float* wk = a.work + static_cast<size_t>(c) * S * S * 3;
      float* sp = a.spill + static_cast<size_t>(c) * (2 * S + 1) * 3;
      for (uint32_t i = half * 64u + lane; i < map_floats; i += 128u) {
        const int tex = static_cast<int>(i / 3u), ch = static_cast<int>(i % 3u);
        const int x = tex % S1, y = tex / S1;
        const float val = map0[i] + map1[i];
        if (x < S && y < S) {
          float* o = wk + (y * S + x) * 3 + ch;
          *o = a.first ? val : *o + val;
        } else {
          const int k = (x == S && y < S) ? y : (y == S && x < S) ? S + x : 2 * S;
          sp[k * 3 + ch] = val;
        }
      }
    }
    __syncthreads();  // the maps are zeroed again for the next caches
  }
}

struct SpecFinalArgs {
  const Counter* counter;
  int S, levels, fill;
  uint32_t per_row;
  uint32_t width;
  uint64_t level_offset[kSpecMaxLevels];
  float norm;
  float* work;
  const float* spill;
  uint32_t* atlas;
};

__device__ __forceinline__ float quantize(float v, int ch) {
  const int m = ch == 2 ? 5 : 6;
  return uf_decode(uf_encode(v, m), m);
}

// Dynamic LDS: the decoded RGB texels of every level of one cache.
__global__ __launch_bounds__(256) void k_specular_finalize(SpecFinalArgs a) {
  extern __shared__ float lv[];
  const int S = a.S, SS = S * S;
  const uint32_t live = a.counter->live;
  const uint32_t N = a.per_row;
  int lofs[kSpecMaxLevels];
  {
    int o = 0;
#pragma unroll
    for (int l = 0; l < kSpecMaxLevels; ++l) {
      lofs[l] = o;
      const int sl = S >> l;
      o += sl * sl;
    }
  }
  const int tid = static_cast<int>(threadIdx.x);
  for (uint32_t c = blockIdx.x; c < live; c += gridDim.x) {
    const uint32_t cx = c % N, cy = c / N;
    const bool has_l = cx > 0u, has_u = cy > 0u;
    float* wk = a.work + static_cast<size_t>(c) * SS * 3;
    const size_t sp_stride = static_cast<size_t>(2 * S + 1) * 3;
    // Level 0: own sums plus the spill of the left, upper and upper-left
    // atlas neighbours, scaled by the base normalization, quantized.
    for (int t = tid; t < SS; t += 256) {
      const int x = t % S, y = t / S;
      float v[3] = {wk[t * 3], wk[t * 3 + 1], wk[t * 3 + 2]};
      const bool from_l = x == 0 && has_l, from_u = y == 0 && has_u;
      if (from_l) {
        const float* s = a.spill + static_cast<size_t>(c - 1u) * sp_stride + y * 3;
        for (int ch = 0; ch < 3; ++ch) v[ch] += s[ch];
      }
      if (from_u) {
        const float* s = a.spill + static_cast<size_t>(c - N) * sp_stride + (S + x) * 3;
        for (int ch = 0; ch < 3; ++ch) v[ch] += s[ch];
      }
      if (from_l && from_u) {
        const float* s = a.spill + static_cast<size_t>(c - N - 1u) * sp_stride + 2 * S * 3;
        for (int ch = 0; ch < 3; ++ch) v[ch] += s[ch];
      }
      if (from_l || from_u)
        for (int ch = 0; ch < 3; ++ch) wk[t * 3 + ch] = v[ch];
      for (int ch = 0; ch < 3; ++ch) lv[t * 3 + ch] = quantize(v[ch] * a.norm, ch);
    }
    __syncthreads();
    // specularenvmap_mipmap.frag: a bilinear tap at the centre of each
    // destination texel = mean of its 2x2 source texels.
    for (int l = 1; l < a.levels; ++l) {
      const int sl = S >> l, sp = S >> (l - 1);
      const float* src = lv + lofs[l - 1] * 3;
      float* dst = lv + lofs[l] * 3;
      for (int t = tid; t < sl * sl; t += 256) {
        const int x = t % sl, y = t / sl;
        const int i00 = ((2 * y) * sp + 2 * x) * 3, i10 = i00 + 3, i01 = i00 + sp * 3, i11 = i01 + 3;
        for (int ch = 0; ch < 3; ++ch)
          dst[t * 3 + ch] =
              quantize(((src[i00 + ch] + src[i10 + ch]) + (src[i01 + ch] + src[i11 + ch])) * 0.25f, ch);
      }
      __syncthreads();
    }
    // specularenvmap_fillholes.frag, levels fill .. 1.
    for (int i = a.fill; i >= 1; --i) {
      const int si = S >> i, sd = S >> (i - 1);
      const float* src = lv + lofs[i] * 3;
      float* dst = lv + lofs[i - 1] * 3;
      for (int t = tid; t < si * si; t += 256) {
        const int x = t % si, y = t / si;
        const float s[3] = {src[t * 3], src[t * 3 + 1], src[t * 3 + 2]};
        if ((s[0] + s[1]) + s[2] < 0.0001f) continue;  // discard
        int di[4];
        di[0] = ((2 * y) * sd + 2 * x) * 3;
        di[1] = di[0] + 3;
        di[2] = di[0] + sd * 3;
        di[3] = di[2] + 3;
        float d[4][3];
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 4; ++k) {
          for (int ch = 0; ch < 3; ++ch) d[k][ch] = dst[di[k] + ch];
          if ((d[k][0] + d[k][1]) + d[k][2] == 0.0f)
            for (int ch = 0; ch < 3; ++ch) d[k][ch] = s[ch];
          for (int ch = 0; ch < 3; ++ch) sum[ch] += d[k][ch];
        }
        float en[3];
        for (int ch = 0; ch < 3; ++ch) en[ch] = 4.0f * s[ch] / (sum[ch] + 0.00001f);
        for (int k = 0; k < 4; ++k)
          for (int ch = 0; ch < 3; ++ch) dst[di[k] + ch] = quantize(d[k][ch] * en[ch], ch);
      }
      __syncthreads();
    }
    // Every level of the cache's atlas region.
    for (int l = 0; l < a.levels; ++l) {
      const int sl = S >> l;
      const uint64_t wl = a.width >> l;
      uint32_t* out = a.atlas + a.level_offset[l];
      const float* src = lv + lofs[l] * 3;
      for (int t = tid; t < sl * sl; t += 256) {
        const int x = t % sl, y = t / sl;
        const uint64_t px = static_cast<uint64_t>(cx) * sl + x, py = static_cast<uint64_t>(cy) * sl + y;
        out[py * wl + px] = pack_r11g11b10(src[t * 3], src[t * 3 + 1], src[t * 3 + 2]);
      }
    }
    __syncthreads();  // lv is reused by the next cache
  }
}

}  // namespace

void launch_specular(const float* cam, const float* caches, int stride, const Counter* counter, const Vpl* vpls,
                     uint32_t n_vpl, const float* vis, uint32_t vis_pitch, uint32_t interval, const SpecAtlas& at,
                     int first, int grid, hipStream_t s) {
  SpecGatherArgs g;
  for (int i = 0; i < 3; ++i) g.cam[i] = cam[i];
  g.caches = caches;
  g.stride = stride;
  g.counter = counter;
  g.vpls = vpls;
  g.n_vpl = n_vpl;
  g.vis = vis;
  g.vis_pitch = vis_pitch;
  g.interval = interval;
  g.interval_log2 = static_cast<uint32_t>(__builtin_ctz(interval));
  g.S = at.S;
  g.work = at.work;
  g.spill = at.spill;
  g.first = first;
  const size_t glds = static_cast<size_t>(4) * (at.S + 1) * (at.S + 1) * 3 * sizeof(float);
  // More workgroups than fit at once (each takes two caches a round): the
  // hardware deals the last ones out as others finish.
  hipLaunchKernelGGL(k_specular_gather, dim3(4 * grid), dim3(256), glds, s, g);

  SpecFinalArgs f;
  f.counter = counter;
  f.S = at.S;
  f.levels = at.levels;
  f.fill = at.fill;
  f.per_row = at.per_row;
  f.width = at.width;
  for (int l = 0; l < kSpecMaxLevels; ++l) f.level_offset[l] = at.level_offset[l];
  f.norm = specular_base_normalization(at.S);
  f.work = at.work;
  f.spill = at.spill;
  f.atlas = at.atlas;
  size_t texels = 0;
  for (int l = 0; l < at.levels; ++l) texels += static_cast<size_t>(at.S >> l) * (at.S >> l);
  hipLaunchKernelGGL(k_specular_finalize, dim3(grid), dim3(256), texels * 3 * sizeof(float), s, f);
}

}  // namespace drvk
