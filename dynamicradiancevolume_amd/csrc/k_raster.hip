// The gather's inputs rendered on the device: a depth-tested triangle raster
// with the G-buffer fill (shader/fillgbuffer.frag) and the RSM fill and
// mip chain (shader/fillrsm.frag:24-62, shader/downsamplersm.frag:15-33,
// driven by rendering/renderer.cpp:1300-1339).
//
// Raster rules (the same as the host producer in csrc/inputs, which the
// tests compare against): triangles are projected and clipped against the
// near plane in FP64, pixel centres are sampled, both windings are accepted,
// a fragment with reverse-Z depth z in (0, 1] is kept, and of all fragments
// of a pixel the one with the largest depth wins, the lowest triangle index
// among equal depths.  On the GPU that is a two-pass visibility buffer: pass
// 1 takes the 64-bit atomic max of the depth bits (non-negative doubles
// order like their bit patterns), pass 2 the atomic min of the index among
// the fragments that reached that depth; a resolve pass then shades each
// pixel from its winning triangle.  One thread rasterizes a small triangle;
// triangles with a large pixel box are listed and rasterized by a workgroup.
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

constexpr float kPiF = 3.14159265358979f;   // shader/utils.glsl:1
constexpr float kPi2F = 6.28318530717958f;  // shader/utils.glsl:2

struct ScreenTri {
  double x[3], y[3], z[3];
};

// Projection and near-plane clip (w - z >= 0 for the reverse-Z projection);
// returns the number of screen triangles (0..2).
__device__ int project_clip(const float* p, const double* vp, int w, int h, ScreenTri* out) {
  double clip[3][4];
  for (int v = 0; v < 3; ++v)
    for (int r = 0; r < 4; ++r)
      clip[v][r] = vp[4 * r + 0] * p[3 * v + 0] + vp[4 * r + 1] * p[3 * v + 1] + vp[4 * r + 2] * p[3 * v + 2] +
                   vp[4 * r + 3];
  double poly[4][4];
  int n = 0;
  for (int v = 0; v < 3; ++v) {
    const double* a = clip[v];
    const double* b = clip[(v + 1) % 3];
    const double da = a[3] - a[2], db = b[3] - b[2];
    if (da >= 0) {
      for (int k = 0; k < 4; ++k) poly[n][k] = a[k];
      ++n;
    }
    if ((da >= 0) != (db >= 0)) {
      const double t = da / (da - db);
      for (int k = 0; k < 4; ++k) poly[n][k] = a[k] + (b[k] - a[k]) * t;
      ++n;
    }
  }
  if (n < 3) return 0;
  double sx[4], sy[4], sz[4];
  for (int i = 0; i < n; ++i) {
    if (poly[i][3] <= 1e-12) return 0;
    sx[i] = (poly[i][0] / poly[i][3] * 0.5 + 0.5) * w;
    sy[i] = (poly[i][1] / poly[i][3] * 0.5 + 0.5) * h;
    sz[i] = poly[i][2] / poly[i][3];
  }
  int m = 0;
  for (int i = 1; i + 1 < n; ++i) {
    const int idx[3] = {0, i, i + 1};
    for (int k = 0; k < 3; ++k) {
      out[m].x[k] = sx[idx[k]];
      out[m].y[k] = sy[idx[k]];
      out[m].z[k] = sz[idx[k]];
    }
    ++m;
  }
  return m;
}

struct RasterArgs {
  const float* positions;  // [n_tri][9]
  uint32_t n_tri;
  double vp[16];
  int w, h;
  unsigned long long* zbits;  // [h][w] depth bits (0 = empty)
  uint32_t* tri;              // [h][w] winning triangle (0xffffffff = none)
  uint32_t* big;              // triangles whose screen box exceeds kBigPixels
  uint32_t* n_big;
};

// Triangles covering more pixel-box area than this are rasterized by a whole
// workgroup instead of one thread (a ground plane's few large triangles would
// otherwise serialize the pass).
constexpr uint32_t kBigPixels = 512;

// Screen-space setup of one clipped triangle piece: pixel box and edge data.
struct Setup {
  ScreenTri q;
  int px0, py0, bw, bh;
  double inv;
  bool ok;
};

__device__ __forceinline__ Setup setup_piece(const ScreenTri& q, int w, int h) {
  Setup st;
  st.q = q;
  st.ok = false;
  const double miny = fmin(fmin(q.y[0], q.y[1]), q.y[2]), maxy = fmax(fmax(q.y[0], q.y[1]), q.y[2]);
  const double minx = fmin(fmin(q.x[0], q.x[1]), q.x[2]), maxx = fmax(fmax(q.x[0], q.x[1]), q.x[2]);
  // Bounds in double first so that off-screen triangles cannot overflow int.
  const double fy0 = fmax(0.0, ceil(miny - 0.5)), fy1 = fmin(static_cast<double>(h - 1), floor(maxy - 0.5));
  const double fx0 = fmax(0.0, ceil(minx - 0.5)), fx1 = fmin(static_cast<double>(w - 1), floor(maxx - 0.5));
  if (!(fy0 <= fy1) || !(fx0 <= fx1)) return st;
  const double area = (q.x[1] - q.x[0]) * (q.y[2] - q.y[0]) - (q.x[2] - q.x[0]) * (q.y[1] - q.y[0]);
  if (area == 0.0) return st;
  st.inv = 1.0 / area;
  st.px0 = static_cast<int>(fx0);
  st.py0 = static_cast<int>(fy0);
  st.bw = static_cast<int>(fx1) - st.px0 + 1;
  st.bh = static_cast<int>(fy1) - st.py0 + 1;
  st.ok = true;
  return st;
}

// Visits pixels start, start + step, ... of the piece's box (row-major) and
// calls f(pixel index, depth) for every kept fragment.
template <typename F>
__device__ __forceinline__ void raster_piece(const Setup& st, int w, uint32_t start, uint32_t step, F&& f) {
  const ScreenTri& q = st.q;
  const uint32_t npix = static_cast<uint32_t>(st.bw) * static_cast<uint32_t>(st.bh);
  for (uint32_t k = start; k < npix; k += step) {
    const int px = st.px0 + static_cast<int>(k % static_cast<uint32_t>(st.bw));
    const int py = st.py0 + static_cast<int>(k / static_cast<uint32_t>(st.bw));
    const double cy = py + 0.5, cx = px + 0.5;
    const double e0 = ((q.x[2] - q.x[1]) * (cy - q.y[1]) - (q.y[2] - q.y[1]) * (cx - q.x[1])) * st.inv;
    const double e1 = ((q.x[0] - q.x[2]) * (cy - q.y[2]) - (q.y[0] - q.y[2]) * (cx - q.x[2])) * st.inv;
    const double e2 = 1.0 - e0 - e1;
    if (e0 < 0 || e1 < 0 || e2 < 0) continue;
    const double z = e0 * q.z[0] + e1 * q.z[1] + e2 * q.z[2];
    if (!(z > 0.0) || z > 1.0) continue;  // a fragment must beat the cleared depth 0
    f(static_cast<size_t>(py) * w + px, z);
  }
}

// Small triangles: one thread each.  Large ones are listed (first pass) or
// skipped (second pass) for the workgroup kernel.
template <typename F>
__device__ __forceinline__ void raster_small(const RasterArgs& a, uint32_t t, bool list_big, F&& f) {
  ScreenTri sc[2];
  const int m = project_clip(a.positions + 9 * static_cast<size_t>(t), a.vp, a.w, a.h, sc);
  Setup st[2];
  bool big = false;
  for (int s = 0; s < m; ++s) {
    st[s] = setup_piece(sc[s], a.w, a.h);
    if (st[s].ok && static_cast<uint32_t>(st[s].bw) * static_cast<uint32_t>(st[s].bh) > kBigPixels) big = true;
  }
  if (big) {
    if (list_big) a.big[atomicAdd(a.n_big, 1u)] = t;
    return;
  }
  for (int s = 0; s < m; ++s)
    if (st[s].ok) raster_piece(st[s], a.w, 0u, 1u, f);
}

// Large triangles: workgroup-strided pixel loops, persistent over the list.
template <typename F>
__device__ __forceinline__ void raster_big(const RasterArgs& a, F&& f) {
  const uint32_t n = *a.n_big;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t t = a.big[i];
    ScreenTri sc[2];
    const int m = project_clip(a.positions + 9 * static_cast<size_t>(t), a.vp, a.w, a.h, sc);
    for (int s = 0; s < m; ++s) {
      const Setup st = setup_piece(sc[s], a.w, a.h);
      if (st.ok) raster_piece(st, a.w, threadIdx.x, blockDim.x, [&](size_t px, double z) { f(t, px, z); });
    }
  }
}

__global__ __launch_bounds__(256) void k_raster_depth(RasterArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tri) return;
  raster_small(a, t, true, [&](size_t i, double z) {
    atomicMax(&a.zbits[i], static_cast<unsigned long long>(__double_as_longlong(z)));
  });
}

__global__ __launch_bounds__(256) void k_raster_depth_big(RasterArgs a) {
  raster_big(a, [&](uint32_t, size_t i, double z) {
    atomicMax(&a.zbits[i], static_cast<unsigned long long>(__double_as_longlong(z)));
  });
}

__global__ __launch_bounds__(256) void k_raster_id(RasterArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_tri) return;
  raster_small(a, t, false, [&](size_t i, double z) {
    if (static_cast<unsigned long long>(__double_as_longlong(z)) == a.zbits[i]) atomicMin(&a.tri[i], t);
  });
}

__global__ __launch_bounds__(256) void k_raster_id_big(RasterArgs a) {
  raster_big(a, [&](uint32_t t, size_t i, double z) {
    if (static_cast<unsigned long long>(__double_as_longlong(z)) == a.zbits[i]) atomicMin(&a.tri[i], t);
  });
}

// ---- formats (bit-exact restatements of the host producer's) ------------

__device__ uint16_t f32_to_f16(float f) {
  const uint32_t x = __float_as_uint(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mant = x & 0x7fffffu;
  const int exp = static_cast<int>((x >> 23) & 0xffu);
  if (exp == 0xff) return static_cast<uint16_t>(sign | 0x7c00u | (mant ? 0x200u : 0u));
  const int e = exp - 127 + 15;
  if (e >= 31) return static_cast<uint16_t>(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return static_cast<uint16_t>(sign);
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return static_cast<uint16_t>(sign | h);
  }
  uint32_t h = (static_cast<uint32_t>(e) << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return static_cast<uint16_t>(sign | h);
}

__device__ __forceinline__ int16_t clamp_i16(float v) {
  // GLSL int() truncation, then the RG16I store (saturating).
  const int i = to_int(v);
  return static_cast<int16_t>(min(32767, max(-32768, i)));
}

// PackNormal16I, shader/utils.glsl:65-89.
__device__ __forceinline__ void pack_normal16i(const float* n, int16_t* out) {
  const float px = (n[0] == 0.0f) ? ((n[1] > 0.0f ? 1.0f : (n[1] < 0.0f ? -1.0f : 0.0f)) * kPiF / 2.0f)
                                  : atan2f(n[1], n[0]);
  out[0] = clamp_i16(px * (32768.0f / kPiF));
  out[1] = clamp_i16(n[2] * 32768.0f);
}

__device__ __forceinline__ uint8_t srgb_encode(float c) {
  c = fminf(1.0f, fmaxf(0.0f, c));
  const double s = c <= 0.0031308 ? 12.92 * c : 1.055 * pow(static_cast<double>(c), 1.0 / 2.4) - 0.055;
  return static_cast<uint8_t>(floor(s * 255.0 + 0.5));
}

// ---- resolves -----------------------------------------------------------

struct GbufArgs {
  int w, h;
  const unsigned long long* zbits;
  const uint32_t* tri;
  const float* albedo;
  const float* normals;
  float* depth;
  int16_t* normal;
  uint8_t* albedo_srgb;
};

// shader/fillgbuffer.frag: depth, packed normal, sRGB albedo.
__global__ __launch_bounds__(256) void k_gbuffer_resolve(GbufArgs a) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(a.w) * a.h) return;
  const uint32_t t = a.tri[i];
  a.depth[i] = static_cast<float>(__longlong_as_double(static_cast<long long>(a.zbits[i])));
  if (t == 0xffffffffu) {
    a.normal[2 * i] = a.normal[2 * i + 1] = 0;
    a.albedo_srgb[3 * i] = a.albedo_srgb[3 * i + 1] = a.albedo_srgb[3 * i + 2] = 0;
    return;
  }
  pack_normal16i(a.normals + 3 * static_cast<size_t>(t), a.normal + 2 * i);
  for (int k = 0; k < 3; ++k) a.albedo_srgb[3 * i + k] = srgb_encode(a.albedo[3 * static_cast<size_t>(t) + k]);
}

struct RsmArgs {
  int res;
  const unsigned long long* zbits;
  const uint32_t* tri;
  const float* albedo;
  const float* normals;
  double ivp[16];
  float light_pos[3], light_dir[3], intensity[3];
  float cos_half;
  uint16_t* flux;
  int16_t* normal;
  uint16_t* dls;
  float* shadow;  // the light's depth buffer (the direct-lighting shadow map)
};

// shader/fillrsm.frag:32-61 at one texel: flux = albedo * I * spot falloff *
// pixel solid angle / pi, the packed normal, and (d, d^2) of the distance
// to the light; the world position is reconstructed from the depth.
__global__ __launch_bounds__(256) void k_rsm_resolve(RsmArgs a) {
  const int res = a.res;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(res) * res) return;
  const int x = static_cast<int>(i % res), y = static_cast<int>(i / res);
  const uint32_t t = a.tri[i];
  a.shadow[i] = static_cast<float>(__longlong_as_double(static_cast<long long>(a.zbits[i])));
  if (t == 0xffffffffu) {
    a.flux[3 * i] = a.flux[3 * i + 1] = a.flux[3 * i + 2] = 0;
    a.normal[2 * i] = a.normal[2 * i + 1] = 0;
    a.dls[2 * i] = a.dls[2 * i + 1] = 0;
    return;
  }
  const float d = static_cast<float>(__longlong_as_double(static_cast<long long>(a.zbits[i])));
  const double ndc[4] = {(x + 0.5) / res * 2.0 - 1.0, (y + 0.5) / res * 2.0 - 1.0, static_cast<double>(d), 1.0};
  double wp[4];
  for (int r = 0; r < 4; ++r)
    wp[r] = a.ivp[4 * r] * ndc[0] + a.ivp[4 * r + 1] * ndc[1] + a.ivp[4 * r + 2] * ndc[2] + a.ivp[4 * r + 3];
  const float pos[3] = {static_cast<float>(wp[0] / wp[3]), static_cast<float>(wp[1] / wp[3]),
                        static_cast<float>(wp[2] / wp[3])};
  float tl[3] = {a.light_pos[0] - pos[0], a.light_pos[1] - pos[1], a.light_pos[2] - pos[2]};
  const float dist = sqrtf(tl[0] * tl[0] + tl[1] * tl[1] + tl[2] * tl[2]);
  for (float& v : tl) v /= dist;
  float cos_to_light = -(tl[0] * a.light_dir[0] + tl[1] * a.light_dir[1] + tl[2] * a.light_dir[2]);
  cos_to_light = fminf(1.0f, fmaxf(0.0f, cos_to_light));
  const float total_sr = kPi2F * (1.0f - a.cos_half);
  const float pixel_sr = total_sr * cos_to_light / static_cast<float>(res) / static_cast<float>(res);
  const float falloff = fminf(1.0f, fmaxf(0.0f, cos_to_light - a.cos_half)) / (1.0f - a.cos_half);
  for (int k = 0; k < 3; ++k)
    a.flux[3 * i + k] = f32_to_f16(a.albedo[3 * static_cast<size_t>(t) + k] * a.intensity[k] * (falloff * pixel_sr / kPiF));
  a.dls[2 * i] = f32_to_f16(dist);
  a.dls[2 * i + 1] = f32_to_f16(dist * dist);
  pack_normal16i(a.normals + 3 * static_cast<size_t>(t), a.normal + 2 * i);
}

// shader/downsamplersm.frag:15-33: flux summed over 2x2, normal the
// normalized mean repacked, (d, d^2) averaged.
__global__ __launch_bounds__(256) void k_rsm_downsample(int res, const uint16_t* flux, const int16_t* normal,
                                                        const uint16_t* dls, uint16_t* flux_out, int16_t* normal_out,
                                                        uint16_t* dls_out) {
  const int r2 = res / 2;
  const size_t o = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (o >= static_cast<size_t>(r2) * r2) return;
  const int x = static_cast<int>(o % r2), y = static_cast<int>(o / r2);
  const size_t s[4] = {static_cast<size_t>(2 * y) * res + 2 * x, static_cast<size_t>(2 * y) * res + 2 * x + 1,
                       static_cast<size_t>(2 * y + 1) * res + 2 * x, static_cast<size_t>(2 * y + 1) * res + 2 * x + 1};
  for (int k = 0; k < 3; ++k) {
    const float a = half_to_float(flux[3 * s[0] + k]) + half_to_float(flux[3 * s[1] + k]);
    const float b = half_to_float(flux[3 * s[2] + k]) + half_to_float(flux[3 * s[3] + k]);
    flux_out[3 * o + k] = f32_to_f16(a + b);
  }
  float n[3] = {0.0f, 0.0f, 0.0f};
  for (int j = 0; j < 4; ++j) {
    const f3 u = unpack_normal16i(normal[2 * s[j]], normal[2 * s[j] + 1]);
    n[0] += u.x;
    n[1] += u.y;
    n[2] += u.z;
  }
  const float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (l > 0.0f)
    for (float& v : n) v /= l;
  pack_normal16i(n, normal_out + 2 * o);
  for (int k = 0; k < 2; ++k) {
    const float a = half_to_float(dls[2 * s[0] + k]) + half_to_float(dls[2 * s[1] + k]);
    const float b = half_to_float(dls[2 * s[2] + k]) + half_to_float(dls[2 * s[3] + k]);
    dls_out[2 * o + k] = f32_to_f16((a + b) * 0.25f);
  }
}

void raster(const float* positions, uint32_t n_tri, const float* vp, int w, int h, unsigned long long* zbits,
            uint32_t* tri, uint32_t* big, hipStream_t s) {
  RasterArgs a;
  a.positions = positions;
  a.n_tri = n_tri;
  for (int i = 0; i < 16; ++i) a.vp[i] = vp[i];
  a.w = w;
  a.h = h;
  a.zbits = zbits;
  a.tri = tri;
  a.n_big = big;      // big[0] counts, the list follows
  a.big = big + 1;
  const size_t n = static_cast<size_t>(w) * h;
  launch_fill_u32(reinterpret_cast<uint32_t*>(zbits), 2 * n, 0u, s);
  launch_fill_u32(tri, n, 0xffffffffu, s);
  launch_fill_u32(big, 1, 0u, s);
  if (n_tri == 0) return;
  const dim3 grid((n_tri + 255) / 256);
  hipLaunchKernelGGL(k_raster_depth, grid, dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_raster_depth_big, dim3(1024), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_raster_id, grid, dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_raster_id_big, dim3(1024), dim3(256), 0, s, a);
}

}  // namespace

void launch_render_gbuffer(const float* positions, const float* albedo, const float* normals, uint32_t n_tri,
                           const float* view_proj, int w, int h, unsigned long long* zbits, uint32_t* tri,
                           uint32_t* big, float* depth, int16_t* normal, uint8_t* albedo_srgb, hipStream_t s) {
  raster(positions, n_tri, view_proj, w, h, zbits, tri, big, s);
  GbufArgs a;
  a.w = w;
  a.h = h;
  a.zbits = zbits;
  a.tri = tri;
  a.albedo = albedo;
  a.normals = normals;
  a.depth = depth;
  a.normal = normal;
  a.albedo_srgb = albedo_srgb;
  const size_t n = static_cast<size_t>(w) * h;
  hipLaunchKernelGGL(k_gbuffer_resolve, dim3((n + 255) / 256), dim3(256), 0, s, a);
}

void launch_render_rsm(const float* positions, const float* albedo, const float* normals, uint32_t n_tri,
                       const drv_spot_light& light, unsigned long long* zbits, uint32_t* tri, uint32_t* big,
                       uint16_t* flux, int16_t* normal, uint16_t* dls, float* shadow, hipStream_t s) {
  const int res = light.rsm_render_resolution;
  raster(positions, n_tri, light.light_view_projection, res, res, zbits, tri, big, s);
  RsmArgs a;
  a.res = res;
  a.zbits = zbits;
  a.tri = tri;
  a.albedo = albedo;
  a.normals = normals;
  for (int i = 0; i < 16; ++i) a.ivp[i] = light.inverse_light_view_projection[i];
  for (int i = 0; i < 3; ++i) {
    a.light_pos[i] = light.light_position[i];
    a.light_dir[i] = light.light_direction[i];
    a.intensity[i] = light.light_intensity[i];
  }
  a.cos_half = light.light_cos_half_angle;
  a.flux = flux;
  a.normal = normal;
  a.dls = dls;
  a.shadow = shadow;
  const size_t n = static_cast<size_t>(res) * res;
  hipLaunchKernelGGL(k_rsm_resolve, dim3((n + 255) / 256), dim3(256), 0, s, a);
  // Mip chain: level l + 1 from level l, packed one after another.
  size_t off = 0;
  for (int r = res; r > 1; r /= 2) {
    const size_t n_in = static_cast<size_t>(r) * r, n_out = n_in / 4;
    hipLaunchKernelGGL(k_rsm_downsample, dim3((n_out + 255) / 256), dim3(256), 0, s, r, flux + 3 * off,
                       normal + 2 * off, dls + 2 * off, flux + 3 * (off + n_in), normal + 2 * (off + n_in),
                       dls + 2 * (off + n_in));
    off += n_in;
  }
}

}  // namespace drvk
