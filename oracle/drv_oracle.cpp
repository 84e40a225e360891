// CPU restatement of the indirect-lighting hot path (test infrastructure only;
// see drv_oracle.h).  Built with -ffp-contract=off so that every FP32
// expression below rounds exactly as written.
#include "drv_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr float kPi = 3.14159265358979f;  // shader/utils.glsl:1 (a float literal in GLSL)

template <typename T>
struct V3 {
  T x, y, z;
};
template <typename T>
inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T>
inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// GLSL normalize(x) = x / length(x).
template <typename T>
inline V3<T> normalize(V3<T> a) { return a / length(a); }
template <typename T>
inline T clamp01(T v) { return std::min(T(1), std::max(T(0), v)); }
// GLSL max(x, y) = (x < y) ? y : x  (so max(x, NaN) = x).
template <typename T>
inline T glsl_max(T x, T y) { return (x < y) ? y : x; }

template <typename T>
inline V3<T> load3(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }

// `vec4(v, w) * M` on a row-major upload: out[r] = sum_c M[r][c] * v[c],
// summed left to right (renderer.cpp:329-339 uploads; GLSL vec*mat).
template <typename T>
inline void mat_vec(const float* m, T x, T y, T z, T w, T* out) {
  for (int r = 0; r < 4; ++r)
    out[r] = x * T(m[4 * r + 0]) + y * T(m[4 * r + 1]) + z * T(m[4 * r + 2]) + w * T(m[4 * r + 3]);
}

// Saturating float -> int (GLSL int() is undefined outside the int range).
template <typename T>
inline int to_int(T f) {
  if (!(f == f)) return 0;
  if (f >= T(2147483647.0)) return 2147483647;
  if (f <= T(-2147483648.0)) return -2147483647 - 1;
  return static_cast<int>(f);
}

float half_to_float(uint16_t h) {
  const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1fu;
  uint32_t mant = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (mant == 0) {
      bits = sign;
    } else {
      int e = -1;
      do {
        ++e;
        mant <<= 1;
      } while (!(mant & 0x400u));
      bits = sign | (static_cast<uint32_t>(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (mant << 13);
  } else {
    bits = sign | ((exp + 127 - 15) << 23) | (mant << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// UnpackNormal16I / UnpackNormal, shader/utils.glsl:43-60.
template <typename T>
V3<T> unpack_normal16i(int16_t a, int16_t b) {
  const T pi = T(kPi);
  T px = T(a) * (pi / T(32768.0));
  T py = T(b) * (T(1.0) / T(32768.0));
  T sin_phi = std::sqrt(T(1.0) - py * py);
  return normalize(mk<T>(std::cos(px) * sin_phi, std::sin(px) * sin_phi, py));
}

void morton_decode(uint32_t morton, uint32_t* ox, uint32_t* oy) {
  // cacheLightingRSM.comp:46-62: x = even bits, y = odd bits.
  uint32_t x = morton, y = morton >> 1;
  x &= 0x55555555u; y &= 0x55555555u;
  x |= x >> 1; y |= y >> 1;
  x &= 0x33333333u; y &= 0x33333333u;
  x |= x >> 2; y |= y >> 2;
  x &= 0x0f0f0f0fu; y &= 0x0f0f0f0fu;
  x |= x >> 4; y |= y >> 4;
  x &= 0x00ff00ffu; y &= 0x00ff00ffu;
  x |= x >> 8; y |= y >> 8;
  x &= 0x0000ffffu; y &= 0x0000ffffu;
  *ox = x;
  *oy = y;
}

int cascade_of(const drv_volume* vol, int n, const float* wp) {
  // lightcache.glsl:109-122
  int c = 0;
  for (; c < n - 1; ++c) {
    const drv_cascade& k = vol->cascades[c];
    bool inside = true;
    for (int i = 0; i < 3; ++i) inside = inside && (wp[i] <= k.decision_max[i]) && (wp[i] >= k.decision_min[i]);
    if (inside) break;
  }
  return c;
}

float cascade_transition(const drv_volume* vol, int c, const float* wp) {
  // lightcache.glsl:125-134
  const drv_cascade& k = vol->cascades[c];
  float dmax[3], dmin[3];
  for (int i = 0; i < 3; ++i) {
    dmax[i] = k.decision_max[i] - wp[i];
    dmin[i] = wp[i] - k.decision_min[i];
  }
  float m = std::min(std::min(std::min(dmax[0], dmax[1]), dmax[2]), std::min(std::min(dmin[0], dmin[1]), dmin[2]));
  return clamp01(1.0f - m / (k.world_voxel_size * vol->cav_transition_zone_size));
}

void unproject(const float* inv_vp, int w, int h, int x, int y, float depth, float* wp) {
  // cacheGather.comp:113-119 (and cacheApply.frag:134-135 with gl_FragCoord).
  const float px = static_cast<float>(x) + 0.5f;
  const float py = static_cast<float>(y) + 0.5f;
  const float sx = px / static_cast<float>(w) * 2.0f - 1.0f;
  const float sy = py / static_cast<float>(h) * 2.0f - 1.0f;
  float v[4];
  mat_vec<float>(inv_vp, sx, sy, depth, 1.0f, v);
  wp[0] = v[0] / v[3];
  wp[1] = v[1] / v[3];
  wp[2] = v[2] / v[3];
}

int cache_1d_coord(const drv_volume* vol, int R, const float* wp, int c) {
  // cacheGather.comp:20-30
  const drv_cascade& k = vol->cascades[c];
  int g[3];
  for (int i = 0; i < 3; ++i) {
    int v = to_int((wp[i] - k.min[i]) / k.world_voxel_size);
    g[i] = std::min(R - 1, std::max(0, v));
  }
  return g[0] + g[1] * R + g[2] * R * R + c * R * R * R;
}

int entry_stride(int sh_bands) { return sh_bands >= 2 ? 32 : 16; }

// ---- voxel sampling (GL 4.5 section 8.14, FP32 or FP64) ----------------

template <typename T>
T texel(const uint8_t* lvl, int w, int x, int y, int z) {
  x = std::min(w - 1, std::max(0, x));
  y = std::min(w - 1, std::max(0, y));
  z = std::min(w - 1, std::max(0, z));
  return T(lvl[(static_cast<size_t>(z) * w + y) * w + x]) / T(255.0);
}

template <typename T>
T sample_level(const uint8_t* lvl, int w, const T* p) {
  T u[3], a[3];
  int i0[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = p[k] * T(w) - T(0.5);
    T f = std::floor(u[k]);
    i0[k] = to_int(f);
    a[k] = u[k] - f;
  }
  const T ax = a[0], ay = a[1], az = a[2];
  const T bx = T(1) - ax, by = T(1) - ay, bz = T(1) - az;
  const int x = i0[0], y = i0[1], z = i0[2];
  return bx * by * bz * texel<T>(lvl, w, x, y, z) + ax * by * bz * texel<T>(lvl, w, x + 1, y, z) +
         bx * ay * bz * texel<T>(lvl, w, x, y + 1, z) + ax * ay * bz * texel<T>(lvl, w, x + 1, y + 1, z) +
         bx * by * az * texel<T>(lvl, w, x, y, z + 1) + ax * by * az * texel<T>(lvl, w, x + 1, y, z + 1) +
         bx * ay * az * texel<T>(lvl, w, x, y + 1, z + 1) + ax * ay * az * texel<T>(lvl, w, x + 1, y + 1, z + 1);
}

int log2i(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) ++l;
  return l;
}

size_t level_offset(int res, int level) {
  size_t off = 0;
  for (int l = 0; l < level; ++l) {
    const size_t w = static_cast<size_t>(res >> l);
    off += w * w * w;
  }
  return off;
}

template <typename T>
T sample_voxels(int res, const uint8_t* mips, const T* p, T lod) {
  // LINEAR_MIPMAP_LINEAR with level_base 0 and q = log2(res); c = 0.
  const int q = log2i(res);
  if (!(lod > T(0))) return sample_level<T>(mips, res, p);
  if (lod >= T(q)) return sample_level<T>(mips + level_offset(res, q), res >> q, p);
  const T fl = std::floor(lod);
  const int d1 = to_int(fl);
  const T f = lod - fl;
  const T t1 = sample_level<T>(mips + level_offset(res, d1), res >> d1, p);
  const T t2 = sample_level<T>(mips + level_offset(res, d1 + 1), res >> (d1 + 1), p);
  return (T(1) - f) * t1 + f * t2;
}

// Bilinear fetch of an RG16F level of w x h texels (clamp to edge),
// component `comp`.
template <typename T>
T bilinear_rg16f(const uint16_t* img, int w, int h, T s, T t, int comp) {
  const T u = s * T(w) - T(0.5), v = t * T(h) - T(0.5);
  const T fu = std::floor(u), fv = std::floor(v);
  const int i0 = to_int(fu), j0 = to_int(fv);
  const T a = u - fu, b = v - fv;
  auto tx = [&](int i, int j) {
    i = std::min(w - 1, std::max(0, i));
    j = std::min(h - 1, std::max(0, j));
    return T(half_to_float(img[2 * (static_cast<size_t>(j) * w + i) + comp]));
  };
  return (T(1) - a) * (T(1) - b) * tx(i0, j0) + a * (T(1) - b) * tx(i0 + 1, j0) + (T(1) - a) * b * tx(i0, j0 + 1) +
         a * b * tx(i0 + 1, j0 + 1);
}

// ---- gather -------------------------------------------------------------

template <typename T>
struct Vpl {
  V3<T> flux;
  T disc_area;
  V3<T> pos;
  V3<T> normal;
};

template <typename T>
struct Block {
  V3<T> avg_pos;
  T rho;  // distToSphereRad
};

// Reconstructs the world position of an RSM texel centre at distance d from
// the light (cacheLightingRSM.comp:154-155, 175-176).
template <typename T>
V3<T> rsm_position(const drv_spot_light* light, T s, T t, T d) {
  T clip[4];
  mat_vec<T>(light->inverse_light_view_projection, s * T(2.0) - T(1.0), t * T(2.0) - T(1.0), T(0.0), T(1.0), clip);
  const V3<T> lp = load3<T>(light->light_position);
  const V3<T> p = mk<T>(clip[0] / clip[3], clip[1] / clip[3], clip[2] / clip[3]);
  return lp + normalize(p - lp) * d;
}

// Rows of the read level: the width for the reference's square RSM, half
// of it for the 2:1 extension (whose Morton decode covers 2h x h).
int rsm_height(const orc_rsm* rsm) { return rsm->read_height > 0 ? rsm->read_height : rsm->read_resolution; }

template <typename T>
void prepare_vpls(const drv_spot_light* light, const orc_rsm* rsm, std::vector<Vpl<T>>& vpls) {
  // cacheLightingRSM.comp:138-158 (one LightInfo per RSM texel, Morton order)
  const int R = rsm->read_resolution, H = rsm_height(rsm);
  const uint32_t n = static_cast<uint32_t>(R) * static_cast<uint32_t>(H);
  vpls.resize(n);
  for (uint32_t m = 0; m < n; ++m) {
    uint32_t x, y;
    morton_decode(m, &x, &y);
    const T s = (T(x) + T(0.5)) / T(R);
    const T t = (T(y) + T(0.5)) / T(H);
    const size_t i = static_cast<size_t>(y) * R + x;
    Vpl<T>& v = vpls[m];
    v.flux = mk<T>(half_to_float(rsm->flux[3 * i]), half_to_float(rsm->flux[3 * i + 1]),
                   half_to_float(rsm->flux[3 * i + 2]));
    // The linear sampler at the exact texel centre returns the texel itself.
    const T d = T(half_to_float(rsm->depth_lin_sq[2 * i]));
    v.disc_area = d * d * T(light->val_area_factor);
    v.pos = rsm_position<T>(light, s, t, d);
    v.normal = unpack_normal16i<T>(rsm->normal[2 * i], rsm->normal[2 * i + 1]);
  }
}

template <typename T>
void prepare_blocks(const drv_spot_light* light, const orc_rsm* rsm, std::vector<Block<T>>& blocks) {
  // cacheLightingRSM.comp:169-192: per shadow block (every SampleInterval-th
  // Morton index) the averaged VAL position and cone aperture.
  const int R = rsm->read_resolution, H = rsm_height(rsm);
  const int interval = light->indirect_shadow_sample_interval;
  const int lod = static_cast<int>(light->indirect_shadow_lod);
  const int Rl = R >> lod, Hl = H >> lod;
  const uint32_t n = static_cast<uint32_t>(R) * static_cast<uint32_t>(H) / static_cast<uint32_t>(interval);
  blocks.resize(n);
  for (uint32_t b = 0; b < n; ++b) {
    uint32_t x, y;
    morton_decode(b * static_cast<uint32_t>(interval), &x, &y);
    const T off = T(light->indirect_shadow_sampling_offset);
    const T s = (T(x) + off) / T(R);
    const T t = (T(y) + off) / T(H);
    const T d = bilinear_rg16f<T>(rsm->depth_lin_sq_shadow, Rl, Hl, s, t, 0);
    const T dsq = bilinear_rg16f<T>(rsm->depth_lin_sq_shadow, Rl, Hl, s, t, 1);
    blocks[b].avg_pos = rsm_position<T>(light, s, t, d);
    const T variance = dsq - d * d;
    blocks[b].rho = glsl_max<T>(T(light->indirect_shadow_super_val_width), std::sqrt(variance) * T(2.0) / d);
  }
}

template <typename T>
T cone_trace(const drv_volume* vol, int res, const uint8_t* mips, V3<T> wp, V3<T> voxel_pos, const Block<T>& blk,
             int* steps) {
  // cacheLightingRSM.comp:195-230
  const T vres = T(res);
  V3<T> to_avg = blk.avg_pos - wp;
  const T light_dist = length(to_avg);
  to_avg = to_avg / light_dist;
  const V3<T> dir = mk<T>(to_avg.x / vres, to_avg.y / vres, to_avg.z / vres);
  V3<T> cur = voxel_pos + dir * T(2.0);
  T occlusion = T(0);
  T step = T(1);
  T dist = T(0);
  const T goal = light_dist / T(vol->voxel_size_in_world) - T(2.0);
  const T k = T(2.0) / (T(1.0) - blk.rho);
  int s = 0;
  for (; s < 32; ++s) {
    cur = cur + dir * step;
    dist = dist + step;
    const T radius = dist * blk.rho;
    const T p[3] = {cur.x, cur.y, cur.z};
    const T o = sample_voxels<T>(res, mips, p, std::log2(radius));
    occlusion = occlusion + (T(1.0) - occlusion) * o;
    if (dist >= goal) break;
    step = glsl_max<T>(T(1.0), radius * k);
  }
  if (steps) *steps = std::min(s + 1, 32);
  return clamp01(T(1.0) - occlusion);
}

template <typename T>
struct ShAcc {
  V3<T> c[9];  // SH00, SH1-1, SH10, SH1+1, SH2-2, SH2-1, SH20, SH2+1, SH2+2
};

template <typename T>
void gather_one(const drv_config* cfg, const drv_volume* vol, const drv_sh_constants* shc,
                const drv_spot_light* light, const uint8_t* mips, const std::vector<Vpl<T>>& vpls,
                const std::vector<Block<T>>& blocks, V3<T> wp, ShAcc<T>& acc, T* vis_out, int64_t* steps_out) {
  const int res = cfg->voxel_resolution;
  const T vres = T(res);
  // cacheLightingRSM.comp:103-104
  const T vsz = T(vol->voxel_size_in_world) * vres;
  const V3<T> voxel_pos = mk<T>((wp.x - T(vol->volume_world_min[0])) / vsz, (wp.y - T(vol->volume_world_min[1])) / vsz,
                                (wp.z - T(vol->volume_world_min[2])) / vsz);
  for (auto& c : acc.c) c = mk<T>(0, 0, 0);
  T shadowing = T(1.0);
  const uint32_t interval = static_cast<uint32_t>(light->indirect_shadow_sample_interval);
  const T c0 = T(shc->eva0), c1 = T(shc->eva1), c2 = T(shc->eva2n2_p1_n1), c20 = T(shc->eva20),
          c22 = T(shc->eva2p2);
  const bool sh2 = cfg->sh_bands >= 2;
  int64_t steps = 0;
  for (uint32_t m = 0; m < vpls.size(); ++m) {
    if (cfg->indirect_shadow && m % interval == 0) {
      int st = 0;
      shadowing = cone_trace<T>(vol, res, mips, wp, voxel_pos, blocks[m / interval], &st);
      steps += st;
      if (vis_out) vis_out[m / interval] = shadowing;
    }
    // cacheLightingRSM.comp:236-277
    const Vpl<T>& v = vpls[m];
    // cacheLightingRSM.comp:240-243: INDIRECT_SPECULAR skips dim VALs.
    if (cfg->indirect_specular &&
        (static_cast<float>(v.flux.x) + static_cast<float>(v.flux.y)) + static_cast<float>(v.flux.z) < 0.001f)
      continue;
    V3<T> to_val = v.pos - wp;
    const T d2 = dot(to_val, to_val);
    to_val = to_val * (T(1.0) / std::sqrt(d2));
    T k = clamp01(dot(v.normal, mk<T>(-to_val.x, -to_val.y, -to_val.z)));
    k = k * shadowing;
    const V3<T> rad = v.flux * (k / (d2 + v.disc_area));
    acc.c[0] = acc.c[0] + c0 * rad;
    acc.c[1] = acc.c[1] - (c1 * to_val.y) * rad;
    acc.c[2] = acc.c[2] + (c1 * to_val.z) * rad;
    acc.c[3] = acc.c[3] - (c1 * to_val.x) * rad;
    if (sh2) {
      acc.c[4] = acc.c[4] - (c2 * to_val.x * to_val.y) * rad;
      acc.c[5] = acc.c[5] + (c2 * to_val.y * to_val.z) * rad;
      acc.c[6] = acc.c[6] + (c20 * (to_val.z * to_val.z * T(3.0) - T(1.0))) * rad;
      acc.c[7] = acc.c[7] + (c2 * to_val.x * to_val.z) * rad;
      acc.c[8] = acc.c[8] + (c22 * (to_val.x * to_val.x - to_val.y * to_val.y)) * rad;
    }
  }
  if (steps_out) *steps_out = steps;
}

int resolve_threads(int n) {
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// Field offsets of the std430 LightCacheEntry (shader/lightcache.glsl:33-58)
// for the canonical coefficient order, per RGB channel.
const int kShOffset[9][3] = {
    {7, 11, 15},   // SH00_r/g/b
    {4, 5, 6},     // SH1neg1
    {8, 9, 10},    // SH10
    {12, 13, 14},  // SH1pos1
    {16, 17, 18},  // SH2neg2
    {20, 21, 22},  // SH2neg1
    {19, 23, 27},  // SH20_r/g/b
    {24, 25, 26},  // SH2pos1
    {28, 29, 30},  // SH2pos2
};

// ---- indirect specular ---------------------------------------------------

int specular_size(const drv_config* cfg) { return cfg->specular_envmap_size == 0 ? 16 : cfg->specular_envmap_size; }

int specular_fill(const drv_config* cfg) {
  return std::min(cfg->specular_hole_fill_level, log2i(specular_size(cfg)));
}

// renderer.cpp:253: 2^ceil(log2(ceil(sqrt(max caches)) * S)).
uint32_t specular_width(const drv_config* cfg) {
  const double side = std::ceil(std::sqrt(static_cast<double>(cfg->max_caches))) * specular_size(cfg);
  return static_cast<uint32_t>(std::pow(2.0, std::ceil(std::log2(side))));
}

size_t specular_level_offset(uint32_t width, int level) {
  size_t o = 0;
  for (int l = 0; l < level; ++l) o += static_cast<size_t>(width >> l) * (width >> l);
  return o;
}

// ComputeLocalViewSpace, lightcache.glsl:98-106 (columns X, Y, Z).
void local_view_space(const float* cam, const float* wp, float X[3], float Y[3], float Z[3]) {
  const V3<float> z = normalize(mk<float>(cam[0] - wp[0], cam[1] - wp[1], cam[2] - wp[2]));
  const V3<float> x = normalize(mk<float>(z.z, 0.0f, -z.x));
  const V3<float> y = mk<float>(z.y * x.z - z.z * x.y, z.z * x.x - z.x * x.z, z.x * x.y - z.y * x.x);
  X[0] = x.x, X[1] = x.y, X[2] = x.z;
  Y[0] = y.x, Y[1] = y.y, Y[2] = y.z;
  Z[0] = z.x, Z[1] = z.y, Z[2] = z.z;
}

// `v * mat3(X, Y, Z)` = (dot(v, X), dot(v, Y), dot(v, Z)).
V3<float> to_local(const float X[3], const float Y[3], const float Z[3], V3<float> v) {
  return mk<float>(dot(v, load3<float>(X)), dot(v, load3<float>(Y)), dot(v, load3<float>(Z)));
}

}  // namespace

extern "C" {

void orc_sh_constants(drv_sh_constants* o) {
  // renderer.cpp:296-306 (float arithmetic with ei::PI as a float).
  const float pi = 3.14159265358979f;
  o->cos_lobe0 = std::sqrt(pi) / 2.0f;
  o->cos_lobe1 = std::sqrt(pi / 3.0f);
  o->cos_lobe2n2_p1_n1 = -std::sqrt(15.0f * pi) / 8.0f;
  o->cos_lobe20 = std::sqrt(5.0f * pi) / 16.0f;
  o->cos_lobe2p2 = std::sqrt(15.0f * pi) / 16.0f;
  o->eva0 = 1.0f / (2.0f * std::sqrt(pi));
  o->eva1 = std::sqrt(3.0f) / (2.0f * std::sqrt(pi));
  o->eva2n2_p1_n1 = std::sqrt(15.0f / (4.0f * pi));
  o->eva20 = std::sqrt(5.0f / (16.0f * pi));
  o->eva2p2 = std::sqrt(15.0f / (16.0f * pi));
}

void orc_morton_decode(uint32_t morton, uint32_t* x, uint32_t* y) { morton_decode(morton, x, y); }

void orc_unpack_normal16i(int16_t a, int16_t b, float out[3]) {
  V3<float> n = unpack_normal16i<float>(a, b);
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
}

void orc_pack_normal16i(const float n[3], int16_t out[2]) {
  // shader/utils.glsl:65-89; the RG16I store clamps to the int16 range.
  float px = (n[0] == 0.0f) ? ((n[1] > 0.0f ? 1.0f : (n[1] < 0.0f ? -1.0f : 0.0f)) * kPi / 2.0f)
                            : std::atan2(n[1], n[0]);
  int a = to_int(px * (32768.0f / kPi));
  int b = to_int(n[2] * 32768.0f);
  out[0] = static_cast<int16_t>(std::min(32767, std::max(-32768, a)));
  out[1] = static_cast<int16_t>(std::min(32767, std::max(-32768, b)));
}

float orc_srgb_to_linear(uint8_t c) {
  const double s = c / 255.0;
  return static_cast<float>(s <= 0.04045 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4));
}

int orc_cascade(const drv_volume* vol, int32_t n, const float wp[3]) { return cascade_of(vol, n, wp); }

float orc_cascade_transition(const drv_volume* vol, int32_t c, const float wp[3]) {
  return cascade_transition(vol, c, wp);
}

void orc_unproject_pixel(const float inv_vp[16], int32_t w, int32_t h, int32_t x, int32_t y, float depth,
                         float wp[3]) {
  unproject(inv_vp, w, h, x, y, depth, wp);
}

void orc_update_volume(const drv_config* cfg, const float cam[3], const float smin[3], const float smax[3],
                       drv_volume* out) {
  // renderer.cpp:350-360
  std::memset(out, 0, sizeof(*out));
  float vmin[3], vmax[3], ext[3];
  for (int i = 0; i < 3; ++i) {
    vmin[i] = smin[i] - 0.001f;
    vmax[i] = smax[i] + 0.001f;
    ext[i] = vmax[i] - vmin[i];
  }
  const float largest = std::max(std::max(ext[0], ext[1]), ext[2]);
  for (int i = 0; i < 3; ++i) {
    vmax[i] = vmax[i] + (largest - ext[i]);
    out->volume_world_min[i] = vmin[i];
    out->volume_world_max[i] = vmax[i];
  }
  out->voxel_size_in_world = (vmax[0] - vmin[0]) / static_cast<float>(cfg->voxel_resolution);
  out->cav_transition_zone_size = cfg->cav_transition_size;
  // renderer.cpp:366-428, cascade sizes from renderer.cpp:1183-1187.
  float size = cfg->cav_cascade0_world_size;
  for (int c = 0; c < cfg->cav_cascades; ++c) {
    const float cell = size / static_cast<float>(cfg->cav_resolution);
    drv_cascade& k = out->cascades[c];
    for (int i = 0; i < 3; ++i) {
      const float snapped = std::round(cam[i] / cell) * cell;
      k.min[i] = snapped - size * 0.5f;
      k.max[i] = snapped + size * 0.5f;
      k.decision_min[i] = cam[i] - size * 0.5f + cell * 1.5f;
      k.decision_max[i] = cam[i] + size * 0.5f - cell * 1.5f;
    }
    k.world_voxel_size = cell;
    size = size * 2.0f;
  }
}

void orc_prepare_spot_scalars(const drv_spot_params* p, drv_spot_light* o) {
  // renderer.cpp:680-719
  float dl = std::sqrt(p->direction[0] * p->direction[0] + p->direction[1] * p->direction[1] +
                       p->direction[2] * p->direction[2]);
  for (int i = 0; i < 3; ++i) {
    o->light_intensity[i] = p->intensity[i];
    o->light_position[i] = p->position[i];
    o->light_direction[i] = p->direction[i] / dl;
  }
  o->light_cos_half_angle = std::cos(p->half_angle);
  const int res_pow2 = 1 << static_cast<int>(std::ceil(std::log2(static_cast<double>(p->rsm_resolution))));
  o->rsm_render_resolution = res_pow2;
  const int read_res = static_cast<int>(res_pow2 / std::pow(2.0, static_cast<double>(p->rsm_read_lod)));
  o->rsm_read_resolution = read_res;
  // the 2:1 extension halves the rows of the render and read levels
  o->rsm_render_height = p->rsm_wide ? res_pow2 / 2 : res_pow2;
  const int read_h = p->rsm_wide ? read_res / 2 : read_res;
  o->rsm_read_height = read_h;
  const float clip_w = std::sin(p->half_angle) * p->near_plane * 2.0f;
  const float vaf = clip_w * clip_w / (p->near_plane * p->near_plane * read_res * read_h);
  o->val_area_factor = vaf;
  o->indirect_shadow_lod = static_cast<float>(p->indirect_shadow_lod);
  const float block = static_cast<float>(1 << p->indirect_shadow_lod);
  o->indirect_shadow_block_size = block;
  o->indirect_shadow_sample_interval = static_cast<int>(block * block);
  o->indirect_shadow_super_val_width = std::sqrt(vaf) * block;
  o->indirect_shadow_sampling_offset = 0.5f + std::sqrt(2.0f) * block / 2.0f;
}

void orc_tile_leaders(const int32_t own[256], const int32_t own2[256], uint8_t lead[256], uint8_t lead2[256]) {
  // cacheGather.comp:141-162.  Arrays are indexed [ly * 16 + lx]; the shader
  // indexes its shared lists as cacheList[x][y], which is the same cell.
  auto L = [&](const int32_t* a, int x, int y) { return a[y * 16 + x]; };
  for (int ly = 0; ly < 16; ++ly)
    for (int lx = 0; lx < 16; ++lx) {
      const int o = L(own, lx, ly);
      const int ux = std::max(0, lx - 1), uy = std::max(0, ly - 1);
      const bool self = (ux == lx) && (uy == ly);
      lead[ly * 16 + lx] = (((L(own, lx, uy) != o && L(own, ux, ly) != o && L(own, ux, uy) != o) || self) &&
                            o != -1)
                               ? 1
                               : 0;
      const int o2 = L(own2, lx, ly);
      const int vx = std::min(15, lx + 1), vy = std::min(15, ly + 1);
      const bool corner = (vx == 15) && (vy == 15);
      lead2[ly * 16 + lx] =
          (((L(own2, lx, vy) != o2 && L(own2, vx, ly) != o2 && L(own2, vx, vy) != o2) || corner) && o2 != -1) ? 1
                                                                                                            : 0;
    }
}

uint32_t orc_allocate(const drv_config* cfg, const drv_volume* vol, const drv_frame* frame, const float* depth,
                      uint32_t* cav, float* entries) {
  const int W = cfg->width, H = cfg->height, R = cfg->cav_resolution, C = cfg->cav_cascades;
  const size_t cav_n = static_cast<size_t>(C) * R * R * R;
  std::memset(cav, 0, cav_n * sizeof(uint32_t));  // renderer.cpp:969-970
  const int stride = entry_stride(cfg->sh_bands);
  const uint32_t cap = static_cast<uint32_t>(cfg->max_caches);
  uint32_t count = 0;
  const int R2 = R * R, R3 = R * R * R;

  // AllocateCaches, cacheGather.comp:32-91.
  auto allocate = [&](int coord, int c) {
    static const int off[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                  {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    coord -= R3 * c;
    int bz = coord / R2;
    int by = (coord - bz * R2) / R;
    int bx = coord % R;
    const drv_cascade& k = vol->cascades[c];
    for (int i = 0; i < 8; ++i) {
      const int lx = bx + off[i][0], ly = by + off[i][1], lz = bz + off[i][2];
      const int gx = lx + c * R;
      const bool inside = gx >= 0 && gx < C * R && ly >= 0 && ly < R && lz >= 0 && lz < R;
      const size_t ci = inside ? (static_cast<size_t>(lz) * R + ly) * (static_cast<size_t>(C) * R) + gx : 0;
      // imageAtomicCompSwap(0 -> 0xFFFFFFFF); outside the image it returns 0.
      if (inside && cav[ci] != 0u) continue;
      const uint32_t idx = count++;
      if (idx < cap) {
        float* e = entries + static_cast<size_t>(idx) * stride;
        e[0] = static_cast<float>(lx) * k.world_voxel_size + k.min[0];
        e[1] = static_cast<float>(ly) * k.world_voxel_size + k.min[1];
        e[2] = static_cast<float>(lz) * k.world_voxel_size + k.min[2];
        for (int j = 3; j < stride; ++j) e[j] = 0.0f;
      }
      if (inside) cav[ci] = idx + 1u;
    }
  };

  const int tiles_x = (W + 15) / 16, tiles_y = (H + 15) / 16;
  int32_t own[256], own2[256], casc[256];
  uint8_t lead[256], lead2[256];
  for (int ty = 0; ty < tiles_y; ++ty)
    for (int tx = 0; tx < tiles_x; ++tx) {
      for (int ly = 0; ly < 16; ++ly)
        for (int lx = 0; lx < 16; ++lx) {
          const int t = ly * 16 + lx;
          own[t] = own2[t] = casc[t] = -1;
          const int x = tx * 16 + lx, y = ty * 16 + ly;
          if (x >= W || y >= H) continue;
          const float d = depth[static_cast<size_t>(y) * W + x];
          if (!(d > 0.0001f)) continue;
          float wp[3];
          unproject(frame->inverse_view_projection, W, H, x, y, d, wp);
          const int c = cascade_of(vol, C, wp);
          casc[t] = c;
          own[t] = cache_1d_coord(vol, R, wp, c);
          if (cfg->cav_transition_size > 0.0f) {
            const float tr = cascade_transition(vol, c, wp);
            if (tr > 0.0f && c < C - 1) own2[t] = cache_1d_coord(vol, R, wp, c + 1);
          }
        }
      orc_tile_leaders(own, own2, lead, lead2);
      for (int t = 0; t < 256; ++t) {
        if (lead[t]) allocate(own[t], casc[t]);
        if (cfg->cav_transition_size > 0.0f && lead2[t]) allocate(own2[t], casc[t] + 1);
      }
    }
  return count;
}

void orc_voxelize(int32_t res, const drv_volume* vol, const float* positions, int64_t n_tri, const float world[16],
                  uint8_t* target) {
  const float h = 1.0f / static_cast<float>(res);
  for (int64_t t = 0; t < n_tri; ++t) {
    // voxelize.vert:15-23
    float P[3][3];
    for (int v = 0; v < 3; ++v) {
      const float* p = positions + 9 * t + 3 * v;
      float w4[4];
      mat_vec<float>(world, p[0], p[1], p[2], 1.0f, w4);
      for (int i = 0; i < 3; ++i)
        P[v][i] = (w4[i] - vol->volume_world_min[i]) / (vol->volume_world_max[i] - vol->volume_world_min[i]) * 2.0f -
                  1.0f;
    }
    // voxelize.geom:43-48 dominant axis
    const float v0[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const float v1[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    const float cr[3] = {v0[1] * v1[2] - v0[2] * v1[1], v0[2] * v1[0] - v0[0] * v1[2], v0[0] * v1[1] - v0[1] * v1[0]};
    const float cl = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const float an[3] = {std::fabs(cr[0] / cl), std::fabs(cr[1] / cl), std::fabs(cr[2] / cl)};
    int side = an[0] > an[1] ? 0 : 1;
    side = an[side] > an[2] ? side : 2;
    // voxelize.geom:53-75 swizzle
    float r[3][3];
    for (int v = 0; v < 3; ++v) {
      if (side == 0) {
        r[v][0] = P[v][2]; r[v][1] = P[v][1]; r[v][2] = P[v][0];
      } else if (side == 1) {
        r[v][0] = P[v][0]; r[v][1] = P[v][2]; r[v][2] = P[v][1];
      } else {
        r[v][0] = P[v][0]; r[v][1] = P[v][1]; r[v][2] = P[v][2];
      }
    }
    // voxelize.geom:81-83 raster AABB in window pixels
    float bb[4];
    bb[0] = std::min(std::min(r[0][0], r[1][0]), r[2][0]) - h;
    bb[1] = std::min(std::min(r[0][1], r[1][1]), r[2][1]) - h;
    bb[2] = std::max(std::max(r[0][0], r[1][0]), r[2][0]) + h;
    bb[3] = std::max(std::max(r[0][1], r[1][1]), r[2][1]) + h;
    for (float& b : bb) b = (b * 0.5f + 0.5f) * static_cast<float>(res);
    // voxelize.geom:86-102 edge planes, oriented and pushed out by h*L1
    const float a[2] = {r[0][0] - r[2][0], r[0][1] - r[2][1]};
    const float b[2] = {r[1][0] - r[0][0], r[1][1] - r[0][1]};
    const float e2[2] = {r[2][0] - r[1][0], r[2][1] - r[1][1]};
    // cross((d, 0), (p, 1)) = (d.y, -d.x, d.x * p.y - d.y * p.x)
    float pl[3][3] = {{a[1], -a[0], a[0] * r[2][1] - a[1] * r[2][0]},
                      {b[1], -b[0], b[0] * r[0][1] - b[1] * r[0][0]},
                      {e2[1], -e2[0], e2[0] * r[1][1] - e2[1] * r[1][0]}};
    const float wind_raw = a[0] * b[1] - b[0] * a[1];
    const float wind = wind_raw > 0.0f ? 1.0f : (wind_raw < 0.0f ? -1.0f : 0.0f);
    if (wind == 0.0f) continue;  // zero-area primitive: nothing is rasterized
    for (auto& p : pl) {
      p[0] = p[0] * wind;
      p[1] = p[1] * wind;
      p[2] = p[2] * wind;
      p[2] = p[2] - (h * std::fabs(p[0]) + h * std::fabs(p[1]));
    }
    // voxelize.geom:118-128: expanded vertices = pairwise line intersections
    auto isect = [&](const float* p, const float* q, float* out) {
      const float cx = p[1] * q[2] - p[2] * q[1];
      const float cy = p[2] * q[0] - p[0] * q[2];
      const float cw = p[0] * q[1] - p[1] * q[0];
      out[0] = cx / cw;
      out[1] = cy / cw;
    };
    float V[3][3];
    isect(pl[0], pl[1], V[0]);
    isect(pl[1], pl[2], V[1]);
    isect(pl[2], pl[0], V[2]);
    for (int v = 0; v < 3; ++v) V[v][2] = r[v][2];
    // Depth plane over the expanded triangle (screen-space linear, w = 1).
    const float ex1 = V[1][0] - V[0][0], ey1 = V[1][1] - V[0][1];
    const float ex2 = V[2][0] - V[0][0], ey2 = V[2][1] - V[0][1];
    const float area = ex1 * ey2 - ex2 * ey1;
    if (!(area != 0.0f) || !std::isfinite(area)) continue;
    const float dz1 = V[1][2] - V[0][2], dz2 = V[2][2] - V[0][2];
    // dz/dx_ndc, dz/dy_ndc == dFdxCoarse/dFdyCoarse of gl_FragCoord.z * res
    const float gx = (dz1 * ey2 - dz2 * ey1) / area;
    const float gy = (dz2 * ex1 - dz1 * ex2) / area;
    const float max_change = std::sqrt(gx * gx + gy * gy) * 1.414f;
    // Pixel range from the expanded AABB (the only pixels that survive).
    const int px0 = std::max(0, to_int(std::floor(bb[0] - 0.5f)));
    const int py0 = std::max(0, to_int(std::floor(bb[1] - 0.5f)));
    const int px1 = std::min(res - 1, to_int(std::ceil(bb[2] - 0.5f)));
    const int py1 = std::min(res - 1, to_int(std::ceil(bb[3] - 0.5f)));
    for (int py = py0; py <= py1; ++py)
      for (int px = px0; px <= px1; ++px) {
        const float fx = static_cast<float>(px) + 0.5f, fy = static_cast<float>(py) + 0.5f;
        // voxelize.frag:26-30 AABB clip
        if (fx < bb[0] || fy < bb[1] || fx > bb[2] || fy > bb[3]) continue;
        // Coverage: pixel centre inside all three pushed edge half-planes
        // (interior is the non-positive side after the winding flip).
        const float nx = fx / static_cast<float>(res) * 2.0f - 1.0f;
        const float ny = fy / static_cast<float>(res) * 2.0f - 1.0f;
        bool in = true;
        for (auto& p : pl) in = in && (p[0] * nx + p[1] * ny + p[2] <= 0.0f);
        if (!in) continue;
        // Interpolated NDC depth at the pixel centre -> window depth.
        const float zn = V[0][2] + gx * (nx - V[0][0]) + gy * (ny - V[0][1]);
        const float zw = zn * 0.5f + 0.5f;
        // voxelize.frag:33-56
        const float vz = zw * static_cast<float>(res);
        const int iz = to_int(vz);
        auto store = [&](int sx, int sy, int sz) {
          int x, y, z;
          if (side == 0) {
            x = sz; y = sy; z = sx;
          } else if (side == 1) {
            x = sx; y = sz; z = sy;
          } else {
            x = sx; y = sy; z = sz;
          }
          if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return;
          target[(static_cast<size_t>(z) * res + y) * res + x] = 255;
        };
        store(px, py, iz);
        if (iz != to_int(vz - max_change)) store(px, py, iz - 1);
        if (iz != to_int(vz + max_change)) store(px, py, iz + 1);
      }
  }
}

void orc_voxel_blend(int32_t res, const uint8_t* target, uint8_t* volume, float f) {
  const size_t n = static_cast<size_t>(res) * res * res;
  for (size_t i = 0; i < n; ++i) {
    const float t = static_cast<float>(target[i]) / 255.0f;
    const float o = static_cast<float>(volume[i]) / 255.0f;
    const float d = t - o;
    const float s = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
    float v = o + s * f;
    v = clamp01(v);
    volume[i] = static_cast<uint8_t>(std::floor(v * 255.0f + 0.5f));
  }
}

void orc_voxel_mips(int32_t res, uint8_t* mips) {
  const int q = log2i(res);
  for (int l = 0; l < q; ++l) {
    const int w = res >> l, w2 = w >> 1;
    const uint8_t* src = mips + level_offset(res, l);
    uint8_t* dst = mips + level_offset(res, l + 1);
    for (int z = 0; z < w2; ++z)
      for (int y = 0; y < w2; ++y)
        for (int x = 0; x < w2; ++x) {
          int sum = 0;
          for (int k = 0; k < 8; ++k)
            sum += src[(static_cast<size_t>(2 * z + (k >> 2)) * w + 2 * y + ((k >> 1) & 1)) * w + 2 * x + (k & 1)];
          dst[(static_cast<size_t>(z) * w2 + y) * w2 + x] = static_cast<uint8_t>((sum + 4) / 8);
        }
  }
}

float orc_sample_voxels_f32(int32_t res, const uint8_t* mips, const float p[3], float lod) {
  return sample_voxels<float>(res, mips, p, lod);
}

double orc_sample_voxels_f64(int32_t res, const uint8_t* mips, const double p[3], double lod) {
  return sample_voxels<double>(res, mips, p, lod);
}

float orc_cone_trace_f32(const drv_volume* vol, int32_t res, const uint8_t* mips, const drv_spot_light* light,
                         const orc_rsm* rsm, const float wp[3], uint32_t block_morton, int32_t* steps_out) {
  std::vector<Block<float>> blocks;
  prepare_blocks<float>(light, rsm, blocks);
  const V3<float> w = mk<float>(wp[0], wp[1], wp[2]);
  const float vsz = vol->voxel_size_in_world * static_cast<float>(res);
  const V3<float> vp = mk<float>((w.x - vol->volume_world_min[0]) / vsz, (w.y - vol->volume_world_min[1]) / vsz,
                                 (w.z - vol->volume_world_min[2]) / vsz);
  int st = 0;
  const float s = cone_trace<float>(vol, res, mips, w, vp,
                                    blocks[block_morton / static_cast<uint32_t>(light->indirect_shadow_sample_interval)],
                                    &st);
  if (steps_out) *steps_out = st;
  return s;
}

float orc_cone_march_f32(const drv_volume* vol, int32_t res, const uint8_t* mips, const float wp[3],
                         const float avg_pos[3], float rho, int32_t* steps_out) {
  const V3<float> w = mk<float>(wp[0], wp[1], wp[2]);
  const float vsz = vol->voxel_size_in_world * static_cast<float>(res);
  const V3<float> vp = mk<float>((w.x - vol->volume_world_min[0]) / vsz, (w.y - vol->volume_world_min[1]) / vsz,
                                 (w.z - vol->volume_world_min[2]) / vsz);
  Block<float> b;
  b.avg_pos = mk<float>(avg_pos[0], avg_pos[1], avg_pos[2]);
  b.rho = rho;
  int st = 0;
  const float s = cone_trace<float>(vol, res, mips, w, vp, b, &st);
  if (steps_out) *steps_out = st;
  return s;
}

void orc_prepare_vpls(const drv_spot_light* light, const orc_rsm* rsm, float* out) {
  std::vector<Vpl<float>> vpls;
  prepare_vpls<float>(light, rsm, vpls);
  for (size_t m = 0; m < vpls.size(); ++m) {
    const Vpl<float>& v = vpls[m];
    float* o = out + 12 * m;
    o[0] = v.pos.x;
    o[1] = v.pos.y;
    o[2] = v.pos.z;
    o[3] = v.disc_area;
    o[4] = v.flux.x;
    o[5] = v.flux.y;
    o[6] = v.flux.z;
    o[7] = 0.0f;
    o[8] = v.normal.x;
    o[9] = v.normal.y;
    o[10] = v.normal.z;
    o[11] = 0.0f;
  }
}

void orc_shadow_blocks(const drv_spot_light* light, const orc_rsm* rsm, float* out) {
  std::vector<Block<float>> blocks;
  prepare_blocks<float>(light, rsm, blocks);
  const int R = rsm->read_resolution;
  const int B = 1 << static_cast<int>(light->indirect_shadow_lod);
  for (size_t b = 0; b < blocks.size(); ++b) {
    uint32_t x, y;
    morton_decode(static_cast<uint32_t>(b) * static_cast<uint32_t>(light->indirect_shadow_sample_interval), &x, &y);
    // Block b covers the B x B texel square with lower corner (x, y).
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int dy = 0; dy < B; ++dy)
      for (int dx = 0; dx < B; ++dx) {
        const size_t t = static_cast<size_t>(y + dy) * R + (x + dx);
        const V3<float> n = unpack_normal16i<float>(rsm->normal[2 * t], rsm->normal[2 * t + 1]);
        sx += n.x;
        sy += n.y;
        sz += n.z;
      }
    const float nl = std::sqrt(sx * sx + sy * sy + sz * sz);
    const float inv = nl > 0.0f ? 1.0f / nl : 0.0f;
    float* o = out + 8 * b;
    o[0] = blocks[b].avg_pos.x;
    o[1] = blocks[b].avg_pos.y;
    o[2] = blocks[b].avg_pos.z;
    o[3] = blocks[b].rho;
    o[4] = sx * inv;
    o[5] = sy * inv;
    o[6] = sz * inv;
    o[7] = 0.0f;
  }
}

void orc_mlp_features(const drv_volume* vol, int32_t res, const uint8_t* mips, const float cam[3], const float wp[3],
                      const float blk[8], float out[32]) {
  const float fres = static_cast<float>(res);
  const float vsz = vol->voxel_size_in_world * fres;
  const float* vmin = vol->volume_world_min;
  // Geometry: cache position in the volume, unit direction to the camera,
  // VAL position in the volume, block normal, unit direction to the VAL,
  // distance, aperture, march goal and two cosines.
  const float vp[3] = {(wp[0] - vmin[0]) / vsz, (wp[1] - vmin[1]) / vsz, (wp[2] - vmin[2]) / vsz};
  float c[3] = {cam[0] - wp[0], cam[1] - wp[1], cam[2] - wp[2]};
  const float cl = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const float ci = cl > 0.0f ? 1.0f / cl : 0.0f;
  for (float& v : c) v *= ci;
  float t[3] = {blk[0] - wp[0], blk[1] - wp[1], blk[2] - wp[2]};
  const float ld = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  const float li = ld > 0.0f ? 1.0f / ld : 0.0f;
  for (float& v : t) v *= li;
  const float dist = ld / vol->voxel_size_in_world;
  const float goal = dist - 2.0f;
  const float G[20] = {vp[0], vp[1], vp[2], c[0], c[1], c[2],
                       (blk[0] - vmin[0]) / vsz, (blk[1] - vmin[1]) / vsz, (blk[2] - vmin[2]) / vsz,
                       blk[4], blk[5], blk[6], t[0], t[1], t[2],
                       dist * (1.0f / 64.0f), blk[3] * 16.0f, goal * (1.0f / 64.0f),
                       -(blk[4] * t[0] + blk[5] * t[1] + blk[6] * t[2]), c[0] * t[0] + c[1] * t[1] + c[2] * t[2]};
  // Occupancy samples along the segment from 2 voxels off the cache towards
  // the VAL: 8 at (i + 0.5) / 8 of it on the level of the cone radius there,
  // 4 at (i + 0.5) / 4 on level 3 (single-level trilinear, no mip blend).
  const int q = log2i(res);
  const float span = goal > 0.0f ? goal : 0.0f;
  float S[12];
  for (int k = 0; k < 12; ++k) {
    float d;
    int lvl;
    if (k < 8) {
      d = 2.0f + span * ((static_cast<float>(k) + 0.5f) * 0.125f);
      const float r = d * blk[3];
      const float lod = std::log2(r > 1.0f ? r : 1.0f);
      lvl = std::min(static_cast<int>(lod), q);
    } else {
      d = 2.0f + span * ((static_cast<float>(k - 8) + 0.5f) * 0.25f);
      lvl = std::min(3, q);
    }
    const float p[3] = {(vp[0] * fres + t[0] * d) / fres, (vp[1] * fres + t[1] * d) / fres,
                        (vp[2] * fres + t[2] * d) / fres};
    S[k] = sample_level<float>(mips + level_offset(res, lvl), res >> lvl, p);
  }
  for (int g = 0; g < 4; ++g) {
    for (int e = 0; e < 5; ++e) out[8 * g + e] = G[5 * g + e];
    for (int e = 0; e < 3; ++e) out[8 * g + 5 + e] = S[3 * g + e];
  }
}

void orc_entry_sh(const float* e, int32_t sh_bands, float out[27]) {
  const int n = sh_bands >= 2 ? 9 : 4;
  for (int k = 0; k < 9; ++k)
    for (int c = 0; c < 3; ++c) out[3 * k + c] = k < n ? e[kShOffset[k][c]] : 0.0f;
}

void orc_gather_f32(const drv_config* cfg, const drv_volume* vol, const drv_sh_constants* shc,
                    const drv_spot_light* light, const orc_rsm* rsm, const uint8_t* mips, float* entries,
                    const uint32_t* ids, int64_t n, float* visibility, int64_t* steps, int32_t n_threads) {
  std::vector<Vpl<float>> vpls;
  std::vector<Block<float>> blocks;
  prepare_vpls<float>(light, rsm, vpls);
  if (cfg->indirect_shadow) prepare_blocks<float>(light, rsm, blocks);
  const int stride = entry_stride(cfg->sh_bands);
  const int nb = cfg->indirect_shadow ? static_cast<int>(blocks.size()) : 0;
  const int nsh = cfg->sh_bands >= 2 ? 9 : 4;
  int64_t total_steps = 0;
  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 2) num_threads(nt) reduction(+ : total_steps)
  for (int64_t i = 0; i < n; ++i) {
    const int64_t id = ids ? ids[i] : i;
    float* e = entries + static_cast<size_t>(id) * stride;
    ShAcc<float> acc;
    int64_t st = 0;
    gather_one<float>(cfg, vol, shc, light, mips, vpls, blocks, load3<float>(e), acc,
                      visibility ? visibility + static_cast<size_t>(i) * nb : nullptr, &st);
    total_steps += st;
    // cacheLightingRSM.comp:357-373: Entry.SH += accumulated
    for (int k = 0; k < nsh; ++k) {
      e[kShOffset[k][0]] += acc.c[k].x;
      e[kShOffset[k][1]] += acc.c[k].y;
      e[kShOffset[k][2]] += acc.c[k].z;
    }
  }
  if (steps) *steps = total_steps;
}

void orc_gather_f64(const drv_config* cfg, const drv_volume* vol, const drv_sh_constants* shc,
                    const drv_spot_light* light, const orc_rsm* rsm, const uint8_t* mips, const float* pos, int64_t n,
                    double* sh_out, double* visibility, int32_t n_threads) {
  std::vector<Vpl<double>> vpls;
  std::vector<Block<double>> blocks;
  prepare_vpls<double>(light, rsm, vpls);
  if (cfg->indirect_shadow) prepare_blocks<double>(light, rsm, blocks);
  const int nb = cfg->indirect_shadow ? static_cast<int>(blocks.size()) : 0;
  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 2) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    ShAcc<double> acc;
    gather_one<double>(cfg, vol, shc, light, mips, vpls, blocks, load3<double>(pos + 3 * i), acc,
                       visibility ? visibility + static_cast<size_t>(i) * nb : nullptr, nullptr);
    for (int k = 0; k < 9; ++k) {
      sh_out[27 * i + 3 * k + 0] = acc.c[k].x;
      sh_out[27 * i + 3 * k + 1] = acc.c[k].y;
      sh_out[27 * i + 3 * k + 2] = acc.c[k].z;
    }
  }
}

void orc_bruteforce_rsm(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                        const orc_rsm* rsm, const float* depth, const int16_t* normal, const uint8_t* albedo,
                        const uint8_t* rm, int32_t diffuse_only, const uint32_t* pixels, int64_t n, float* out,
                        int32_t n_threads) {
  const int W = cfg->width, H = cfg->height, R = rsm->read_resolution, RH = rsm_height(rsm);
  if (!pixels) n = static_cast<int64_t>(W) * H;
  // The VALs of the read level (bruteforcersm.frag:49-53), shader order.
  std::vector<Vpl<float>> vals(static_cast<size_t>(R) * RH);
  for (int x = 0; x < R; ++x)
    for (int y = 0; y < RH; ++y) {
      const size_t i = static_cast<size_t>(y) * R + x;
      Vpl<float>& v = vals[static_cast<size_t>(x) * RH + y];
      const float s = (static_cast<float>(x) + 0.5f) / static_cast<float>(R);
      const float t = (static_cast<float>(y) + 0.5f) / static_cast<float>(RH);
      const float d = half_to_float(rsm->depth_lin_sq[2 * i]);
      v.pos = rsm_position<float>(light, s, t, d);
      v.flux = mk<float>(half_to_float(rsm->flux[3 * i]), half_to_float(rsm->flux[3 * i + 1]),
                         half_to_float(rsm->flux[3 * i + 2]));
      v.normal = unpack_normal16i<float>(rsm->normal[2 * i], rsm->normal[2 * i + 1]);
    }
  float lut[256];
  for (int i = 0; i < 256; ++i) lut[i] = orc_srgb_to_linear(static_cast<uint8_t>(i));
  const V3<float> lpos = mk<float>(light->light_position[0], light->light_position[1], light->light_position[2]);
  const V3<float> cam = mk<float>(frame->camera_position[0], frame->camera_position[1], frame->camera_position[2]);
  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    const size_t p = pixels ? pixels[k] : static_cast<size_t>(k);
    const int x = static_cast<int>(p % W), y = static_cast<int>(p / W);
    float* o = out + 3 * k;
    o[0] = o[1] = o[2] = 0.0f;
    if (depth[p] < 0.00001f) continue;
    float wpa[3];
    unproject(frame->inverse_view_projection, W, H, x, y, depth[p], wpa);
    const V3<float> wp = mk<float>(wpa[0], wpa[1], wpa[2]);
    const V3<float> nrm = unpack_normal16i<float>(normal[2 * p], normal[2 * p + 1]);
    const V3<float> to_cam = normalize(cam - wp);
    const float base[3] = {lut[albedo[3 * p]], lut[albedo[3 * p + 1]], lut[albedo[3 * p + 2]]};
    const float rough = rm ? static_cast<float>(rm[2 * p]) / 255.0f : 1.0f;
    const float metal = rm ? static_cast<float>(rm[2 * p + 1]) / 255.0f : 0.0f;
    // lightingfunctions.glsl:26-44
    const float rsq = rough * rough;
    const float blinn = 2.0f / (rsq * rsq + 0.0005f);
    const float norm_f = (blinn + 8.0f) / (8.0f * kPi);
    float diff[3], spec[3];
    for (int c = 0; c < 3; ++c) {
      diff[c] = base[c] * (1.0f - metal) + 0.02f * metal;     // mix(baseColor, 0.02, metallic)
      spec[c] = 0.04f * (1.0f - metal) + base[c] * metal;     // mix(0.04, baseColor, metallic)
    }
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (const Vpl<float>& v : vals) {
      // bruteforcersm.frag:58-80
      V3<float> to_val = v.pos - wp;
      const float dsq = dot(to_val, to_val);
      to_val = to_val * (1.0f / std::sqrt(dsq));
      const float cos_theta = std::min(1.0f, std::max(0.0f, dot(to_val, nrm)));
      const V3<float> vl = v.pos - lpos;
      const float val_area = dot(vl, vl) * light->val_area_factor;
      const float to_intensity = std::min(1.0f, std::max(0.0f, dot(v.normal, to_val * -1.0f)));
      const float f = to_intensity * cos_theta / (dsq + val_area);
      float bs = 0.0f;
      if (!diffuse_only) {  // lightingfunctions.glsl:54-57
        const V3<float> h = normalize(to_cam + to_val);
        bs = norm_f * std::pow(std::min(1.0f, std::max(0.0f, dot(nrm, h))), blinn);
      }
      const float fl[3] = {v.flux.x, v.flux.y, v.flux.z};
      for (int c = 0; c < 3; ++c) acc[c] += (diff[c] / kPi + bs * spec[c]) * (fl[c] * f);
    }
    for (int c = 0; c < 3; ++c) o[c] = acc[c];
  }
}

void orc_direct_lighting(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                         const float* depth, const int16_t* normal, const uint8_t* albedo, const uint8_t* rm,
                         const float* shadow, int32_t sres, float normal_offset, float bias, float* out,
                         int32_t n_threads) {
  const int W = cfg->width, H = cfg->height;
  float lut[256];
  for (int i = 0; i < 256; ++i) lut[i] = orc_srgb_to_linear(static_cast<uint8_t>(i));
  const V3<float> lpos = load3<float>(light->light_position);
  const V3<float> ldir = load3<float>(light->light_direction);
  const V3<float> cam = load3<float>(frame->camera_position);
  // textureProjLod on the shadow sampler: per-texel GREATER comparisons,
  // bilinearly weighted, texels outside the map read the border depth 0.
  auto pcf = [&](float s, float t, float ref) {
    const float u = s * static_cast<float>(sres) - 0.5f, v = t * static_cast<float>(sres) - 0.5f;
    const float fu = std::floor(u), fv = std::floor(v);
    const int i0 = to_int(fu), j0 = to_int(fv);
    const float al = u - fu, be = v - fv;
    auto lit = [&](int i, int j) {
      const float d = (i >= 0 && j >= 0 && i < sres && j < sres) ? shadow[static_cast<size_t>(j) * sres + i] : 0.0f;
      return ref > d ? 1.0f : 0.0f;
    };
    return (1.0f - al) * (1.0f - be) * lit(i0, j0) + al * (1.0f - be) * lit(i0 + 1, j0) +
           (1.0f - al) * be * lit(i0, j0 + 1) + al * be * lit(i0 + 1, j0 + 1);
  };
  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = static_cast<size_t>(y) * W + x;
      float* o = out + 3 * p;
      o[0] = o[1] = o[2] = 0.0f;
      if (depth[p] < 0.00001f) continue;
      float wpa[3];
      unproject(frame->inverse_view_projection, W, H, x, y, depth[p], wpa);
      const V3<float> wp = load3<float>(wpa);
      const V3<float> n = unpack_normal16i<float>(normal[2 * p], normal[2 * p + 1]);
      // directdeferredlighting.frag:20-25
      V3<float> to_light = lpos - wp;
      const float dsq = dot(to_light, to_light);
      to_light = to_light * (1.0f / std::sqrt(dsq));
      const float cos_theta = std::min(1.0f, std::max(0.0f, dot(to_light, n)));
      // directdeferredlighting.frag:27-36
      const float off = normal_offset * (1.0f - cos_theta);
      const V3<float> wo = wp + n * off;
      float po[4], pw[4];
      mat_vec<float>(light->light_view_projection, wo.x, wo.y, wo.z, 1.0f, po);
      mat_vec<float>(light->light_view_projection, wp.x, wp.y, wp.z, 1.0f, pw);
      const float sx = po[0] * 0.5f + 0.5f * pw[3], sy = po[1] * 0.5f + 0.5f * pw[3];
      const float sz = pw[2] + bias;
      const float shadowing = pcf(sx / pw[3], sy / pw[3], sz / pw[3]);
      if (shadowing == 0.0f) continue;  // discard
      const V3<float> to_cam = normalize(cam - wp);
      const float rough = rm ? static_cast<float>(rm[2 * p]) / 255.0f : 1.0f;
      const float metal = rm ? static_cast<float>(rm[2 * p + 1]) / 255.0f : 0.0f;
      // lightingfunctions.glsl:3-10, 26-58
      const float falloff = std::min(1.0f, std::max(0.0f, dot(to_light * -1.0f, ldir) - light->light_cos_half_angle)) /
                            (1.0f - light->light_cos_half_angle);
      const float rsq = rough * rough;
      const float blinn = 2.0f / (rsq * rsq + 0.0005f);
      const float norm = (blinn + 8.0f) / (8.0f * kPi);
      const V3<float> h = normalize(to_cam + to_light);
      const float bp = norm * std::pow(std::min(1.0f, std::max(0.0f, dot(n, h))), blinn);
      const float k = shadowing * falloff * cos_theta / dsq;
      for (int c = 0; c < 3; ++c) {
        const float base = lut[albedo[3 * p + c]];
        const float diff = base * (1.0f - metal) + 0.02f * metal;
        const float spec = 0.04f * (1.0f - metal) + base * metal;
        o[c] = (diff / kPi + bp * spec) * (light->light_intensity[c] * k);
      }
    }
}

// ---- indirect specular -----------------------------------------------------

uint32_t orc_uf_encode(float x, int32_t m) {
  const uint32_t max_finite = (30u << m) | ((1u << m) - 1u);
  if (std::isnan(x)) return (31u << m) | 1u;
  if (!(x > 0.0f)) return 0u;
  if (std::isinf(x)) return 31u << m;
  // Candidates: the two representable neighbours of x, found from the
  // binade; nearest wins, ties to the even mantissa.
  int e2;
  (void)std::frexp(x, &e2);  // x = f * 2^e2, f in [0.5, 1)
  const int e = std::max(e2 - 1, -14);  // binade exponent (denormals share -14)
  const double ulp = std::ldexp(1.0, e - m);
  const double q = static_cast<double>(x) / ulp;  // exact
  double lo = std::floor(q);
  const double frac = q - lo;
  double k = lo;
  if (frac > 0.5 || (frac == 0.5 && std::fmod(lo, 2.0) != 0.0)) k = lo + 1.0;
  const double v = k * ulp;  // the rounded value
  if (v > static_cast<double>(std::ldexp(1.0, 15) * (2.0 - std::ldexp(1.0, -m)))) return max_finite;
  // Encode v: exponent field from its binade.
  if (v < std::ldexp(1.0, -14)) return static_cast<uint32_t>(v / std::ldexp(1.0, -14 - m));
  int ev;
  const double fv = std::frexp(v, &ev);  // v = fv * 2^ev, fv in [0.5, 1)
  const uint32_t exp_field = static_cast<uint32_t>(ev - 1 + 15);
  const uint32_t mant = static_cast<uint32_t>((fv * 2.0 - 1.0) * std::ldexp(1.0, m));
  return (exp_field << m) | mant;
}

float orc_uf_decode(uint32_t c, int32_t m) {
  const uint32_t e = c >> m, f = c & ((1u << m) - 1u);
  if (e == 0u) return static_cast<float>(std::ldexp(static_cast<double>(f), -14 - m));
  if (e == 31u) return f ? std::nanf("") : INFINITY;
  return static_cast<float>(std::ldexp(1.0 + std::ldexp(static_cast<double>(f), -m), static_cast<int>(e) - 15));
}

uint32_t orc_pack_r11g11b10(const float rgb[3]) {
  return orc_uf_encode(rgb[0], 6) | (orc_uf_encode(rgb[1], 6) << 11) | (orc_uf_encode(rgb[2], 5) << 22);
}

void orc_unpack_r11g11b10(uint32_t v, float rgb[3]) {
  rgb[0] = orc_uf_decode(v & 0x7ffu, 6);
  rgb[1] = orc_uf_decode((v >> 11) & 0x7ffu, 6);
  rgb[2] = orc_uf_decode(v >> 22, 5);
}

void orc_hemi_projection(const float d[3], float out[2]) {
  // utils.glsl:146-177
  const float r = std::sqrt(1.0f - d[2]) * 0.5f;
  float phi = std::atan2(d[1], d[0]) * (4.0f / kPi);
  if (phi < -1.0f) phi += 8.0f;
  float px, py;
  if (phi < 3.0f) {
    if (phi < 1.0f) {
      px = r;
      py = phi * r;
    } else {
      px = -(phi - 2.0f) * r;
      py = r;
    }
  } else {
    if (phi < 5.0f) {
      px = -r;
      py = -(phi - 4.0f) * r;
    } else {
      px = (phi - 6.0f) * r;
      py = -r;
    }
  }
  out[0] = px + 0.5f;
  out[1] = py + 0.5f;
}

void orc_local_view_space(const float cam[3], const float wp[3], float x[3], float y[3], float z[3]) {
  local_view_space(cam, wp, x, y, z);
}

void orc_specular_atlas_geometry(const drv_config* cfg, int32_t* width, int32_t* per_row, int32_t* levels) {
  const int S = specular_size(cfg);
  const uint32_t w = specular_width(cfg);
  if (width) *width = static_cast<int32_t>(w);
  if (per_row) *per_row = static_cast<int32_t>(w / static_cast<uint32_t>(S));
  if (levels) *levels = log2i(S) + 1;
}

void orc_specular_gather(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                         const orc_rsm* rsm, const float* pos, int64_t n, const float* visibility, float* maps,
                         int32_t n_threads) {
  std::vector<Vpl<float>> vpls;
  prepare_vpls<float>(light, rsm, vpls);
  const int S = specular_size(cfg), S1 = S + 1;
  const uint32_t interval = static_cast<uint32_t>(light->indirect_shadow_sample_interval);
  const int64_t nb = cfg->indirect_shadow ? static_cast<int64_t>(vpls.size() / interval) : 0;
  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
  for (int64_t c = 0; c < n; ++c) {
    float* map = maps + static_cast<size_t>(c) * S1 * S1 * 3;
    std::fill(map, map + static_cast<size_t>(S1) * S1 * 3, 0.0f);
    const float* wp = pos + 3 * c;
    float X[3], Y[3], Z[3];
    local_view_space(frame->camera_position, wp, X, Y, Z);  // cacheLightingRSM.comp:122-124
    const V3<float> w = load3<float>(wp);
    float shadowing = 1.0f;
    for (uint32_t m = 0; m < vpls.size(); ++m) {
      if (cfg->indirect_shadow && m % interval == 0)
        shadowing = visibility ? visibility[static_cast<size_t>(c) * nb + m / interval] : 1.0f;
      const Vpl<float>& v = vpls[m];
      // cacheLightingRSM.comp:240-264
      if ((v.flux.x + v.flux.y) + v.flux.z < 0.001f) continue;
      V3<float> to_val = v.pos - w;
      const float d2 = dot(to_val, to_val);
      to_val = to_val * (1.0f / std::sqrt(d2));
      float k = clamp01(dot(v.normal, mk<float>(-to_val.x, -to_val.y, -to_val.z)));
      k = k * shadowing;
      const V3<float> rad = v.flux * (k / (d2 + v.disc_area));
      // cacheLightingRSM.comp:281-335
      const V3<float> l = to_local(X, Y, Z, to_val);
      const V3<float> h = normalize(mk<float>(l.x, l.y, l.z + 1.0f));
      const float hv[3] = {h.x, h.y, h.z};
      float proj[2];
      orc_hemi_projection(hv, proj);
      const int tx = std::min(std::max(to_int(proj[0] * static_cast<float>(S) + 0.5f), 0), S);
      const int ty = std::min(std::max(to_int(proj[1] * static_cast<float>(S) + 0.5f), 0), S);
      const V3<float> out = rad * clamp01(dot(h, l));
      float* t = map + (static_cast<size_t>(ty) * S1 + tx) * 3;
      t[0] += out.x;
      t[1] += out.y;
      t[2] += out.z;
    }
  }
}

void orc_specular_finalize(const drv_config* cfg, int64_t n, const float* maps, float* work, int32_t accumulate,
                           uint32_t* atlas) {
  const int S = specular_size(cfg), S1 = S + 1, levels = log2i(S) + 1, fill = specular_fill(cfg);
  const uint32_t W = specular_width(cfg), N = W / static_cast<uint32_t>(S);
  const float norm = ((static_cast<float>(S) * static_cast<float>(S) - 1.0f) + 8.0f) / (8.0f * kPi);
  const size_t SS = static_cast<size_t>(S) * S;
  // Own sums of the light (the in-map texels of each cache's map).
  for (int64_t c = 0; c < n; ++c) {
    const float* map = maps + static_cast<size_t>(c) * S1 * S1 * 3;
    float* wk = work + static_cast<size_t>(c) * SS * 3;
    for (int y = 0; y < S; ++y)
      for (int x = 0; x < S; ++x)
        for (int ch = 0; ch < 3; ++ch) {
          const float v = map[(static_cast<size_t>(y) * S1 + x) * 3 + ch];
          float& o = wk[(static_cast<size_t>(y) * S + x) * 3 + ch];
          o = accumulate ? o + v : v;
        }
  }
  // Texel index S lands in the atlas neighbour (dropped outside the atlas
  // or on a cache past the live count): left column, upper row, corner -
  // added after the neighbour's own sum in that order.
  for (int64_t c = 0; c < n; ++c) {
    const uint32_t cx = static_cast<uint32_t>(c) % N, cy = static_cast<uint32_t>(c) / N;
    float* wk = work + static_cast<size_t>(c) * SS * 3;
    const auto spill = [&](int64_t src, int sx, int sy) {
      return maps + (static_cast<size_t>(src) * S1 * S1 + static_cast<size_t>(sy) * S1 + sx) * 3;
    };
    if (cx > 0)
      for (int y = 0; y < S; ++y)
        for (int ch = 0; ch < 3; ++ch) wk[(static_cast<size_t>(y) * S) * 3 + ch] += spill(c - 1, S, y)[ch];
    if (cy > 0)
      for (int x = 0; x < S; ++x)
        for (int ch = 0; ch < 3; ++ch) wk[static_cast<size_t>(x) * 3 + ch] += spill(c - N, x, S)[ch];
    if (cx > 0 && cy > 0)
      for (int ch = 0; ch < 3; ++ch) wk[ch] += spill(c - N - 1, S, S)[ch];
  }
  // renderer.cpp:994-1045 per cache: level 0, mips, hole fill, store.
  std::vector<float> lv;
  std::vector<size_t> lofs(levels);
  size_t total = 0;
  for (int l = 0; l < levels; ++l) {
    lofs[l] = total;
    total += static_cast<size_t>(S >> l) * (S >> l);
  }
  lv.resize(total * 3);
  const auto quantize = [](float v, int ch) {
    const int m = ch == 2 ? 5 : 6;
    return orc_uf_decode(orc_uf_encode(v, m), m);
  };
  for (int64_t c = 0; c < n; ++c) {
    const uint32_t cx = static_cast<uint32_t>(c) % N, cy = static_cast<uint32_t>(c) / N;
    const float* wk = work + static_cast<size_t>(c) * SS * 3;
    for (size_t t = 0; t < SS; ++t)
      for (int ch = 0; ch < 3; ++ch) lv[t * 3 + ch] = quantize(wk[t * 3 + ch] * norm, ch);
    // specularenvmap_mipmap.frag
    for (int l = 1; l < levels; ++l) {
      const int sl = S >> l, sp = S >> (l - 1);
      const float* src = lv.data() + lofs[l - 1] * 3;
      float* dst = lv.data() + lofs[l] * 3;
      for (int y = 0; y < sl; ++y)
        for (int x = 0; x < sl; ++x) {
          const size_t i00 = (static_cast<size_t>(2 * y) * sp + 2 * x) * 3, i10 = i00 + 3, i01 = i00 + sp * 3,
                       i11 = i01 + 3;
          for (int ch = 0; ch < 3; ++ch)
            dst[(static_cast<size_t>(y) * sl + x) * 3 + ch] =
                quantize(((src[i00 + ch] + src[i10 + ch]) + (src[i01 + ch] + src[i11 + ch])) * 0.25f, ch);
        }
    }
    // specularenvmap_fillholes.frag:8-51, levels fill .. 1
    for (int i = fill; i >= 1; --i) {
      const int si = S >> i, sd = S >> (i - 1);
      const float* src = lv.data() + lofs[i] * 3;
      float* dst = lv.data() + lofs[i - 1] * 3;
      for (int y = 0; y < si; ++y)
        for (int x = 0; x < si; ++x) {
          const float* s = src + (static_cast<size_t>(y) * si + x) * 3;
          if (s[0] + s[1] + s[2] < 0.0001f) continue;
          size_t di[4];
          di[0] = (static_cast<size_t>(2 * y) * sd + 2 * x) * 3;
          di[1] = di[0] + 3;
          di[2] = di[0] + static_cast<size_t>(sd) * 3;
          di[3] = di[2] + 3;
          float d[4][3];
          float sum[3] = {0.0f, 0.0f, 0.0f};
          for (int k = 0; k < 4; ++k) {
            for (int ch = 0; ch < 3; ++ch) d[k][ch] = dst[di[k] + ch];
            if (d[k][0] + d[k][1] + d[k][2] == 0.0f)
              for (int ch = 0; ch < 3; ++ch) d[k][ch] = s[ch];
            for (int ch = 0; ch < 3; ++ch) sum[ch] += d[k][ch];
          }
          float en[3];
          for (int ch = 0; ch < 3; ++ch) en[ch] = 4.0f * s[ch] / (sum[ch] + 0.00001f);
          for (int k = 0; k < 4; ++k)
            for (int ch = 0; ch < 3; ++ch) dst[di[k] + ch] = quantize(d[k][ch] * en[ch], ch);
        }
    }
    for (int l = 0; l < levels; ++l) {
      const int sl = S >> l;
      const size_t wl = W >> l;
      uint32_t* out = atlas + specular_level_offset(W, l);
      for (int y = 0; y < sl; ++y)
        for (int x = 0; x < sl; ++x) {
          const float* v = lv.data() + (lofs[l] + static_cast<size_t>(y) * sl + x) * 3;
          out[(static_cast<size_t>(cy) * sl + y) * wl + static_cast<size_t>(cx) * sl + x] = orc_pack_r11g11b10(v);
        }
    }
  }
}

void orc_specular_sample(const drv_config* cfg, const uint32_t* atlas, float s, float t, float lod, float out[3]) {
  // textureLod with LINEAR_MIPMAP_LINEAR, clamp to edge (GL 4.5 section 8.14)
  const int S = specular_size(cfg);
  const uint32_t AW = specular_width(cfg);
  const int levels = log2i(S) + 1;
  const auto bilinear = [&](int l, float* o) {
    const int wl = static_cast<int>(AW >> l);
    const uint32_t* lvl = atlas + specular_level_offset(AW, l);
    const float u = s * static_cast<float>(wl) - 0.5f, v = t * static_cast<float>(wl) - 0.5f;
    const float fu = std::floor(u), fv = std::floor(v);
    const int i0 = to_int(fu), j0 = to_int(fv);
    const float al = u - fu, be = v - fv;
    const auto tex = [&](int i, int j, float* rgb) {
      i = std::min(std::max(i, 0), wl - 1);
      j = std::min(std::max(j, 0), wl - 1);
      orc_unpack_r11g11b10(lvl[static_cast<size_t>(j) * wl + i], rgb);
    };
    float a[3], b[3], c[3], d[3];
    tex(i0, j0, a);
    tex(i0 + 1, j0, b);
    tex(i0, j0 + 1, c);
    tex(i0 + 1, j0 + 1, d);
    for (int ch = 0; ch < 3; ++ch)
      o[ch] = (1.0f - al) * (1.0f - be) * a[ch] + al * (1.0f - be) * b[ch] + (1.0f - al) * be * c[ch] +
              al * be * d[ch];
  };
  const int q = levels - 1;
  if (lod <= 0.0f) {
    bilinear(0, out);
  } else if (lod >= static_cast<float>(q)) {
    bilinear(q, out);
  } else {
    const float d1 = std::floor(lod), f = lod - d1;
    float t1[3], t2[3];
    bilinear(to_int(d1), t1);
    bilinear(to_int(d1) + 1, t2);
    for (int ch = 0; ch < 3; ++ch) out[ch] = (1.0f - f) * t1[ch] + f * t2[ch];
  }
}

void orc_apply(const drv_config* cfg, const drv_volume* vol, const drv_frame* frame, const drv_sh_constants* shc,
               const float* depth, const int16_t* normal, const uint8_t* albedo, const uint32_t* cav,
               const float* entries, uint32_t live, const orc_specular* spec, float* out, int32_t n_threads) {
  const int W = cfg->width, H = cfg->height, R = cfg->cav_resolution, C = cfg->cav_cascades;
  const int stride = entry_stride(cfg->sh_bands);
  const bool sh2 = cfg->sh_bands >= 2;
  float lut[256];
  for (int i = 0; i < 256; ++i) lut[i] = orc_srgb_to_linear(static_cast<uint8_t>(i));
  const float CL0 = shc->cos_lobe0, CL1 = shc->cos_lobe1, CL2 = shc->cos_lobe2n2_p1_n1, CL20 = shc->cos_lobe20,
              CL22 = shc->cos_lobe2p2;

  // SampleCacheIrradiance, lightcache.glsl:137-183
  auto sample_cache = [&](uint32_t addr, const float* n, float* irr) {
    if (addr >= live) {  // out-of-range SSBO read returns zeros
      irr[0] = irr[1] = irr[2] = 0.0f;
      return;
    }
    const float* e = entries + static_cast<size_t>(addr) * stride;
    for (int ch = 0; ch < 3; ++ch) {
      float v = e[kShOffset[0][ch]] * CL0;
      v = v - e[kShOffset[1][ch]] * (CL1 * n[1]);
      v = v + e[kShOffset[2][ch]] * (CL1 * n[2]);
      v = v - e[kShOffset[3][ch]] * (CL1 * n[0]);
      if (sh2) {
        v = v - e[kShOffset[4][ch]] * (CL2 * n[0] * n[1]);
        v = v + e[kShOffset[5][ch]] * (CL2 * n[1] * n[2]);
        v = v + e[kShOffset[6][ch]] * (CL20 * (n[2] * n[2] * 3.0f - 1.0f));
        v = v + e[kShOffset[7][ch]] * (CL2 * n[0] * n[2]);
        v = v + e[kShOffset[8][ch]] * (CL22 * (n[0] * n[0] - n[1] * n[1]));
      }
      irr[ch] = std::max(v, 0.0f);
    }
  };

  // Specular atlas: textureLod with the linear-clamp trilinear sampler
  // (renderer.cpp:28-29, 1070-1071).
  const bool use_spec = spec && cfg->indirect_specular;
  const int S = use_spec ? specular_size(cfg) : 0;
  const uint32_t AW = use_spec ? specular_width(cfg) : 0u;
  const uint32_t per_row = use_spec ? AW / static_cast<uint32_t>(S) : 1u;
  const int levels = use_spec ? log2i(S) + 1 : 0;
  // textureLod on the env-map of cache (cx, cy): the texture coordinate
  // (lookup + offset) * S_l - 0.5 is evaluated exactly, as the in-region part
  // lookup * S_l - 0.5 (exact in FP32 for lookups at least half a texel in)
  // plus the integer offset; the FP32 sum lookup + offset of the GLSL would
  // round away up to 2^-15 of a cache, below GL's sub-texel precision.
  const auto bilinear_cache = [&](int l, uint32_t cx, uint32_t cy, const float* lookup, float* o) {
    const int sl = S >> l;
    const int wl = static_cast<int>(AW >> l);
    const uint32_t* lvl = spec->atlas + specular_level_offset(AW, l);
    const float u = lookup[0] * static_cast<float>(sl) - 0.5f, v = lookup[1] * static_cast<float>(sl) - 0.5f;
    const float fu = std::floor(u), fv = std::floor(v);
    const int i0 = static_cast<int>(cx) * sl + to_int(fu), j0 = static_cast<int>(cy) * sl + to_int(fv);
    const float al = u - fu, be = v - fv;
    const auto tex = [&](int i, int j, float* rgb) {
      i = std::min(std::max(i, 0), wl - 1);
      j = std::min(std::max(j, 0), wl - 1);
      orc_unpack_r11g11b10(lvl[static_cast<size_t>(j) * wl + i], rgb);
    };
    float a[3], b[3], c[3], d[3];
    tex(i0, j0, a);
    tex(i0 + 1, j0, b);
    tex(i0, j0 + 1, c);
    tex(i0 + 1, j0 + 1, d);
    for (int ch = 0; ch < 3; ++ch)
      o[ch] = (1.0f - al) * (1.0f - be) * a[ch] + al * (1.0f - be) * b[ch] + (1.0f - al) * be * c[ch] +
              al * be * d[ch];
  };
  const auto trilinear = [&](uint32_t addr, const float* lookup, float lod, float* o) {
    const uint32_t cx = addr % per_row, cy = addr / per_row;
    const int q = levels - 1;
    if (lod <= 0.0f) {
      bilinear_cache(0, cx, cy, lookup, o);
    } else if (lod >= static_cast<float>(q)) {
      bilinear_cache(q, cx, cy, lookup, o);
    } else {
      const float d1 = std::floor(lod), f = lod - d1;
      float t1[3], t2[3];
      bilinear_cache(to_int(d1), cx, cy, lookup, t1);
      bilinear_cache(to_int(d1) + 1, cx, cy, lookup, t2);
      for (int ch = 0; ch < 3; ++ch) o[ch] = (1.0f - f) * t1[ch] + f * t2[ch];
    }
  };

  // ComputeLightingFromCaches, cacheApply.frag:28-118
  auto lighting = [&](const float* wp, const float* n, int c, const float* diffuse, const float* lookup, float lod,
                      const float* spec_color, float* outc) {
    const drv_cascade& k = vol->cascades[c];
    float ac[3];
    int a0[3];
    for (int i = 0; i < 3; ++i) {
      ac[i] = (wp[i] - k.min[i]) / k.world_voxel_size;
      a0[i] = to_int(ac[i]);
    }
    float it[3], iv[3];
    for (int i = 0; i < 3; ++i) {
      it[i] = ac[i] - static_cast<float>(a0[i]);
      iv[i] = 1.0f - it[i];
    }
    float w[8];
    w[0] = iv[0] * iv[1] * iv[2];
    w[1] = it[0] * iv[1] * iv[2];
    w[2] = iv[0] * it[1] * iv[2];
    w[3] = it[0] * it[1] * iv[2];
    w[4] = iv[0] * iv[1] * it[2];
    w[5] = it[0] * iv[1] * it[2];
    w[6] = iv[0] * it[1] * it[2];
    w[7] = it[0] * it[1] * it[2];
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float sacc[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 8; ++i) {
      const int x = a0[0] + off[i][0] + R * c, y = a0[1] + off[i][1], z = a0[2] + off[i][2];
      uint32_t addr = 0;
      if (x >= 0 && y >= 0 && z >= 0 && x < C * R && y < R && z < R)
        addr = cav[(static_cast<size_t>(z) * R + y) * (static_cast<size_t>(C) * R) + x];
      addr -= 1u;
      if (use_spec && addr < live) {  // cacheApply.frag:102-106 (no env-map past the live caches)
        float smp[3];
        trilinear(addr, lookup, lod, smp);
        for (int ch = 0; ch < 3; ++ch) sacc[ch] = sacc[ch] + smp[ch] * w[i];
      }
      float irr[3];
      sample_cache(addr, n, irr);
      for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] + irr[ch] * w[i];
    }
    for (int ch = 0; ch < 3; ++ch) outc[ch] = acc[ch] * diffuse[ch] / kPi;
    if (use_spec)
      for (int ch = 0; ch < 3; ++ch) outc[ch] = outc[ch] + sacc[ch] * spec_color[ch];
  };

  const int nt = resolve_threads(n_threads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = static_cast<size_t>(y) * W + x;
      float* o = out + 3 * p;
      const float d = depth[p];
      if (d < 0.00001f) {  // cacheApply.frag:127-131 discard
        o[0] = o[1] = o[2] = 0.0f;
        continue;
      }
      float wp[3];
      unproject(frame->inverse_view_projection, W, H, x, y, d, wp);
      const int c = cascade_of(vol, C, wp);
      float n[3];
      orc_unpack_normal16i(normal[2 * p], normal[2 * p + 1], n);
      float diffuse[3] = {lut[albedo[3 * p]], lut[albedo[3 * p + 1]], lut[albedo[3 * p + 2]]};
      float lookup[2] = {0.0f, 0.0f}, lod = 0.0f, spec_color[3] = {0.0f, 0.0f, 0.0f};
      if (use_spec) {
        // cacheApply.frag:139-154, lightingfunctions.glsl:26-44, utils.glsl:193-201
        const uint8_t* rm = spec->roughness_metallic;
        const float rough = rm ? static_cast<float>(rm[2 * p]) / 255.0f : 1.0f;
        const float metal = rm ? static_cast<float>(rm[2 * p + 1]) / 255.0f : 0.0f;
        for (int ch = 0; ch < 3; ++ch) {
          const float base = diffuse[ch];
          diffuse[ch] = base * (1.0f - metal) + 0.02f * metal;
          spec_color[ch] = 0.04f * (1.0f - metal) + base * metal;
        }
        const float rsq = rough * rough;
        const float blinn = 2.0f / (rsq * rsq + 0.0005f);
        const float fs = static_cast<float>(S);
        lod = std::max(0.0f, std::log2(fs * fs / (1.0f + blinn)) * 0.5f);
        const float max_half = 0.5f / (std::exp2(-std::ceil(lod)) * fs);
        float X[3], Y[3], Z[3];
        local_view_space(frame->camera_position, wp, X, Y, Z);
        const V3<float> cn = to_local(X, Y, Z, load3<float>(n));
        const float cv[3] = {cn.x, cn.y, cn.z};
        orc_hemi_projection(cv, lookup);
        for (int k = 0; k < 2; ++k) lookup[k] = std::min(std::max(lookup[k], max_half), 1.0f - max_half);
      }
      float col[3];
      lighting(wp, n, c, diffuse, lookup, lod, spec_color, col);
      if (cfg->cav_transition_size > 0.0f) {  // cacheApply.frag:172-185
        const float t = cascade_transition(vol, c, wp);
        if (t > 0.0f && c < C - 1) {
          float col2[3];
          lighting(wp, n, c + 1, diffuse, lookup, lod, spec_color, col2);
          for (int ch = 0; ch < 3; ++ch) col[ch] = col[ch] * (1.0f - t) + col2[ch] * t;
        }
      }
      o[0] = col[0];
      o[1] = col[1];
      o[2] = col[2];
    }
}

}  // extern "C"
