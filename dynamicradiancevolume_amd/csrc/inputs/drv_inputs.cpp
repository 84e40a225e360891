// Seeded procedural scenes and a CPU rasterizer producing the pipeline's
// G-buffer and RSM inputs in the reference's texture formats.
//
// Formats follow rendering/renderer.cpp:466-471 (G-buffer: SRGB8 albedo,
// RG16I normal, D32F reverse-Z depth) and renderer.cpp:1287-1290 (RSM:
// RGB16F flux, RG16I normal, RG16F depth/depth^2), the per-texel values follow
// shader/fillgbuffer.frag:21-43 and shader/fillrsm.frag:24-62, and the RSM
// mip step follows shader/downsamplersm.frag:15-33.  Normal maps are flat
// (every procedural surface uses its geometric face normal).
#include "drv_inputs.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../mat4.hpp"

namespace {

constexpr float kPi = 3.14159265358979f;    // shader/utils.glsl:1
constexpr float kPi2 = 6.28318530717958f;   // shader/utils.glsl:2

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double uniform() { return static_cast<double>(g() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct Soup {
  std::vector<float> pos;     // 9 per triangle
  std::vector<float> albedo;  // 3 per triangle
  std::vector<float> normal;  // 3 per triangle

  void tri(const double* a, const double* b, const double* c, const double* col) {
    for (const double* p : {a, b, c})
      for (int k = 0; k < 3; ++k) pos.push_back(static_cast<float>(p[k]));
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) {
      normal.push_back(static_cast<float>(l > 0 ? n[k] / l : 0.0));
      albedo.push_back(static_cast<float>(col[k]));
    }
  }

  // Quad o + u*s + v*t, s,t in [0,1], split into tess x tess cells of two
  // triangles; the face normal is u x v.
  void quad(const double* o, const double* u, const double* v, int tess, const double* col) {
    for (int i = 0; i < tess; ++i)
      for (int j = 0; j < tess; ++j) {
        double s0 = double(i) / tess, s1 = double(i + 1) / tess;
        double t0 = double(j) / tess, t1 = double(j + 1) / tess;
        double p00[3], p10[3], p01[3], p11[3];
        for (int k = 0; k < 3; ++k) {
          p00[k] = o[k] + u[k] * s0 + v[k] * t0;
          p10[k] = o[k] + u[k] * s1 + v[k] * t0;
          p01[k] = o[k] + u[k] * s0 + v[k] * t1;
          p11[k] = o[k] + u[k] * s1 + v[k] * t1;
        }
        tri(p00, p10, p11, col);
        tri(p00, p11, p01, col);
      }
  }

  // Triangle a,b,c subdivided into tess^2 similar triangles.
  void tri_tess(const double* a, const double* b, const double* c, int tess, const double* col) {
    auto pt = [&](int i, int j, double* out) {  // a + (b-a)*i/t + (c-a)*j/t
      for (int k = 0; k < 3; ++k) out[k] = a[k] + (b[k] - a[k]) * i / tess + (c[k] - a[k]) * j / tess;
    };
    for (int i = 0; i < tess; ++i)
      for (int j = 0; i + j < tess; ++j) {
        double p0[3], p1[3], p2[3];
        pt(i, j, p0);
        pt(i + 1, j, p1);
        pt(i, j + 1, p2);
        tri(p0, p1, p2, col);
        if (i + j + 1 < tess) {
          double p3[3];
          pt(i + 1, j + 1, p3);
          tri(p1, p3, p2, col);
        }
      }
  }
};

// Oriented box: centre c, half extents h, rotation `ang` about +y; outward normals.
void add_box(Soup& s, const double* c, const double* h, double ang, int tess, const double* col) {
  const double ca = std::cos(ang), sa = std::sin(ang);
  const double ax[3][3] = {{ca * h[0], 0.0, -sa * h[0]}, {0.0, h[1], 0.0}, {sa * h[2], 0.0, ca * h[2]}};
  for (int axis = 0; axis < 3; ++axis)
    for (int sign = -1; sign <= 1; sign += 2) {
      const double* n = ax[axis];
      const double* u = ax[(axis + 1) % 3];
      const double* v = ax[(axis + 2) % 3];
      // Face centre c + sign*n; corners c + sign*n +- u +- v.  Orientation:
      // (2u) x (2v) is parallel to +n, so swap u/v for the negative face.
      double o[3], uu[3], vv[3];
      const double* a = sign > 0 ? u : v;
      const double* b = sign > 0 ? v : u;
      for (int k = 0; k < 3; ++k) {
        o[k] = c[k] + sign * n[k] - a[k] - b[k];
        uu[k] = 2.0 * a[k];
        vv[k] = 2.0 * b[k];
      }
      s.quad(o, uu, vv, tess, col);
    }
}

// Ramp (wedge) on the ground: footprint 2*h[0] x 2*h[2] around c (y = c[1] is
// the base), rising along local +z from 0 to height 2*h[1]; rotated about +y.
void add_ramp(Soup& s, const double* c, const double* h, double ang, int tess, const double* col) {
  const double ca = std::cos(ang), sa = std::sin(ang);
  auto P = [&](double x, double y, double z, double* out) {
    out[0] = c[0] + ca * x + sa * z;
    out[1] = c[1] + y;
    out[2] = c[2] - sa * x + ca * z;
  };
  double A[3], B[3], C[3], D[3], E[3], F[3];
  P(-h[0], 0, -h[2], A);          // base low edge
  P(h[0], 0, -h[2], B);
  P(-h[0], 0, h[2], C);           // base high edge
  P(h[0], 0, h[2], D);
  P(-h[0], 2 * h[1], h[2], E);    // top edge (above C, D)
  P(h[0], 2 * h[1], h[2], F);
  auto sub = [](const double* a, const double* b, double* o) { for (int k = 0; k < 3; ++k) o[k] = a[k] - b[k]; };
  double u[3], v[3];
  // bottom (normal -y): A + (B-A) s + (C-A) t
  sub(B, A, u); sub(C, A, v); s.quad(A, u, v, tess, col);
  // back wall at +z (normal +z in local frame): C + (D-C) s + (E-C) t
  sub(D, C, u); sub(E, C, v); s.quad(C, u, v, tess, col);
  // slope: A + (E-C) ... use A, B, F, E: A + (B-A) s + (E-A) t (normal up/-z)
  sub(B, A, u); sub(E, A, v); s.quad(A, u, v, tess, col);
  // side triangles
  s.tri_tess(A, C, E, tess, col);   // x = -h: normal -x
  s.tri_tess(B, F, D, tess, col);   // x = +h: normal +x
}

void build_cornell(Soup& s, uint64_t seed, int tess) {
  Rng r(seed);
  double col[3];
  auto rand_col = [&]() { for (double& x : col) x = r.uniform(0.2, 0.9); };
  const double L = 4.0;
  // Walls facing inwards: floor, ceiling, back, left, right (front open).
  struct W { double o[3], u[3], v[3]; };
  const W walls[5] = {
      {{-L, -L, -L}, {0, 0, 2 * L}, {2 * L, 0, 0}},   // floor y=-L, normal +y
      {{-L, L, -L}, {2 * L, 0, 0}, {0, 0, 2 * L}},    // ceiling y=+L, normal -y
      {{-L, -L, -L}, {2 * L, 0, 0}, {0, 2 * L, 0}},   // back z=-L, normal +z
      {{-L, -L, -L}, {0, 2 * L, 0}, {0, 0, 2 * L}},   // left x=-L, normal +x
      {{L, -L, -L}, {0, 0, 2 * L}, {0, 2 * L, 0}},    // right x=+L, normal -x
  };
  for (const W& w : walls) {
    rand_col();
    s.quad(w.o, w.u, w.v, tess, col);
  }
  for (int b = 0; b < 2; ++b) {
    double h[3] = {r.uniform(0.5, 1.5), r.uniform(0.75, 2.5), r.uniform(0.5, 1.5)};
    double cx = (b == 0 ? -1.0 : 1.0) * r.uniform(1.6, L - 1.6);
    double cz = r.uniform(-L + 1.6, L - 2.0);
    double c[3] = {cx, -L + h[1], cz};
    double ang = r.uniform(-0.6, 0.6);
    rand_col();
    add_box(s, c, h, ang, tess, col);
  }
}

void build_boxes(Soup& s, uint64_t seed, int n_objects, int tess) {
  Rng r(seed);
  double col[3];
  auto rand_col = [&]() { for (double& x : col) x = r.uniform(0.2, 0.9); };
  const double L = 20.0;
  rand_col();
  const double o[3] = {-L, 0.0, -L}, u[3] = {0.0, 0.0, 2 * L}, v[3] = {2 * L, 0.0, 0.0};
  s.quad(o, u, v, std::max(1, tess * 4), col);
  for (int i = 0; i < n_objects; ++i) {
    const bool ramp = r.uniform() < 0.4;
    double h[3] = {r.uniform(0.15, 0.8), r.uniform(0.15, 2.0), r.uniform(0.15, 0.8)};
    double c[3] = {r.uniform(-L + 1.0, L - 1.0), 0.0, r.uniform(-L + 1.0, L - 1.0)};
    // Keep a clearing of radius 3.5 around the origin for the camera.
    if (c[0] * c[0] + c[2] * c[2] < 3.5 * 3.5) {
      const double a = r.uniform(0.0, 2.0 * 3.14159265358979);
      c[0] = (3.5 + r.uniform(0.0, L - 4.5)) * std::cos(a);
      c[2] = (3.5 + r.uniform(0.0, L - 4.5)) * std::sin(a);
    }
    double ang = r.uniform(0.0, 2.0 * 3.14159265358979);
    rand_col();
    if (ramp) {
      add_ramp(s, c, h, ang, tess, col);
    } else {
      c[1] = h[1];
      add_box(s, c, h, ang, tess, col);
    }
  }
}

// ---- formats -----------------------------------------------------------

uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
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

float f16_to_f32(uint16_t h) {
  const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1fu;
  uint32_t mant = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (mant == 0) {
      bits = sign;
    } else {
      int e = -1;
      do { ++e; mant <<= 1; } while (!(mant & 0x400u));
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

int16_t clamp_i16(float v) {
  // GLSL int() truncation, then the RG16I store; values outside the int16
  // range (z = 1 gives 32768) are clamped.
  int i = static_cast<int>(v);
  return static_cast<int16_t>(std::min(32767, std::max(-32768, i)));
}

// PackNormal16I, shader/utils.glsl:65-89.
void pack_normal16i(const float* n, int16_t* out) {
  float px = (n[0] == 0.0f) ? ((n[1] > 0.0f ? 1.0f : (n[1] < 0.0f ? -1.0f : 0.0f)) * kPi / 2.0f)
                            : std::atan2(n[1], n[0]);
  out[0] = clamp_i16(px * (32768.0f / kPi));
  out[1] = clamp_i16(n[2] * 32768.0f);
}

// UnpackNormal16I, shader/utils.glsl:43-60.
void unpack_normal16i(int16_t a, int16_t b, float* out) {
  float px = a * (kPi / 32768.0f);
  float py = b * (1.0f / 32768.0f);
  float sin_phi = std::sqrt(1.0f - py * py);
  float v[3] = {std::cos(px) * sin_phi, std::sin(px) * sin_phi, py};
  float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) out[k] = v[k] / l;
}

uint8_t srgb_encode(float c) {
  c = std::min(1.0f, std::max(0.0f, c));
  double s = c <= 0.0031308 ? 12.92 * c : 1.055 * std::pow(static_cast<double>(c), 1.0 / 2.4) - 0.055;
  return static_cast<uint8_t>(std::floor(s * 255.0 + 0.5));
}

// ---- rasterizer --------------------------------------------------------

struct ScreenTri {
  double x[3], y[3], z[3];
  int64_t id;
};

// Projects one triangle with `vp`, clips it against the near plane (w - z >= 0
// for the reverse-Z projection) and appends the screen-space pieces.
void project_clip(const float* p, const double* vp, int w, int h, int64_t id, std::vector<ScreenTri>& out) {
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
    if (da >= 0) std::memcpy(poly[n++], a, sizeof(double) * 4);
    if ((da >= 0) != (db >= 0)) {
      const double t = da / (da - db);
      for (int k = 0; k < 4; ++k) poly[n][k] = a[k] + (b[k] - a[k]) * t;
      ++n;
    }
  }
  if (n < 3) return;
  double sx[4], sy[4], sz[4];
  for (int i = 0; i < n; ++i) {
    if (poly[i][3] <= 1e-12) return;
    sx[i] = (poly[i][0] / poly[i][3] * 0.5 + 0.5) * w;
    sy[i] = (poly[i][1] / poly[i][3] * 0.5 + 0.5) * h;
    sz[i] = poly[i][2] / poly[i][3];
  }
  for (int i = 1; i + 1 < n; ++i) {
    ScreenTri t;
    const int idx[3] = {0, i, i + 1};
    for (int k = 0; k < 3; ++k) {
      t.x[k] = sx[idx[k]];
      t.y[k] = sy[idx[k]];
      t.z[k] = sz[idx[k]];
    }
    t.id = id;
    out.push_back(t);
  }
}

// Depth-tested raster of all triangles: GREATER test on reverse-Z depth,
// pixel centres sampled, both windings accepted.  Rows are split across
// threads.
void raster(const float* positions, int64_t n_tri, const double* vp, int w, int h, std::vector<float>& depth,
            std::vector<int64_t>& tri_id, int n_threads) {
  std::vector<ScreenTri> tris;
  tris.reserve(static_cast<size_t>(n_tri) + 16);
  for (int64_t t = 0; t < n_tri; ++t) project_clip(positions + 9 * t, vp, w, h, t, tris);
  depth.assign(static_cast<size_t>(w) * h, 0.0f);
  tri_id.assign(static_cast<size_t>(w) * h, -1);
  if (n_threads <= 0) n_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  n_threads = std::min(n_threads, std::max(1, h));
  std::vector<double> zbuf(static_cast<size_t>(w) * h, 0.0);
  auto work = [&](int tid) {
    const int y0 = static_cast<int>(static_cast<int64_t>(h) * tid / n_threads);
    const int y1 = static_cast<int>(static_cast<int64_t>(h) * (tid + 1) / n_threads);
    for (const ScreenTri& t : tris) {
      double miny = std::min({t.y[0], t.y[1], t.y[2]}), maxy = std::max({t.y[0], t.y[1], t.y[2]});
      int py0 = std::max(y0, static_cast<int>(std::ceil(miny - 0.5)));
      int py1 = std::min(y1 - 1, static_cast<int>(std::floor(maxy - 0.5)));
      if (py0 > py1) continue;
      double minx = std::min({t.x[0], t.x[1], t.x[2]}), maxx = std::max({t.x[0], t.x[1], t.x[2]});
      int px0 = std::max(0, static_cast<int>(std::ceil(minx - 0.5)));
      int px1 = std::min(w - 1, static_cast<int>(std::floor(maxx - 0.5)));
      if (px0 > px1) continue;
      const double area = (t.x[1] - t.x[0]) * (t.y[2] - t.y[0]) - (t.x[2] - t.x[0]) * (t.y[1] - t.y[0]);
      if (area == 0.0) continue;
      const double inv = 1.0 / area;
      for (int py = py0; py <= py1; ++py) {
        const double cy = py + 0.5;
        for (int px = px0; px <= px1; ++px) {
          const double cx = px + 0.5;
          double e0 = ((t.x[2] - t.x[1]) * (cy - t.y[1]) - (t.y[2] - t.y[1]) * (cx - t.x[1])) * inv;
          double e1 = ((t.x[0] - t.x[2]) * (cy - t.y[2]) - (t.y[0] - t.y[2]) * (cx - t.x[2])) * inv;
          double e2 = 1.0 - e0 - e1;
          if (e0 < 0 || e1 < 0 || e2 < 0) continue;
          const double z = e0 * t.z[0] + e1 * t.z[1] + e2 * t.z[2];
          if (z < 0.0 || z > 1.0) continue;
          const size_t i = static_cast<size_t>(py) * w + px;
          if (z > zbuf[i]) {
            zbuf[i] = z;
            tri_id[i] = t.id;
          }
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < n_threads; ++i) th.emplace_back(work, i);
  for (auto& t : th) t.join();
  for (size_t i = 0; i < zbuf.size(); ++i) depth[i] = static_cast<float>(zbuf[i]);
}

}  // namespace

extern "C" {

uint16_t drvin_f32_to_f16(float f) { return f32_to_f16(f); }
float drvin_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
void drvin_pack_normal16i(const float n[3], int16_t out[2]) { pack_normal16i(n, out); }

int64_t drvin_scene_build(int32_t kind, uint64_t seed, int32_t n_objects, int32_t tess, float* positions,
                          float* albedo, float* normals, int64_t cap, float* bbox) {
  if (tess < 1) return -1;
  Soup s;
  if (kind == DRVIN_SCENE_CORNELL) {
    build_cornell(s, seed, tess);
  } else if (kind == DRVIN_SCENE_BOXES) {
    if (n_objects < 0) return -1;
    build_boxes(s, seed, n_objects, tess);
  } else {
    return -1;
  }
  const int64_t n = static_cast<int64_t>(s.albedo.size() / 3);
  const int64_t m = std::min(n, cap);
  if (positions && m > 0) std::memcpy(positions, s.pos.data(), sizeof(float) * 9 * m);
  if (albedo && m > 0) std::memcpy(albedo, s.albedo.data(), sizeof(float) * 3 * m);
  if (normals && m > 0) std::memcpy(normals, s.normal.data(), sizeof(float) * 3 * m);
  if (bbox) {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (size_t i = 0; i < s.pos.size(); ++i) {
      lo[i % 3] = std::min(lo[i % 3], s.pos[i]);
      hi[i % 3] = std::max(hi[i % 3], s.pos[i]);
    }
    for (int k = 0; k < 3; ++k) {
      bbox[k] = lo[k];
      bbox[3 + k] = hi[k];
    }
  }
  return n;
}

int drvin_camera(const float eye[3], const float target[3], float hfov, float aspect, float near_plane,
                 float far_plane, float view_proj[16], float inverse_view_proj[16]) {
  double e[3] = {eye[0], eye[1], eye[2]}, t[3] = {target[0], target[1], target[2]}, up[3] = {0.0, 1.0, 0.0};
  double view[16], proj[16], vp[16], ivp[16];
  drvm::look_at(e, t, up, view);
  drvm::perspective_reverse_z(hfov, aspect, near_plane, far_plane, proj);
  drvm::mat4_mul(proj, view, vp);
  if (!drvm::mat4_invert(vp, ivp)) return -1;
  drvm::to_float(vp, view_proj);
  drvm::to_float(ivp, inverse_view_proj);
  return 0;
}

int drvin_raster_gbuffer(const float* positions, const float* albedo, const float* normals, int64_t n_tri,
                         const float view_proj[16], int32_t width, int32_t height, float* depth,
                         int16_t* normal, uint8_t* albedo_srgb, int32_t n_threads) {
  if (!positions || !albedo || !normals || width <= 0 || height <= 0) return -1;
  double vp[16];
  for (int i = 0; i < 16; ++i) vp[i] = view_proj[i];
  std::vector<float> d;
  std::vector<int64_t> id;
  raster(positions, n_tri, vp, width, height, d, id, n_threads);
  const size_t n = static_cast<size_t>(width) * height;
  for (size_t i = 0; i < n; ++i) {
    if (depth) depth[i] = d[i];
    const int64_t t = id[i];
    if (t < 0) {
      if (normal) normal[2 * i] = normal[2 * i + 1] = 0;
      if (albedo_srgb) albedo_srgb[3 * i] = albedo_srgb[3 * i + 1] = albedo_srgb[3 * i + 2] = 0;
      continue;
    }
    if (normal) pack_normal16i(normals + 3 * t, normal + 2 * i);
    if (albedo_srgb)
      for (int k = 0; k < 3; ++k) albedo_srgb[3 * i + k] = srgb_encode(albedo[3 * t + k]);
  }
  return 0;
}

int drvin_raster_rsm(const float* positions, const float* albedo, const float* normals, int64_t n_tri,
                     const drv_spot_light* light, uint16_t* flux, int16_t* normal, uint16_t* depth_lin_sq,
                     int32_t n_threads) {
  if (!positions || !albedo || !normals || !light) return -1;
  const int res = light->rsm_render_resolution;
  // rows: the width for a square RSM, half of it for a 2:1 RSM (the square
  // light frustum rendered with texels twice as tall as wide)
  const int res_h = light->rsm_render_height > 0 ? light->rsm_render_height : res;
  if (res <= 0 || res_h <= 0) return -1;
  double vp[16], ivp[16];
  for (int i = 0; i < 16; ++i) {
    vp[i] = light->light_view_projection[i];
    ivp[i] = light->inverse_light_view_projection[i];
  }
  std::vector<float> d;
  std::vector<int64_t> id;
  raster(positions, n_tri, vp, res, res_h, d, id, n_threads);
  const float cos_half = light->light_cos_half_angle;
  for (int y = 0; y < res_h; ++y)
    for (int x = 0; x < res; ++x) {
      const size_t i = static_cast<size_t>(y) * res + x;
      const int64_t t = id[i];
      if (t < 0) {
        flux[3 * i] = flux[3 * i + 1] = flux[3 * i + 2] = 0;
        normal[2 * i] = normal[2 * i + 1] = 0;
        depth_lin_sq[2 * i] = depth_lin_sq[2 * i + 1] = 0;
        continue;
      }
      // World position of the texel centre from its depth (the raster pass
      // interpolates the vertex position; both give the surface point).
      const double ndc[4] = {(x + 0.5) / res * 2.0 - 1.0, (y + 0.5) / res_h * 2.0 - 1.0, d[i], 1.0};
      double wp[4];
      for (int r = 0; r < 4; ++r)
        wp[r] = ivp[4 * r] * ndc[0] + ivp[4 * r + 1] * ndc[1] + ivp[4 * r + 2] * ndc[2] + ivp[4 * r + 3];
      const float pos[3] = {static_cast<float>(wp[0] / wp[3]), static_cast<float>(wp[1] / wp[3]),
                            static_cast<float>(wp[2] / wp[3])};
      // shader/fillrsm.frag:32-48
      float to_light[3] = {light->light_position[0] - pos[0], light->light_position[1] - pos[1],
                           light->light_position[2] - pos[2]};
      const float dist = std::sqrt(to_light[0] * to_light[0] + to_light[1] * to_light[1] + to_light[2] * to_light[2]);
      for (float& v : to_light) v /= dist;
      float cos_to_light = -(to_light[0] * light->light_direction[0] + to_light[1] * light->light_direction[1] +
                             to_light[2] * light->light_direction[2]);
      cos_to_light = std::min(1.0f, std::max(0.0f, cos_to_light));
      const float total_sr = kPi2 * (1.0f - cos_half);
      const float pixel_sr = total_sr * cos_to_light / res / res_h;
      const float falloff = std::min(1.0f, std::max(0.0f, cos_to_light - cos_half)) / (1.0f - cos_half);
      for (int k = 0; k < 3; ++k)
        flux[3 * i + k] = f32_to_f16(albedo[3 * t + k] * light->light_intensity[k] * (falloff * pixel_sr / kPi));
      depth_lin_sq[2 * i] = f32_to_f16(dist);
      depth_lin_sq[2 * i + 1] = f32_to_f16(dist * dist);
      pack_normal16i(normals + 3 * t, normal + 2 * i);
    }
  return 0;
}

int drvin_rsm_downsample(int32_t res, int32_t res_h, const uint16_t* flux, const int16_t* normal,
                         const uint16_t* depth_lin_sq, uint16_t* flux_out, int16_t* normal_out,
                         uint16_t* depth_lin_sq_out) {
  if (res < 2 || (res & 1) || res_h < 2 || (res_h & 1)) return -1;
  const int r2 = res / 2, rh2 = res_h / 2;
  for (int y = 0; y < rh2; ++y)
    for (int x = 0; x < r2; ++x) {
      const size_t o = static_cast<size_t>(y) * r2 + x;
      const size_t s[4] = {static_cast<size_t>(2 * y) * res + 2 * x, static_cast<size_t>(2 * y) * res + 2 * x + 1,
                           static_cast<size_t>(2 * y + 1) * res + 2 * x,
                           static_cast<size_t>(2 * y + 1) * res + 2 * x + 1};
      for (int k = 0; k < 3; ++k) {
        float a = f16_to_f32(flux[3 * s[0] + k]) + f16_to_f32(flux[3 * s[1] + k]);
        float b = f16_to_f32(flux[3 * s[2] + k]) + f16_to_f32(flux[3 * s[3] + k]);
        flux_out[3 * o + k] = f32_to_f16(a + b);
      }
      float n[3] = {0.0f, 0.0f, 0.0f};
      for (int j = 0; j < 4; ++j) {
        float u[3];
        unpack_normal16i(normal[2 * s[j]], normal[2 * s[j] + 1], u);
        for (int k = 0; k < 3; ++k) n[k] += u[k];
      }
      const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (l > 0.0f)
        for (float& v : n) v /= l;
      pack_normal16i(n, normal_out + 2 * o);
      for (int k = 0; k < 2; ++k) {
        float a = f16_to_f32(depth_lin_sq[2 * s[0] + k]) + f16_to_f32(depth_lin_sq[2 * s[1] + k]);
        float b = f16_to_f32(depth_lin_sq[2 * s[2] + k]) + f16_to_f32(depth_lin_sq[2 * s[3] + k]);
        depth_lin_sq_out[2 * o + k] = f32_to_f16((a + b) * 0.25f);
      }
    }
  return 0;
}

}  // extern "C"
