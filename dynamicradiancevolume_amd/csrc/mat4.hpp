// Host-side 4x4 matrix helpers shared by the pipeline library and the input
// producer.  Row-major storage, column-vector convention: out = M * v.
//
// The reference builds its matrices with the epsilon library
// (ei::camera / ei::perspectiveDX / ei::invert, camera/camera.hpp:28-30,
// rendering/renderer.cpp:687-692), which is an un-vendored submodule.  These
// are this project's own definitions; both the HIP path and the oracle consume
// the resulting float matrices, so only internal consistency matters.
#pragma once

#include <cmath>
#include <cstring>

namespace drvm {

inline void mat4_mul(const double* a, const double* b, double* out) {
  double r[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
      r[4 * i + j] = s;
    }
  std::memcpy(out, r, sizeof(r));
}

// Gauss-Jordan inverse with partial pivoting in double.  Returns false for a
// singular matrix.
inline bool mat4_invert(const double* m, double* out) {
  double a[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      a[i][j] = m[4 * i + j];
      a[i][j + 4] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (std::fabs(a[p][c]) < 1e-300) return false;
    if (p != c)
      for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 8; ++j) a[c][j] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = a[i][j + 4];
  return true;
}

// Right-handed look-at: the camera looks down its -z axis.
inline void look_at(const double* eye, const double* target, const double* up_hint, double* out) {
  double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& v : f) v /= fl;
  double up[3] = {up_hint[0], up_hint[1], up_hint[2]};
  // Fall back to another up vector when the hint is (nearly) parallel.
  double d = f[0] * up[0] + f[1] * up[1] + f[2] * up[2];
  if (std::fabs(d) > 0.999) {
    up[0] = 0.0;
    up[1] = 0.0;
    up[2] = 1.0;
    if (std::fabs(f[2]) > 0.999) {
      up[0] = 1.0;
      up[2] = 0.0;
    }
  }
  double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double& v : s) v /= sl;
  double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  const double m[16] = {
      s[0], s[1], s[2], -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]),
      u[0], u[1], u[2], -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
      -f[0], -f[1], -f[2], (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]),
      0.0, 0.0, 0.0, 1.0};
  std::memcpy(out, m, sizeof(m));
}

// Reverse-Z perspective (near plane -> depth 1, far plane -> depth 0, clip
// depth range [0, 1]); the reference gets the same effect by passing far and
// near swapped to a DirectX-style projection (camera/camera.hpp:28,
// renderer.cpp:688).  `hfov` is the horizontal field of view in radians.
inline void perspective_reverse_z(double hfov, double aspect, double near_plane, double far_plane, double* out) {
  const double sx = 1.0 / std::tan(hfov * 0.5);
  const double sy = sx * aspect;
  const double a = near_plane / (far_plane - near_plane);
  const double b = near_plane * far_plane / (far_plane - near_plane);
  const double m[16] = {sx, 0.0, 0.0, 0.0,
                        0.0, sy, 0.0, 0.0,
                        0.0, 0.0, a, b,
                        0.0, 0.0, -1.0, 0.0};
  std::memcpy(out, m, sizeof(m));
}

inline void to_float(const double* m, float* out) {
  for (int i = 0; i < 16; ++i) out[i] = static_cast<float>(m[i]);
}

}  // namespace drvm
