// Cone-traced indirect shadow of one (cache, shadow block) pair
// (shader/cacheLightingRSM.comp:195-231), shared by the visibility kernels
// and the batched gather.
#pragma once

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_sample.hpp"

namespace drvk {

// Below this transmittance a ray counts as fully occluded: the remaining
// samples can lower the returned visibility by at most this much.
constexpr float kOpaque = 1.0f / 1048576.0f;
// Steps planned together by the fused batch kernel; the single-view cone
// trace takes its count from the launch tuning (steps_per_batch).
constexpr int kStepBatch = 2;

// Per-lane state of one cone trace (cacheLightingRSM.comp:195-210).
struct Trace {
  float px, py, pz;  // current sample position, level-0 texel units
  float dx, dy, dz;  // unit direction, texel units per step unit
  float dist, step, goal, rho, k;
  float occ;
  int s;             // samples taken
  bool done;         // dist reached goal (the reference's break)
};

// The direction and the distance come from one hardware reciprocal square
// root (d / |d| = d * rsq, |d| = |d|^2 * rsq) and the goal is scaled by the
// reciprocal of the voxel size (uniform, hoisted by the compiler): a few
// ulp from the GLSL's normalize and divisions, far below the visibility
// tolerance, and cheap enough to set up a ray in every batch.
__device__ __forceinline__ Trace trace_setup(const VoxelView& vox, float voxel_size, f3 wp, f3 vpos,
                                             const ShadowBlock& b) {
  const float fres = static_cast<float>(vox.res);
  float tx = b.x - wp.x, ty = b.y - wp.y, tz = b.z - wp.z;
  const float d2 = tx * tx + ty * ty + tz * tz;
  const float inv = __builtin_amdgcn_rsqf(d2);
  const float ld = d2 * inv;
  tx = tx * inv;
  ty = ty * inv;
  tz = tz * inv;
  Trace t;
  // dirInVoxel = toAverageVal / volumeSize; positions are kept scaled by res.
  t.dx = tx;
  t.dy = ty;
  t.dz = tz;
  t.px = vpos.x * fres + tx * 2.0f;
  t.py = vpos.y * fres + ty * 2.0f;
  t.pz = vpos.z * fres + tz * 2.0f;
  t.dist = 0.0f;
  t.step = 1.0f;
  t.goal = ld * (1.0f / voxel_size) - 2.0f;
  t.rho = b.rho;
  t.k = b.k;
  t.occ = 0.0f;
  t.s = 0;
  t.done = false;
  return t;
}

// Advances one trace by up to kStepBatch samples (cacheLightingRSM.comp:211-228).
// The sample positions depend only on the geometry, so all samples of the
// batch are planned first and their loads issued together; occlusion is then
// accumulated in step order.  Returns true when the trace has finished.
template <int kBatch = kStepBatch>
__device__ __forceinline__ bool trace_batch(const VoxelView& vx, __amdgpu_buffer_rsrc_t cr, const LevelInfo* lvi,
                                            Trace& t) {
  const int q = vx.levels - 1;
  const float qf = static_cast<float>(q);
  float sx[kBatch], sy[kBatch], sz[kBatch], frac[kBatch];
  int lv[kBatch];
  Footprint f1[kBatch];
  bool take[kBatch];
  bool need2 = false;
#pragma unroll
  for (int i = 0; i < kBatch; ++i) {
    take[i] = !t.done && (t.s + i) < 32;
    t.px = __builtin_fmaf(t.dx, t.step, t.px);
    t.py = __builtin_fmaf(t.dy, t.step, t.py);
    t.pz = __builtin_fmaf(t.dz, t.step, t.pz);
    t.dist = t.dist + t.step;
    const float r = t.dist * t.rho;
    // LINEAR_MIPMAP_LINEAR level pair for lod = log2(r), clamped to [0, q]
    // (GL 4.5 section 8.14.3); frac = 0 means a single level.
    const float lod = __builtin_amdgcn_fmed3f(__builtin_amdgcn_logf(r), 0.0f, qf);
    const float fl = floorf(lod);
    lv[i] = static_cast<int>(fl);
    frac[i] = take[i] ? lod - fl : 0.0f;
    need2 = need2 || frac[i] > 0.0f;
    sx[i] = t.px;
    sy[i] = t.py;
    sz[i] = t.pz;
    f1[i] = footprint(lvi, lv[i], t.px, t.py, t.pz);
    if (take[i]) {
      if (t.dist >= t.goal) t.done = true;
      else t.step = fmaxf(1.0f, r * t.k);
    }
  }
  corner_t c1[kBatch], c2[kBatch];
#pragma unroll
  for (int i = 0; i < kBatch; ++i) c1[i] = take[i] ? corner_load(cr, f1[i].index) : corner_t{0u, 0u, 0u, 0u};
  float o[kBatch];
  // The second level is only planned when some lane of the wave blends two.
  if (__any(need2)) {
    Footprint f2[kBatch];
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      f2[i] = footprint(lvi, min(lv[i] + 1, q), sx[i], sy[i], sz[i]);
      c2[i] = frac[i] > 0.0f ? corner_load(cr, f2[i].index) : corner_t{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const float o1 = blend_corner(c1[i], f1[i].ax, f1[i].ay, f1[i].az);
      const float o2 = blend_corner(c2[i], f2[i].ax, f2[i].ay, f2[i].az);
      o[i] = __builtin_fmaf(frac[i], o2 - o1, o1);
    }
  } else {
#pragma unroll
    for (int i = 0; i < kBatch; ++i) o[i] = blend_corner(c1[i], f1[i].ax, f1[i].ay, f1[i].az);
  }
#pragma unroll
  for (int i = 0; i < kBatch; ++i) {
    if (take[i]) {
      const float v = o[i] * (1.0f / 255.0f);
      t.occ = t.occ + (1.0f - t.occ) * v;
      ++t.s;
    }
  }
  return t.done || t.s >= 32 || (1.0f - t.occ) <= kOpaque;
}

// The visibility a finished trace returns, saturate(1 - occlusion)
// (cacheLightingRSM.comp:229-231).  A trace stopped at transmittance
// <= kOpaque returns exactly 0: later samples only lower the reference's
// value, so it lies in [0, kOpaque] and 0 is as close to it as the value at
// the stop.  Exact zeros let the SH gather skip a shadow block that no lane
// of a wave sees.
__device__ __forceinline__ float trace_visibility(const Trace& t) {
  const float v = 1.0f - t.occ;
  return v <= kOpaque ? 0.0f : fminf(1.0f, fmaxf(0.0f, v));
}

// Steps the reference still takes after a trace stopped early (geometry only).
__device__ __forceinline__ int reference_steps(Trace t) {
  int s = t.s;
  if (t.done) return s;
  while (s < 32) {
    t.dist = t.dist + t.step;
    ++s;
    if (t.dist >= t.goal) break;
    t.step = fmaxf(1.0f, t.dist * t.rho * t.k);
  }
  return s;
}

}  // namespace drvk
