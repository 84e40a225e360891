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
  float trans;       // transmittance, 1 - the reference's occlusion
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
  t.trans = 1.0f;
  t.s = 0;
  t.done = false;
  return t;
}

// Advances one trace by up to kBatch samples (cacheLightingRSM.comp:211-228);
// it must only be called on a trace that has not finished (the callers stop
// at the returned flag).  The sample positions depend only on the geometry,
// so all samples of the batch are planned first and their loads issued
// together; transmittance is then accumulated in step order.  Returns true
// when the trace has finished.
//
// Per sample, the reference's occlusion update o' = o + (1 - o) v is kept as
// the transmittance T = 1 - o: T' = T - T v, one FMA, with v = the blend
// times 1/255 folded into T (the same value up to FP32 rounding).
template <int kBatch = kStepBatch>
__device__ __forceinline__ bool trace_batch(const VoxelView& vx, __amdgpu_buffer_rsrc_t cr, const LevelInfo* lvi,
                                            Trace& t) {
  const float qf = static_cast<float>(vx.levels - 1);
  float sx[kBatch], sy[kBatch], sz[kBatch], frac[kBatch];
  const LevelInfo* lp[kBatch];
  Footprint f1[kBatch];
  bool take[kBatch];
#pragma unroll
  for (int i = 0; i < kBatch; ++i) {
    take[i] = i == 0 || (!t.done && (t.s + i) < 32);
    t.px = __builtin_fmaf(t.dx, t.step, t.px);
    t.py = __builtin_fmaf(t.dy, t.step, t.py);
    t.pz = __builtin_fmaf(t.dz, t.step, t.pz);
    t.dist = t.dist + t.step;
    const float r = t.dist * t.rho;
    // LINEAR_MIPMAP_LINEAR level pair for lod = log2(r), clamped to [0, q]
    // (GL 4.5 section 8.14.3); frac = 0 means a single level.  lod >= 0, so
    // truncation is its floor and v_fract_f32 is lod - floor(lod) exactly.
    const float lod = __builtin_amdgcn_fmed3f(__builtin_amdgcn_logf(r), 0.0f, qf);
    lp[i] = lvi + static_cast<int>(lod);
    frac[i] = take[i] ? __builtin_amdgcn_fractf(lod) : 0.0f;
    sx[i] = t.px;
    sy[i] = t.py;
    sz[i] = t.pz;
    f1[i] = footprint(lp[i][0], t.px, t.py, t.pz);
    if (take[i]) {
      if (t.dist >= t.goal) t.done = true;
      t.step = fmaxf(1.0f, r * t.k);  // unused once done
    }
  }
  corner_t c1[kBatch], c2[kBatch];
  Footprint f2[kBatch];
#pragma unroll
  for (int i = 0; i < kBatch; ++i) c1[i] = take[i] ? corner_load(cr, f1[i].offset) : corner_t{0u, 0u, 0u, 0u};
  // The second level only where the lane blends two (frac > 0 implies
  // lod < q, so entry l + 1 is a real level).
#pragma unroll
  for (int i = 0; i < kBatch; ++i) {
    if (frac[i] > 0.0f) {
      f2[i] = footprint(lp[i][1], sx[i], sy[i], sz[i]);
      c2[i] = corner_load(cr, f2[i].offset);
    }
  }
#pragma unroll
  for (int i = 0; i < kBatch; ++i) {
    if (take[i]) {
      const float o1 = blend_corner(c1[i], f1[i].ax, f1[i].ay, f1[i].az);
      const float o2 = frac[i] > 0.0f ? blend_corner(c2[i], f2[i].ax, f2[i].ay, f2[i].az) : o1;
      const float o = __builtin_fmaf(frac[i], o2 - o1, o1);
      t.trans = __builtin_fmaf(-(t.trans * (1.0f / 255.0f)), o, t.trans);
      ++t.s;
    }
  }
  return t.done || t.s >= 32 || t.trans <= kOpaque;
}

// Per-block records read through the constant address space: the rays of a
// wave share their block, so these loads are wave-uniform and go through the
// scalar cache into SGPRs.
typedef const __attribute__((address_space(4))) MarchStep* MarchSteps;
typedef const __attribute__((address_space(4))) ShadowBlock* BlockRecord;

__device__ __forceinline__ MarchSteps march_steps(const MarchStep* march, uint32_t block) {
  return (MarchSteps)(march + static_cast<size_t>(block) * kMarchSteps);  // NOLINT(old-style-cast)
}
// (Member by member: the implicit copy constructor takes a generic reference.)
__device__ __forceinline__ ShadowBlock block_record(const ShadowBlock* blocks, uint32_t block) {
  const BlockRecord r = (BlockRecord)(blocks + block);  // NOLINT(old-style-cast)
  return ShadowBlock{r->x, r->y, r->z, r->rho, r->nx, r->ny, r->nz, r->k};
}
// One 64-byte record, one s_load_dwordx16.  (Elements are copied to scalars
// before the bit casts: a bit cast of an ext_vector element lvalue read the
// vector's first element on this compiler.)
__device__ __forceinline__ MarchStep march_step(MarchSteps pl, int n) {
  typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
  typedef const __attribute__((address_space(4))) u32x16* Raw;
  const u32x16 v = *((Raw)(pl + n));  // NOLINT(old-style-cast)
  uint32_t w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = v[i];
  MarchStep e;
  e.step = __uint_as_float(w[0]);
  e.dist = __uint_as_float(w[1]);
  e.frac = __uint_as_float(w[2]);
  e.two = w[3];
#pragma unroll
  for (int h = 0; h < 2; ++h)
    e.lv[h] = MarchLevel{w[4 + 5 * h], __uint_as_float(w[5 + 5 * h]), __uint_as_float(w[6 + 5 * h]),
                         __uint_as_float(w[7 + 5 * h]), __uint_as_float(w[8 + 5 * h])};
  e.pad[0] = e.pad[1] = 0u;
  return e;
}

// Samples n0 .. n1 - 1 of a ray set up by trace_setup (the ray has taken
// samples 0 .. n0 - 1), with the samples of its block's march plan `pl`:
// every lane of the calling wave traces towards the same block (pl is
// wave-uniform) and starts at n0.  Each sample is the one trace_batch takes
// (same positions, levels, weights and transmittance updates, so the same
// values), without the ray computing the plan itself: per step the lanes
// only advance, fetch and blend.  Returns true when the ray stopped at one
// of the samples (its march reached the goal or it turned opaque).  t.s
// (samples taken) is kept only with kCount: it costs a VALU instruction per
// sample and only the step counters read it.
template <bool kCount = false, bool kFloatIndex = false>
__device__ __forceinline__ bool march_planned(MarchSteps pl, __amdgpu_buffer_rsrc_t cr, Trace& t, int n0, int n1) {
  // The lanes that march stay in lockstep (all start at sample n0 and take
  // one sample per iteration until they stop), so the sample index is
  // wave-uniform and the record load is scalar.
  bool stopped = false;
  MarchStep next = march_step(pl, __builtin_amdgcn_readfirstlane(n0));
  for (int n = n0; n < n1; ++n) {
    const MarchStep cur = next;
    next = march_step(pl, __builtin_amdgcn_readfirstlane(min(n + 1, kMarchSteps - 1)));  // a step ahead
    t.px = __builtin_fmaf(t.dx, cur.step, t.px);
    t.py = __builtin_fmaf(t.dy, cur.step, t.py);
    t.pz = __builtin_fmaf(t.dz, cur.step, t.pz);
    // Both loads are issued before either record is blended.
    const Footprint f1 = footprint<kFloatIndex>(cur.lv[0], t.px, t.py, t.pz);
    const Footprint f2 = footprint<kFloatIndex>(cur.lv[1], t.px, t.py, t.pz);
    const corner_t c1 = corner_load(cr, f1.offset);
    corner_t c2;  // read only when loaded (cur.two is wave-uniform)
    if (cur.two) c2 = corner_load(cr, f2.offset);
    const float o1 = blend_corner(c1, f1.ax, f1.ay, f1.az);
    float o = o1;
    if (cur.two) o = __builtin_fmaf(cur.frac, blend_corner(c2, f2.ax, f2.ay, f2.az) - o1, o1);
    t.trans = __builtin_fmaf(-(t.trans * (1.0f / 255.0f)), o, t.trans);
    if (kCount) t.s = n + 1;
    if (cur.dist >= t.goal || t.trans <= kOpaque) {
      stopped = true;
      break;
    }
  }
  return stopped;
}

// The whole trace (samples 0 .. 31) of a ray; lanes with live = false take
// no samples.
template <bool kCount = false, bool kFloatIndex = false>
__device__ __forceinline__ void trace_planned(MarchSteps pl, __amdgpu_buffer_rsrc_t cr, Trace& t, bool live) {
  if (live) march_planned<kCount, kFloatIndex>(pl, cr, t, 0, kMarchSteps);
}

// Steps the reference takes for a ray traced by trace_planned (geometry
// only: the first sample whose march distance reaches the goal, at most 32).
__device__ __forceinline__ int reference_steps_planned(const MarchStep* pl, const Trace& t) {
  int s = t.s;
  if (s > 0 && pl[s - 1].dist >= t.goal) return s;  // the march reached the goal
  while (s < kMarchSteps && pl[s].dist < t.goal) ++s;
  return min(s + 1, kMarchSteps);
}

// The visibility a finished trace returns, saturate(1 - occlusion)
// (cacheLightingRSM.comp:229-231).  A trace stopped at transmittance
// <= kOpaque returns exactly 0: later samples only lower the reference's
// value, so it lies in [0, kOpaque] and 0 is as close to it as the value at
// the stop.  Exact zeros let the SH gather skip a shadow block that no lane
// of a wave sees.
__device__ __forceinline__ float trace_visibility(const Trace& t) {
  const float v = t.trans;
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
