// Per-VPL radiance and its SH projection (shader/cacheLightingRSM.comp:235-279)
// as raw sums, and their write-out into a std430 LightCacheEntry
// (cacheLightingRSM.comp:341-374); shared by the single-view gather and the
// batched gather.
#pragma once

#include "drv_internal.hpp"

namespace drvk {

// Raw sums per channel (r, g, b) of one cache: k = 3 * coefficient + channel.
//   0: sum rad                 (SH00 / c0)
//   1: sum t.y rad             (-SH1neg1 / c1)
//   2: sum t.z rad             (SH10 / c1)
//   3: sum t.x rad             (-SH1pos1 / c1)
//   4: sum t.x t.y rad         (-SH2neg2 / c2)
//   5: sum t.y t.z rad         (SH2neg1 / c2)
//   6: sum t.z^2 rad           (SH20 = c20 (3 s6 - s0))
//   7: sum t.x t.z rad         (SH2pos1 / c2)
//   8: sum (t.x^2 - t.y^2) rad (SH2pos2 / c22)
struct ShSums {
  float s[27];
};

__device__ __forceinline__ void sh_zero(ShSums& a) {
#pragma unroll
  for (int k = 0; k < 27; ++k) a.s[k] = 0.0f;
}

// One cache.VPL interaction with the block's shadowing term.
template <bool kSh2>
__device__ __forceinline__ void sh_accumulate(ShSums& a, const Vpl& v, float wx, float wy, float wz,
                                              float shadowing) {
  const float tx0 = v.px - wx, ty0 = v.py - wy, tz0 = v.pz - wz;
  const float d2 = __builtin_fmaf(tz0, tz0, __builtin_fmaf(ty0, ty0, tx0 * tx0));
  const float inv = __builtin_amdgcn_rsqf(d2);
  const float tx = tx0 * inv, ty = ty0 * inv, tz = tz0 * inv;
  const float nd = -__builtin_fmaf(v.nz, tz, __builtin_fmaf(v.ny, ty, v.nx * tx));
  float k = fminf(1.0f, fmaxf(0.0f, nd));
  k = k * shadowing;
  const float f = k * __builtin_amdgcn_rcpf(d2 + v.disc_area);
  const float rad[3] = {v.fr * f, v.fg * f, v.fb * f};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a.s[c] += rad[c];
    a.s[3 + c] = __builtin_fmaf(ty, rad[c], a.s[3 + c]);
    a.s[6 + c] = __builtin_fmaf(tz, rad[c], a.s[6 + c]);
    a.s[9 + c] = __builtin_fmaf(tx, rad[c], a.s[9 + c]);
  }
  if (kSh2) {
    const float bxy = tx * ty, byz = ty * tz, bzz = tz * tz, bxz = tx * tz;
    const float bdd = __builtin_fmaf(tx, tx, -(ty * ty));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a.s[12 + c] = __builtin_fmaf(bxy, rad[c], a.s[12 + c]);
      a.s[15 + c] = __builtin_fmaf(byz, rad[c], a.s[15 + c]);
      a.s[18 + c] = __builtin_fmaf(bzz, rad[c], a.s[18 + c]);
      a.s[21 + c] = __builtin_fmaf(bxz, rad[c], a.s[21 + c]);
      a.s[24 + c] = __builtin_fmaf(bdd, rad[c], a.s[24 + c]);
    }
  }
}

// Adds coefficient j (0..8), channel ch of raw sums v (v0 = the channel's
// sum of rad, needed by SH20) into the entry e with the SH evaluation
// factors of the Constant UBO (renderer.cpp:296-306).
__device__ __forceinline__ void sh_add_coefficient(float* e, uint32_t j, uint32_t ch, float v, float v0,
                                                   const drv_sh_constants& shc) {
  switch (j) {
    case 0: e[7 + 4 * ch] += shc.eva0 * v; break;                       // SH00_r/g/b at 7, 11, 15
    case 1: e[4 + ch] += -(shc.eva1 * v); break;                        // SH1neg1
    case 2: e[8 + ch] += shc.eva1 * v; break;                           // SH10
    case 3: e[12 + ch] += -(shc.eva1 * v); break;                       // SH1pos1
    case 4: e[16 + ch] += -(shc.eva2n2_p1_n1 * v); break;               // SH2neg2
    case 5: e[20 + ch] += shc.eva2n2_p1_n1 * v; break;                  // SH2neg1
    case 6: e[19 + 4 * ch] += shc.eva20 * (3.0f * v - v0); break;       // SH20_r/g/b at 19, 23, 27
    case 7: e[24 + ch] += shc.eva2n2_p1_n1 * v; break;                  // SH2pos1
    default: e[28 + ch] += shc.eva2p2 * v; break;                       // SH2pos2
  }
}

}  // namespace drvk
