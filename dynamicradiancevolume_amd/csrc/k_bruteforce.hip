// Brute-force RSM lighting: every pixel sums the light of every VAL of the
// RSM directly (shader/bruteforcersm.frag:15-88, drawn by
// Renderer::ApplyRSMsBruteForce, rendering/renderer.cpp:851-874).  It is the
// reference's own unshadowed ground truth for the cache path.
//
// One lane per pixel (16x16-pixel workgroups); every lane visits the VALs in
// the same order, so the VAL record is wave-uniform (scalar loads) and the
// BRDF of shader/lightingfunctions.glsl:46-58 runs on FP32 VALU.
#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

struct BruteArgs {
  float inv_vp[16];
  float cam[3];
  float light_pos[3];
  float val_area_factor;
  const float* depth;
  const int16_t* normal;
  const uint8_t* albedo;
  const uint8_t* roughness_metallic;  // RG8 or null (roughness 1, metallic 0)
  const float* srgb_lut;
  const Vpl* vals;  // records of k_prepare_vpls (position, flux, normal)
  uint32_t n_vals;
  float* out;
  int width, height, accumulate;
};

template <bool kDiffuseOnly>
__global__ __launch_bounds__(256) void k_bruteforce(BruteArgs a) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const bool inside = x < a.width && y < a.height;
  const size_t p = inside ? static_cast<size_t>(y) * a.width + x : 0u;
  const float d = inside ? a.depth[p] : 0.0f;
  const bool lit = inside && d >= 0.00001f;  // background pixels (cleared depth) get no light
  // bruteforcersm.frag:17-28
  const f3 wp = unproject_pixel(a.inv_vp, a.width, a.height, x, y, d);
  const f3 n = unpack_normal16i(a.normal[2 * p], a.normal[2 * p + 1]);
  float cx = a.cam[0] - wp.x, cy = a.cam[1] - wp.y, cz = a.cam[2] - wp.z;
  const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
  cx /= cl;
  cy /= cl;
  cz /= cl;
  const float base[3] = {a.srgb_lut[a.albedo[3 * p]], a.srgb_lut[a.albedo[3 * p + 1]], a.srgb_lut[a.albedo[3 * p + 2]]};
  float rough = 1.0f, metal = 0.0f;
  if (a.roughness_metallic) {
    rough = static_cast<float>(a.roughness_metallic[2 * p]) / 255.0f;
    metal = static_cast<float>(a.roughness_metallic[2 * p + 1]) / 255.0f;
  }
  // lightingfunctions.glsl:26-44: Blinn exponent, normalization, colours.
  const float rsq = rough * rough;
  const float blinn = 2.0f / (rsq * rsq + 0.0005f);
  const float norm = (blinn + 8.0f) / (8.0f * kPi);
  float diff[3], spec[3];
  for (int k = 0; k < 3; ++k) {
    diff[k] = base[k] + (0.02f - base[k]) * metal;
    spec[k] = 0.04f + (base[k] - 0.04f) * metal;
  }
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t i = 0; i < a.n_vals; ++i) {
    const Vpl v = a.vals[i];
    // bruteforcersm.frag:58-80
    float tx = v.px - wp.x, ty = v.py - wp.y, tz = v.pz - wp.z;
    const float dsq = __builtin_fmaf(tz, tz, __builtin_fmaf(ty, ty, tx * tx));
    const float inv = __builtin_amdgcn_rsqf(dsq);
    tx *= inv;
    ty *= inv;
    tz *= inv;
    const float cos_theta = fminf(1.0f, fmaxf(0.0f, __builtin_fmaf(tz, n.z, __builtin_fmaf(ty, n.y, tx * n.x))));
    const float lx = v.px - a.light_pos[0], ly = v.py - a.light_pos[1], lz = v.pz - a.light_pos[2];
    const float val_area = __builtin_fmaf(lz, lz, __builtin_fmaf(ly, ly, lx * lx)) * a.val_area_factor;
    const float to_int = fminf(1.0f, fmaxf(0.0f, -__builtin_fmaf(v.nz, tz, __builtin_fmaf(v.ny, ty, v.nx * tx))));
    const float f = to_int * cos_theta * __builtin_amdgcn_rcpf(dsq + val_area);
    float brdf_s = 0.0f;
    if (!kDiffuseOnly) {
      // lightingfunctions.glsl:54-57: normalized Blinn-Phong of the half vector
      float hx = cx + tx, hy = cy + ty, hz = cz + tz;
      const float hi = __builtin_amdgcn_rsqf(__builtin_fmaf(hz, hz, __builtin_fmaf(hy, hy, hx * hx)));
      const float nh = fminf(1.0f, fmaxf(0.0f, (hx * n.x + hy * n.y + hz * n.z) * hi));
      brdf_s = norm * __builtin_amdgcn_exp2f(blinn * __builtin_amdgcn_logf(nh));
    }
    acc[0] = __builtin_fmaf(diff[0] * (1.0f / kPi) + brdf_s * spec[0], v.fr * f, acc[0]);
    acc[1] = __builtin_fmaf(diff[1] * (1.0f / kPi) + brdf_s * spec[1], v.fg * f, acc[1]);
    acc[2] = __builtin_fmaf(diff[2] * (1.0f / kPi) + brdf_s * spec[2], v.fb * f, acc[2]);
  }
  if (!inside) return;
  float* o = a.out + 3 * p;
  for (int k = 0; k < 3; ++k) {
    const float c = lit ? acc[k] : 0.0f;
    o[k] = a.accumulate ? o[k] + c : c;
  }
}

}  // namespace

void launch_bruteforce(const drv_config& cfg, const drv_frame& frame, const drv_spot_light& light,
                       const drv_gbuffer& gb, const uint8_t* roughness_metallic, const float* srgb_lut,
                       const Vpl* vals, uint32_t n_vals, int diffuse_only, float* out_rgb, int accumulate,
                       hipStream_t s) {
  BruteArgs a;
  for (int i = 0; i < 16; ++i) a.inv_vp[i] = frame.inverse_view_projection[i];
  for (int i = 0; i < 3; ++i) {
    a.cam[i] = frame.camera_position[i];
    a.light_pos[i] = light.light_position[i];
  }
  a.val_area_factor = light.val_area_factor;
  a.depth = gb.depth;
  a.normal = gb.normal;
  a.albedo = gb.albedo;
  a.roughness_metallic = roughness_metallic;
  a.srgb_lut = srgb_lut;
  a.vals = vals;
  a.n_vals = n_vals;
  a.out = out_rgb;
  a.width = cfg.width;
  a.height = cfg.height;
  a.accumulate = accumulate;
  const dim3 grid((cfg.width + 15) / 16, (cfg.height + 15) / 16);
  if (diffuse_only)
    hipLaunchKernelGGL(k_bruteforce<true>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_bruteforce<false>, grid, dim3(256), 0, s, a);
}

}  // namespace drvk
