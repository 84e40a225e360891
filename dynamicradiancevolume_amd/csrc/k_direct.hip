// Direct lighting of a spot light and the HDR output pass, so that frames
// can be composed on the device: shader/directdeferredlighting.frag:13-52
// (Renderer::ApplyDirectLighting, rendering/renderer.cpp:827-849) with the
// light's RSM depth as its shadow map, and shader/tonemapping.frag:21-34
// (Renderer::OutputHDRTextureToBackbuffer, rendering/renderer.cpp:745-757).
#include "drv_internal.hpp"
#include "k_common.hpp"

namespace drvk {
namespace {

struct DirectArgs {
  float inv_vp[16];
  float lvp[16];
  float cam[3];
  float light_pos[3], light_dir[3], intensity[3];
  float cos_half;
  float normal_offset, bias;
  const float* depth;
  const int16_t* normal;
  const uint8_t* albedo;
  const uint8_t* roughness_metallic;
  const float* srgb_lut;
  const float* shadow;  // reverse-Z depth of the light's view, res x res
  int shadow_res;
  float* out;
  int width, height, accumulate;
};

// textureProjLod on a sampler2DShadow with LINEAR filtering, compare GREATER
// and a border depth of 0 (the shadow sampler of renderer.cpp:33-34): the
// four texel comparisons weighted bilinearly.
__device__ __forceinline__ float shadow_pcf(const DirectArgs& a, float s, float t, float ref) {
  const int w = a.shadow_res;
  const float u = s * static_cast<float>(w) - 0.5f, v = t * static_cast<float>(w) - 0.5f;
  const float fu = floorf(u), fv = floorf(v);
  const int i0 = to_int(fu), j0 = to_int(fv);
  const float al = u - fu, be = v - fv;
  auto lit = [&](int i, int j) {
    const float d = (i >= 0 && j >= 0 && i < w && j < w) ? a.shadow[static_cast<size_t>(j) * w + i] : 0.0f;
    return ref > d ? 1.0f : 0.0f;
  };
  return (1.0f - al) * (1.0f - be) * lit(i0, j0) + al * (1.0f - be) * lit(i0 + 1, j0) +
         (1.0f - al) * be * lit(i0, j0 + 1) + al * be * lit(i0 + 1, j0 + 1);
}

__global__ __launch_bounds__(256) void k_direct(DirectArgs a) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= a.width || y >= a.height) return;
  const size_t p = static_cast<size_t>(y) * a.width + x;
  float* o = a.out + 3 * p;
  float col[3] = {0.0f, 0.0f, 0.0f};
  const float d = a.depth[p];
  if (d >= 0.00001f) {  // background pixels receive no light
    const f3 wp = unproject_pixel(a.inv_vp, a.width, a.height, x, y, d);
    const f3 n = unpack_normal16i(a.normal[2 * p], a.normal[2 * p + 1]);
    // directdeferredlighting.frag:20-25
    float lx = a.light_pos[0] - wp.x, ly = a.light_pos[1] - wp.y, lz = a.light_pos[2] - wp.z;
    const float dsq = lx * lx + ly * ly + lz * lz;
    const float inv = 1.0f / sqrtf(dsq);
    lx *= inv;
    ly *= inv;
    lz *= inv;
    const float cos_theta = fminf(1.0f, fmaxf(0.0f, lx * n.x + ly * n.y + lz * n.z));
    // directdeferredlighting.frag:27-36: normal-offset shadow lookup
    const float off = a.normal_offset * (1.0f - cos_theta);
    const float ox = wp.x + off * n.x, oy = wp.y + off * n.y, oz = wp.z + off * n.z;
    float po[4], pw[4];
    mat_vec(a.lvp, ox, oy, oz, 1.0f, po);
    mat_vec(a.lvp, wp.x, wp.y, wp.z, 1.0f, pw);
    const float sx = po[0] * 0.5f + 0.5f * pw[3], sy = po[1] * 0.5f + 0.5f * pw[3];
    const float sz = pw[2] + a.bias;
    const float shadowing = shadow_pcf(a, sx / pw[3], sy / pw[3], sz / pw[3]);
    if (shadowing != 0.0f) {  // directdeferredlighting.frag:38-39: discard
      float cx = a.cam[0] - wp.x, cy = a.cam[1] - wp.y, cz = a.cam[2] - wp.z;
      const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
      cx /= cl;
      cy /= cl;
      cz /= cl;
      float rough = 1.0f, metal = 0.0f;
      if (a.roughness_metallic) {
        rough = static_cast<float>(a.roughness_metallic[2 * p]) / 255.0f;
        metal = static_cast<float>(a.roughness_metallic[2 * p + 1]) / 255.0f;
      }
      // lightingfunctions.glsl:3-10 spot falloff, 46-58 BRDF
      const float falloff =
          fminf(1.0f, fmaxf(0.0f, -(lx * a.light_dir[0] + ly * a.light_dir[1] + lz * a.light_dir[2]) - a.cos_half)) /
          (1.0f - a.cos_half);
      const float rsq = rough * rough;
      const float blinn = 2.0f / (rsq * rsq + 0.0005f);
      const float norm = (blinn + 8.0f) / (8.0f * kPi);
      float hx = cx + lx, hy = cy + ly, hz = cz + lz;
      const float hl = sqrtf(hx * hx + hy * hy + hz * hz);
      hx /= hl;
      hy /= hl;
      hz /= hl;
      const float nh = fminf(1.0f, fmaxf(0.0f, n.x * hx + n.y * hy + n.z * hz));
      const float bp = norm * powf(nh, blinn);
      const float k = shadowing * falloff * cos_theta / dsq;
      for (int c = 0; c < 3; ++c) {
        const float base = a.srgb_lut[a.albedo[3 * p + c]];
        const float diff = base * (1.0f - metal) + 0.02f * metal;
        const float spec = 0.04f * (1.0f - metal) + base * metal;
        col[c] = (diff / kPi + bp * spec) * (a.intensity[c] * k);
      }
    }
  }
  for (int c = 0; c < 3; ++c) o[c] = a.accumulate ? o[c] + col[c] : col[c];
}

// tonemapping.frag:21-34: Drago operator log2(c * exposure + 1) / log2(LMax + 1).
__global__ __launch_bounds__(256) void k_tonemap(const float* hdr, size_t n, float exposure, float divider,
                                                 float* out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = log2f(hdr[i] * exposure + 1.0f) / divider;
}

}  // namespace

void launch_direct(const drv_config& cfg, const drv_frame& frame, const drv_spot_light& light, const drv_gbuffer& gb,
                   const uint8_t* roughness_metallic, const float* srgb_lut, const float* shadow, int shadow_res,
                   float normal_offset, float bias, float* out_rgb, int accumulate, hipStream_t s) {
  DirectArgs a;
  for (int i = 0; i < 16; ++i) {
    a.inv_vp[i] = frame.inverse_view_projection[i];
    a.lvp[i] = light.light_view_projection[i];
  }
  for (int i = 0; i < 3; ++i) {
    a.cam[i] = frame.camera_position[i];
    a.light_pos[i] = light.light_position[i];
    a.light_dir[i] = light.light_direction[i];
    a.intensity[i] = light.light_intensity[i];
  }
  a.cos_half = light.light_cos_half_angle;
  a.normal_offset = normal_offset;
  a.bias = bias;
  a.depth = gb.depth;
  a.normal = gb.normal;
  a.albedo = gb.albedo;
  a.roughness_metallic = roughness_metallic;
  a.srgb_lut = srgb_lut;
  a.shadow = shadow;
  a.shadow_res = shadow_res;
  a.out = out_rgb;
  a.width = cfg.width;
  a.height = cfg.height;
  a.accumulate = accumulate;
  const dim3 grid((cfg.width + 15) / 16, (cfg.height + 15) / 16);
  hipLaunchKernelGGL(k_direct, grid, dim3(256), 0, s, a);
}

void launch_tonemap(const float* hdr, size_t n, float exposure, float divider, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tonemap, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, hdr, n, exposure,
                     divider, out);
}

}  // namespace drvk
