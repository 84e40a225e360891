/*
 * drv_oracle.h - CPU restatement of DynamicRadianceVolume's indirect-lighting
 * hot path.  TEST INFRASTRUCTURE ONLY: it is the checker the HIP path is
 * compared against and the CPU baseline the benchmark times.  Nothing in the
 * product links, loads or calls it.
 *
 * Every function restates one reference shader or host routine in its own
 * statement order, in IEEE single precision with no FMA contraction
 * (built with -ffp-contract=off), plus an IEEE double version of the gather
 * used to state the SH error of every FP32 path.  Each function cites the
 * reference file:line it follows (paths relative to
 * DynamicRadianceVolume/).
 *
 * GL semantics that the shaders rely on and that are fixed here:
 *  - image atomics / stores outside the image are no-ops, and an atomic
 *    outside the image returns 0 (GL 4.5 core section 8.26): an allocation
 *    corner outside the address volume therefore "wins" its lock and appends
 *    a cache that no CAV cell points to;
 *  - SSBO reads/writes beyond the buffer read 0 / are dropped (robust buffer
 *    access): caches past max_caches are counted but not stored, and apply
 *    treats such addresses (and unclaimed cells, address 0xFFFFFFFF) as zero;
 *  - texture filtering is evaluated in FP32 with exact weights (the GL 4.5
 *    formulas of section 8.14), LINEAR_MIPMAP_LINEAR level selection, clamp
 *    to edge;
 *  - unorm8 conversion rounds to nearest (ties up); the 2x2x2 voxel mip mean
 *    is computed exactly in integers.
 */
#ifndef DRV_ORACLE_H
#define DRV_ORACLE_H

#include <stdint.h>

#include "../include/drv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- small pure functions -------------------------------------------- */
void orc_sh_constants(drv_sh_constants* out);                          /* renderer.cpp:290-306 */
void orc_morton_decode(uint32_t morton, uint32_t* x, uint32_t* y);     /* cacheLightingRSM.comp:46-62 */
void orc_unpack_normal16i(int16_t a, int16_t b, float out[3]);          /* utils.glsl:43-60 */
void orc_pack_normal16i(const float n[3], int16_t out[2]);              /* utils.glsl:65-89 */
float orc_srgb_to_linear(uint8_t c);                                    /* SRGB8 texture decode */
int orc_cascade(const drv_volume* vol, int32_t n_cascades, const float wp[3]);            /* lightcache.glsl:109-122 */
float orc_cascade_transition(const drv_volume* vol, int32_t cascade, const float wp[3]);  /* lightcache.glsl:125-134 */
/* cacheGather.comp:113-119 / cacheApply.frag:134-135: pixel centre -> world. */
void orc_unproject_pixel(const float inv_vp[16], int32_t width, int32_t height, int32_t x, int32_t y,
                         float depth, float wp_out[3]);
/* renderer.cpp:346-431 */
void orc_update_volume(const drv_config* cfg, const float camera_position[3], const float scene_min[3],
                       const float scene_max[3], drv_volume* out);
/* Scalar fields of renderer.cpp:694-719 (matrices are left untouched). */
void orc_prepare_spot_scalars(const drv_spot_params* p, drv_spot_light* out);

/* ---- stage A: cache allocation ---------------------------------------- */
/* renderer.cpp:951-992 + cacheGather.comp:93-164 + cacheGather.comp:20-91.
 * Clears cav (uint32[(C*R) * R * R]) and writes entries (std430 stride:
 * 16 floats SH1 / 32 floats SH2, capacity cfg->max_caches).  Tiles and
 * threads run in dispatch order.  Returns the demanded cache count
 * (TotalLightCacheCount). */
uint32_t orc_allocate(const drv_config* cfg, const drv_volume* vol, const drv_frame* frame, const float* depth,
                      uint32_t* cav, float* entries);
/* The per-thread leader decision of cacheGather.comp:141-162 for one 16x16
 * tile of cell ids (-1 = none); writes 0/1 flags for both rules. */
void orc_tile_leaders(const int32_t own[256], const int32_t own2[256], uint8_t lead[256], uint8_t lead2[256]);

/* ---- stage B: voxelization -------------------------------------------- */
/* voxelize.vert/.geom/.frag (voxelize.vert:15-23, voxelize.geom:19-112,
 * voxelize.frag:21-57): marks voxels of `target` (uint8[res^3], x fastest)
 * with 255.  positions: float[9 * n_tri] (non-indexed). */
void orc_voxelize(int32_t res, const drv_volume* vol, const float* positions, int64_t n_tri,
                  const float world[16], uint8_t* target);
/* voxelblend.comp:8-18 */
void orc_voxel_blend(int32_t res, const uint8_t* target, uint8_t* volume, float adaption_factor);
/* voxelmipmap.comp:8-12 + voxelization.cpp:154-172: levels packed one after
 * another in `mips` (level 0 first, res^3 bytes, then (res/2)^3 ...);
 * computes levels 1..log2(res) from level 0. */
void orc_voxel_mips(int32_t res, uint8_t* mips);
/* textureLod(VoxelVolume, p, lod) with the linear-clamp sampler bound at
 * renderer.cpp:920-921; single and double precision. */
float orc_sample_voxels_f32(int32_t res, const uint8_t* mips, const float p[3], float lod);
double orc_sample_voxels_f64(int32_t res, const uint8_t* mips, const double p[3], double lod);

/* ---- stage C: gather --------------------------------------------------- */
typedef struct orc_rsm {
  const uint16_t* flux;                /* RGB16F, read_resolution^2 */
  const int16_t* normal;               /* RG16I */
  const uint16_t* depth_lin_sq;        /* RG16F at the read level */
  const uint16_t* depth_lin_sq_shadow; /* RG16F at read level + lod */
  int32_t read_resolution;             /* width of the read level */
  int32_t read_height;                 /* rows: 0 or read_resolution (square), read_resolution / 2 (2:1) */
} orc_rsm;

/* cacheLightingRSM.comp:80-391 (SH1/SH2, INDIRECT_SHADOW) for the caches
 * listed in `cache_ids` (NULL = 0..n-1): adds the light's contribution into
 * `entries` (std430 stride as orc_allocate).  Optional outputs:
 * visibility float[n * n_blocks] (the shadowing term per shadow block) and
 * steps (total cone march steps).  n_threads <= 0 uses all cores. */
void orc_gather_f32(const drv_config* cfg, const drv_volume* vol, const drv_sh_constants* shc,
                    const drv_spot_light* light, const orc_rsm* rsm, const uint8_t* voxel_mips, float* entries,
                    const uint32_t* cache_ids, int64_t n, float* visibility, int64_t* steps, int32_t n_threads);
/* FP64 evaluation of the same sum for cache positions pos[3*n]; writes the
 * 27 SH coefficients per cache (order: SH00, SH1-1, SH10, SH1+1, SH2-2,
 * SH2-1, SH20, SH2+1, SH2+2; each RGB) into sh_out[27*n] (overwrites). */
void orc_gather_f64(const drv_config* cfg, const drv_volume* vol, const drv_sh_constants* shc,
                    const drv_spot_light* light, const orc_rsm* rsm, const uint8_t* voxel_mips, const float* pos,
                    int64_t n, double* sh_out, double* visibility, int32_t n_threads);
/* Cone-traced shadowing of one (cache position, shadow block), FP32
 * (cacheLightingRSM.comp:167-231); returns shadowing, writes the step count. */
float orc_cone_trace_f32(const drv_volume* vol, int32_t voxel_res, const uint8_t* voxel_mips,
                         const drv_spot_light* light, const orc_rsm* rsm, const float wp[3], uint32_t block_morton,
                         int32_t* steps_out);
/* The cone march alone for a given averaged VAL position and aperture
 * (cacheLightingRSM.comp:195-230), FP32. */
float orc_cone_march_f32(const drv_volume* vol, int32_t voxel_res, const uint8_t* voxel_mips, const float wp[3],
                         const float avg_pos[3], float rho, int32_t* steps_out);
/* The LightInfo records of one light (cacheLightingRSM.comp:138-158), in
 * Morton order: out[12 * n_vpl] = position x, y, z, DiscArea, flux r, g, b,
 * 0, normal x, y, z, 0 (the layout the gather sums over). */
void orc_prepare_vpls(const drv_spot_light* light, const orc_rsm* rsm, float* out);
/* ---- surrogate visibility features (no counterpart in the reference) ---- */
/* Shadow blocks of one light: per block the averaged VAL position and cone
 * aperture of cacheLightingRSM.comp:169-192, plus the normalized mean VPL
 * normal of its B x B texels; out[8 * n_blocks] = x, y, z, rho, nx, ny, nz, 0. */
void orc_shadow_blocks(const drv_spot_light* light, const orc_rsm* rsm, float* out);
/* The 32 network inputs of the pair (cache at wp, block blk[8]), in the
 * network's slot order (slot 8g + e: e < 5 geometry feature 5g + e, else
 * voxel sample 3g + e - 5). */
void orc_mlp_features(const drv_volume* vol, int32_t voxel_res, const uint8_t* voxel_mips, const float cam[3],
                      const float wp[3], const float blk[8], float out[32]);

/* Extracts the 27 SH coefficients (canonical order above) from an entry. */
void orc_entry_sh(const float* entry, int32_t sh_bands, float out27[27]);

/* ---- indirect specular (INDIRECT_SPECULAR, DIRECT_SPECULAR_MAP_WRITE) ---- */
/* GL unsigned 11/10-bit float conversion (GL 4.5 section 2.3.4.3-4): round
 * to the nearest representable finite value, ties to even; negative -> 0,
 * finite overflow -> largest finite, +inf -> inf, NaN -> NaN.  m = mantissa
 * bits (6 or 5).  Written by nearest-neighbour search in double. */
uint32_t orc_uf_encode(float x, int32_t m);
float orc_uf_decode(uint32_t code, int32_t m);
/* GL_UNSIGNED_INT_10F_11F_11F_REV packing: red bits 0-10, green 11-21, blue 22-31. */
uint32_t orc_pack_r11g11b10(const float rgb[3]);
void orc_unpack_r11g11b10(uint32_t v, float rgb[3]);
/* HemisphericalProjection, concentric quad (utils.glsl:102-178). */
void orc_hemi_projection(const float dir[3], float out[2]);
/* ComputeLocalViewSpace (lightcache.glsl:98-106): the columns X, Y, Z. */
void orc_local_view_space(const float cam[3], const float wp[3], float x[3], float y[3], float z[3]);
/* Atlas of renderer.cpp:253-286 for cfg (specular_envmap_size 0 = 16). */
void orc_specular_atlas_geometry(const drv_config* cfg, int32_t* width, int32_t* per_row, int32_t* levels);
/* cacheLightingRSM.comp:87-99, 122-124, 240-335 for caches at pos[3*n]:
 * each cache's (S+1)^2 RGB map maps[n][(S+1)^2][3] (x fastest) of the light's
 * reflected radiance before the base normalization, VALs in the shader's
 * order; texel index S is the column / row that the reference's store puts
 * into the next cache of the atlas.  visibility[n][n_blocks] (as
 * orc_gather_f32 writes it) is the shadowing per block; NULL = unshadowed. */
void orc_specular_gather(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                         const orc_rsm* rsm, const float* pos, int64_t n, const float* visibility, float* maps,
                         int32_t n_threads);
/* The atlas writes of one light for the n live caches: adds the in-map
 * texels into the FP32 sums work[n][S*S][3] (accumulate = 0 overwrites, the
 * frame's first light), routes index-S texels to the left/upper/upper-left
 * atlas neighbours' edge sums, then per cache: level 0 = R11G11B10F of
 * sum * BlinnNormalization(S^2 - 1), the mips of specularenvmap_mipmap.frag
 * and the hole fill of specularenvmap_fillholes.frag
 * (renderer.cpp:994-1045), packed into atlas (levels back to back). */
void orc_specular_finalize(const drv_config* cfg, int64_t n, const float* maps, float* work, int32_t accumulate,
                           uint32_t* atlas);

/* textureLod(CacheSpecularEnvmap, (s, t), lod) on the atlas with the
 * linear-clamp sampler (LINEAR_MIPMAP_LINEAR, clamp to edge). */
void orc_specular_sample(const drv_config* cfg, const uint32_t* atlas, float s, float t, float lod, float out[3]);

typedef struct orc_specular {
  const uint32_t* atlas;             /* as orc_specular_finalize writes it */
  const uint8_t* roughness_metallic; /* RG8 or NULL (roughness 1, metallic 0) */
} orc_specular;

/* ---- stage D: apply ----------------------------------------------------- */
/* cacheApply.frag:28-195 + lightcache.glsl:137-183: out_rgb float[3*W*H];
 * discarded pixels (depth < 1e-5) get 0.  spec (NULL = none) adds the
 * INDIRECT_SPECULAR lookup when cfg->indirect_specular is set; caches past
 * live_count contribute no specular. */
void orc_apply(const drv_config* cfg, const drv_volume* vol, const drv_frame* frame, const drv_sh_constants* shc,
               const float* depth, const int16_t* normal, const uint8_t* albedo, const uint32_t* cav,
               const float* entries, uint32_t live_count, const orc_specular* spec, float* out_rgb,
               int32_t n_threads);

/* ---- brute-force RSM reference ----------------------------------------- */
/* bruteforcersm.frag:15-88 with BRDF of lightingfunctions.glsl:46-58 for the
 * pixels listed in `pixels` (y * width + x; NULL = all), VALs visited in the
 * shader's order (x outer, y inner); roughness_metallic (RG8) may be NULL
 * (roughness 1, metallic 0); diffuse_only drops the Blinn-Phong term.
 * Pixels with depth < 1e-5 get 0.  out_rgb[3 * n]. */
void orc_bruteforce_rsm(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                        const orc_rsm* rsm, const float* depth, const int16_t* normal, const uint8_t* albedo,
                        const uint8_t* roughness_metallic, int32_t diffuse_only, const uint32_t* pixels, int64_t n,
                        float* out_rgb, int32_t n_threads);

/* ---- direct lighting and HDR output -------------------------------------- */
/* directdeferredlighting.frag:13-52 with the LINEAR compare-GREATER PCF of
 * the shadow sampler (renderer.cpp:33-34, border depth 0); out_rgb[3*W*H],
 * background pixels 0.  roughness_metallic may be NULL (1, 0). */
void orc_direct_lighting(const drv_config* cfg, const drv_frame* frame, const drv_spot_light* light,
                         const float* depth, const int16_t* normal, const uint8_t* albedo,
                         const uint8_t* roughness_metallic, const float* shadow_map, int32_t shadow_res,
                         float normal_offset_bias, float shadow_bias, float* out_rgb, int32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif
