/*
 * drv.h - C ABI of the MI355X indirect-lighting pipeline.
 *
 * The four per-frame compute passes of DynamicRadianceVolume's renderer
 * (Renderer::Draw, DynamicRadianceVolume/rendering/renderer.cpp:539-594):
 *
 *   VoxelizeScene -> AllocateCaches -> LightCachesRSM -> ApplyCaches
 *
 * are exported here as plain C entry points over device pointers.  The
 * reference keeps these passes private to `Renderer` and binds its data as GL
 * state (UBOs 0-4, SSBOs 0-1, image unit 0, texture units 0-4); each struct
 * below restates one of those GL bindings as a plain C record.
 *
 * Conventions (shared by every entry point):
 *  - Every function returns an int status: DRV_OK (0) or a negative DRV_ERR_*
 *    code.  Nothing throws across the ABI.  drv_last_error() returns a text
 *    description of the last failure on a context.
 *  - Image/buffer pointers marked "device" are HIP device pointers that stay
 *    owned by the caller.  The library owns the cache buffer, the cache address
 *    volume, the voxel volume and all scratch buffers.
 *  - Matrices are 16 floats, row-major, applied to column vectors:
 *    out[r] = sum_c m[4*r + c] * v[c].  This is what the reference's GLSL
 *    expression `vec4 * mat4` computes on a row-major ei::Mat4x4 upload.
 *  - 2D images are row-major with row 0 at normalized texture coordinate
 *    v = 0 (the GL texture origin, NDC y = -1).
 *  - A context is single-threaded; all work is enqueued on one HIP stream
 *    (drv_set_stream) and only the drv_read_* / drv_get_cache_count calls
 *    synchronize with the host.
 */
#ifndef DRV_H
#define DRV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRV_OK 0
#define DRV_ERR_INVALID_ARG (-1)
#define DRV_ERR_HIP (-2)
#define DRV_ERR_NO_MEMORY (-3)
#define DRV_ERR_STATE (-4)

#define DRV_MAX_CASCADES 4 /* MAX_NUM_ADDRESS_VOLUME_CASCADES, shader/globalubos.glsl:46 */

typedef struct drv_ctx drv_ctx;

/* Static configuration.  Replaces the Renderer constructor defaults
 * (rendering/renderer.cpp:36-51, 86-90), SetMaxCacheCount
 * (renderer.cpp:248-288), SetCAVCascades (renderer.cpp:1174-1193) and the
 * shader #define prefix of ReloadLightingSettingDependentCacheShader
 * (renderer.cpp:197-246). */
typedef struct drv_config {
  int32_t width;               /* backbuffer width in pixels */
  int32_t height;              /* backbuffer height in pixels */
  int32_t max_caches;          /* capacity of the light cache buffer */
  int32_t cav_resolution;      /* cache address volume cells per axis and cascade */
  int32_t cav_cascades;        /* 1..DRV_MAX_CASCADES */
  int32_t voxel_resolution;    /* voxel volume cells per axis, power of two in [2, 512] */
  int32_t sh_bands;            /* 1 = INDDIFFUSE_VIA_SH1 (64 B/cache), 2 = INDDIFFUSE_VIA_SH2 (128 B/cache) */
  int32_t indirect_shadow;     /* INDIRECT_SHADOW: cone-traced visibility on/off */
  float cav_transition_size;   /* CAVTransitionZoneSize; > 0 defines ADDRESSVOL_CASCADE_TRANSITIONS */
  float cav_cascade0_world_size; /* world size of cascade 0 (renderer.cpp:1185: 4.0); each further cascade doubles */
  /* INDIRECT_SPECULAR (SetIndirectSpecular, renderer.hpp:80; off by default,
   * renderer.cpp:49) with DIRECT_SPECULAR_MAP_WRITE: per-cache hemisphere
   * env-maps in an R11G11B10F atlas, read by drv_apply_caches. */
  int32_t indirect_specular;
  int32_t specular_envmap_size;  /* SetPerCacheSpecularEnvMapSize (renderer.cpp:454-460): power of two
                                    1..32; 0 selects the reference default 16.  The atlas side
                                    (see drv_get_specular_atlas) must stay <= 32768 texels. */
  int32_t specular_hole_fill_level; /* SetSpecularEnvMapHoleFillLevel (renderer.hpp:97): clamped to
                                       log2(size); 0 = no hole fill (default) */
} drv_config;

/* One CAV cascade, struct CAVCascade (shader/globalubos.glsl:48-60). */
typedef struct drv_cascade {
  float min[3];
  float world_voxel_size;
  float max[3];
  float pad0;
  float decision_min[3];
  float pad1;
  float decision_max[3];
  float pad2;
} drv_cascade;

/* UBO VolumeInfo (shader/globalubos.glsl:65-79). */
typedef struct drv_volume {
  float volume_world_min[3];
  float voxel_size_in_world;
  float volume_world_max[3];
  float cav_transition_zone_size;
  drv_cascade cascades[DRV_MAX_CASCADES];
} drv_volume;

/* The fields of UBO PerFrame (shader/globalubos.glsl:33-43) the path reads. */
typedef struct drv_frame {
  float inverse_view_projection[16];
  float camera_position[3];
} drv_frame;

/* UBO Constant SH factors (shader/globalubos.glsl:4-16), filled as in
 * Renderer::UpdateConstantUBO (renderer.cpp:290-306). */
typedef struct drv_sh_constants {
  float cos_lobe0, cos_lobe1, cos_lobe2n2_p1_n1, cos_lobe20, cos_lobe2p2;
  float eva0, eva1, eva2n2_p1_n1, eva20, eva2p2;
} drv_sh_constants;

/* Spot light description, struct Light (scene/light.hpp:8-54). */
typedef struct drv_spot_params {
  float intensity[3];
  float position[3];
  float direction[3];          /* need not be normalized */
  float half_angle;            /* radians */
  uint32_t rsm_resolution;     /* rounded up to a power of two */
  uint32_t rsm_read_lod;
  uint32_t indirect_shadow_lod;
  float near_plane;            /* Light::nearPlane = 0.1 (scene/scene.cpp:6) */
  float far_plane;             /* Light::farPlane = 10000 (scene/scene.cpp:7) */
  /* 1: a 2:1 RSM (render and read levels twice as wide as tall, texels of
   * the square light frustum twice as tall as wide), e.g. 256 x 128 = 32768
   * VPLs at the read level; 0: the reference's square RSM.  No counterpart
   * in the reference, whose Morton order (cacheLightingRSM.comp:142) is
   * defined for square reads: the decode of m < 2 h^2 covers the 2h x h
   * texels exactly once (x takes the extra bit). */
  uint32_t rsm_wide;
} drv_spot_params;

/* UBO SpotLight (shader/globalubos.glsl:88-112), as filled by
 * Renderer::PrepareLights (renderer.cpp:664-725). */
typedef struct drv_spot_light {
  float light_intensity[3];
  float light_position[3];
  float light_direction[3];
  float light_cos_half_angle;
  float light_view_projection[16];
  float inverse_light_view_projection[16];
  int32_t rsm_render_resolution;
  int32_t rsm_read_resolution;
  float val_area_factor;
  float indirect_shadow_lod;
  float indirect_shadow_block_size;
  int32_t indirect_shadow_sample_interval;
  float indirect_shadow_super_val_width;
  float indirect_shadow_sampling_offset;
  int32_t rsm_render_height;  /* = rsm_render_resolution (square) or half of it (rsm_wide) */
  int32_t rsm_read_height;    /* = rsm_read_resolution (square) or half of it (rsm_wide) */
} drv_spot_light;

/* G-buffer (renderer.cpp:466-471, texture units 0-3 of shader/gbuffer.glsl). */
typedef struct drv_gbuffer {
  const float* depth;     /* device, DEPTH_COMPONENT32F, reverse-Z (clear 0) */
  const int16_t* normal;  /* device, RG16I packed by PackNormal16I (shader/utils.glsl:83-89) */
  const uint8_t* albedo;  /* device, SRGB8, 3 bytes per pixel, sRGB-encoded */
  int32_t width;
  int32_t height;
  /* device, RG8 (roughness, metallic), texture unit 3 (GBuffer_RoughnessMetalic);
   * read by drv_apply_caches with indirect_specular only.  NULL reads as
   * roughness 1, metallic 0. */
  const uint8_t* roughness_metallic;
} drv_gbuffer;

/* Reflective shadow map of one light at its read level (texture units 0-2 of
 * shader/cacheLightingRSM.comp:11-13; the base level is set to rsmReadLod at
 * renderer.cpp:807-809). */
typedef struct drv_rsm {
  const uint16_t* flux;              /* device, RGB16F, read_resolution^2 texels (flux / pi) */
  const int16_t* normal;             /* device, RG16I */
  const uint16_t* depth_lin_sq;      /* device, RG16F (d, d^2) at the read level */
  const uint16_t* depth_lin_sq_shadow; /* device, RG16F (d, d^2) at read level + indirect_shadow_lod */
  int32_t read_resolution;           /* power of two; RSMReadResolution (the width) */
  int32_t read_height;               /* rows of the read level: read_resolution (square), read_resolution / 2
                                        (a 2:1 RSM), or 0 for square */
} drv_rsm;

/* Per-stage timings of the last frame when profiling is on (the scopes of
 * the reference's FrameProfiler: VoxelizeScene, VoxelBlendMipMap,
 * AllocateCaches, LightCaches, ApplyCaches). */
typedef struct drv_stage_times {
  float voxelize_ms;     /* target clear + raster of all meshes */
  float voxel_mips_ms;   /* blend + mip chain */
  float allocate_ms;     /* counter/CAV clear + allocation + dispatch args */
  float light_caches_ms; /* whole LightCachesRSM of the last light */
  float prepare_ms;      /* VPL and shadow-block records */
  float visibility_ms;   /* cone-traced visibility kernel */
  float gather_ms;       /* SH gather kernel */
  float reduce_ms;       /* split reduction into the cache entries */
  float apply_ms;
  uint64_t cone_traces;  /* (cache, shadow block) traces of the last light, when counting */
  uint64_t cone_steps;   /* march steps the reference takes for those traces, when counting */
  uint64_t cone_steps_executed; /* steps actually sampled (rays stop once opaque) */
  float specular_ms;     /* specular env-map gather + mips/hole fill (ProcessSpecularEnvmap), inside light_caches_ms */
} drv_stage_times;

/* ---- lifetime -------------------------------------------------------- */

/* Renderer::Renderer + AllocateCacheData + SetCAVCascades
 * (renderer.cpp:26-120, 248-288, 1174-1193). */
int drv_create(const drv_config* config, drv_ctx** out_ctx);
int drv_destroy(drv_ctx* ctx);
const char* drv_last_error(const drv_ctx* ctx);
/* Sets the HIP stream (hipStream_t) all work is enqueued on; NULL = default stream. */
int drv_set_stream(drv_ctx* ctx, void* hip_stream);
int drv_synchronize(drv_ctx* ctx);

/* ---- per-frame state (the UBOs) -------------------------------------- */

/* UpdateConstantUBO (renderer.cpp:290-306): the SH factors in use. */
int drv_get_sh_constants(const drv_ctx* ctx, drv_sh_constants* out);
/* UpdatePerFrameUBO (renderer.cpp:324-344). */
int drv_set_frame(drv_ctx* ctx, const drv_frame* frame);
/* UpdateVolumeUBO (renderer.cpp:346-431): voxel volume from the scene bounding
 * box, CAV cascades centred on the camera. */
int drv_update_volume(drv_ctx* ctx, const float camera_position[3],
                      const float scene_min[3], const float scene_max[3]);
int drv_set_volume(drv_ctx* ctx, const drv_volume* volume);
int drv_get_volume(const drv_ctx* ctx, drv_volume* out);
/* PrepareLights (renderer.cpp:664-725): SpotLight UBO of one light.  Pure host
 * function, no context needed. */
int drv_prepare_spot_light(const drv_spot_params* params, drv_spot_light* out);

/* ---- the four stages -------------------------------------------------- */

/* Voxelization::VoxelizeScene (rendering/voxelization.cpp:90-176).
 * begin clears the voxel target; mesh rasterizes one entity's triangles
 * (positions: device float[3 * n_vertices]; indices: device uint32[3 * n_tri],
 * or NULL for non-indexed float[9 * n_tri] positions) through `world`;
 * end blends the target into the volume with `adaption_factor`
 * (voxelblend.comp) and rebuilds the mip chain (voxelmipmap.comp). */
int drv_voxelize_begin(drv_ctx* ctx);
/* Off by default: a voxelization starts after all work already enqueued on
 * the stream.  On: drv_voxelize_begin starts it right after the last call
 * that read the previous voxels (the previous light pass), so it runs beside
 * the rest of that frame (SH gather, reduction, apply).  The caller then must
 * not write, on the stream after that light pass, the positions / indices the
 * following drv_voxelize_mesh calls read (static geometry, or geometry written
 * before the light pass, qualifies).  Graph capture always waits for the
 * whole stream.  No counterpart in the reference (GL orders the passes). */
int drv_set_voxel_overlap(drv_ctx* ctx, int32_t on);
int drv_voxelize_mesh(drv_ctx* ctx, const float* positions, const uint32_t* indices,
                      uint32_t n_tri, const float world[16]);
int drv_voxelize_end(drv_ctx* ctx, float adaption_factor);

/* Renderer::AllocateCaches (renderer.cpp:951-992) = clear of counter and CAV,
 * shader/cacheGather.comp and shader/cachePrepareLighting.comp. */
int drv_allocate_caches(drv_ctx* ctx, const drv_gbuffer* gbuffer);

/* Renderer::LightCachesRSM (renderer.cpp:899-933) for one light:
 * shader/cacheLightingRSM.comp.  Results accumulate (+=) over calls. */
int drv_light_caches(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light);

/* drv_light_caches without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It lights the caches with the fused work-queue kernel (each (cache, shadow
 * block) pair cone-traced inside the VPL sum, as drv_batch_light_caches does)
 * and keeps the VPL and shadow-block records, its queue and its split partials
 * in one static device array of 32 MiB per GPU, held from the first context on
 * that GPU until the process ends.  Limits: at most 65536 VPLs (RSM read
 * resolution <= 256; DRV_ERR_INVALID_ARG beyond); no indirect specular and no
 * surrogate (DRV_ERR_STATE).  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (other streams, other threads) race. */
int drv_light_caches_noalloc(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light);

/* Renderer::ApplyCaches (renderer.cpp:1047-1079): shader/cacheApply.frag.
 * out_rgb: device float[3 * width * height]; accumulate != 0 adds to it
 * (the reference's GL_ONE/GL_ONE blend), 0 overwrites covered pixels and
 * writes 0 to discarded ones. */
int drv_apply_caches(drv_ctx* ctx, const drv_gbuffer* gbuffer, float* out_rgb, int32_t accumulate);

/* ---- indirect specular env-map atlas ----------------------------------- */

/* Geometry of the specular atlas (renderer.cpp:253-286): level-0 width in
 * texels (2^ceil(log2(ceil(sqrt(max_caches)) * size))), caches per atlas row
 * (cache i sits at ((i % per_row) * size, (i / per_row) * size)) and mip
 * levels (log2(size) + 1).  DRV_ERR_STATE when indirect_specular is off. */
int drv_get_specular_atlas(drv_ctx* ctx, int32_t* width, int32_t* caches_per_row, int32_t* levels);
/* Host copy of one atlas level: (width >> level)^2 packed texels
 * (GL_UNSIGNED_INT_10F_11F_11F_REV: red bits 0-10, green 11-21, blue 22-31),
 * row-major.  Only the regions of the live caches of the last light pass
 * are defined. */
int drv_read_specular_atlas(drv_ctx* ctx, int32_t level, uint32_t* host_out);

/* ---- introspection (synchronous) --------------------------------------- */

/* TotalLightCacheCount (shader/lightcache.glsl:89; read back at
 * renderer.cpp:960-966).  demanded = appends attempted; live = min(demanded,
 * max_caches).  demanded > max_caches means the buffer overflowed. */
int drv_get_cache_count(drv_ctx* ctx, uint32_t* demanded, uint32_t* live);
/* Copies the first `count` LightCacheEntry records (std430,
 * shader/lightcache.glsl:33-76: 16 floats per cache for SH1, 32 for SH2). */
int drv_read_caches(drv_ctx* ctx, float* host_out, uint32_t count);
/* Copies the CAV atlas, uint32[(cascades * R) * R * R], x fastest. */
int drv_read_cav(drv_ctx* ctx, uint32_t* host_out);
/* Copies one mip level of the voxel volume, uint8[res_l^3], x fastest. */
int drv_read_voxels(drv_ctx* ctx, int32_t level, uint8_t* host_out);
/* Copies the per-(cache, shadow block) visibility of the last
 * drv_light_caches call, float[live * n_blocks], cache-major. */
int drv_read_visibility(drv_ctx* ctx, float* host_out, uint64_t count);

/* ---- inputs rendered on the device ------------------------------------ */

/* Scene geometry for the device-side raster passes: a triangle soup on the
 * device, positions[n_tri][9] (world space), albedo[n_tri][3] (linear) and
 * normals[n_tri][3] per triangle. */
typedef struct drv_scene {
  const float* positions;
  const float* albedo;
  const float* normals;
  uint32_t n_tri;
} drv_scene;

/* G-buffer fill (shader/fillgbuffer.frag; Renderer::Draw's G-buffer pass,
 * rendering/renderer.cpp:539-560) into library-owned buffers of the
 * configured width x height; *out receives their device pointers. */
int drv_render_gbuffer(drv_ctx* ctx, const drv_scene* scene, const float view_projection[16], drv_gbuffer* out);
/* RSM of one spot light (Renderer::PrepareRSM, rendering/renderer.cpp:1300-
 * 1339: shader/fillrsm.frag at the render resolution, then the
 * shader/downsamplersm.frag mip chain) into library-owned buffers; *out
 * receives the read level (rsm_read_lod) and the indirect-shadow level, as
 * drv_light_caches expects them. */
int drv_render_rsm(drv_ctx* ctx, const drv_scene* scene, const drv_spot_light* light, drv_rsm* out);
/* Host copies of the last rendered G-buffer (width x height texels). */
int drv_read_gbuffer(drv_ctx* ctx, float* depth, int16_t* normal, uint8_t* albedo);
/* Host copy of mip `level` of the last rendered RSM (flux RGB16F,
 * normal RG16I, depth/depth^2 RG16F; (render resolution >> level)^2 texels). */
int drv_read_rsm(drv_ctx* ctx, int32_t level, uint16_t* flux, int16_t* normal, uint16_t* depth_lin_sq);

/* The depth buffer of the last drv_render_rsm (reverse-Z, resolution^2
 * floats, device): the light's shadow map for drv_direct_lighting. */
int drv_get_shadow_map(drv_ctx* ctx, const float** shadow_map, int32_t* resolution);
/* Host copy of that depth buffer (resolution^2 floats). */
int drv_read_shadow_map(drv_ctx* ctx, float* host_out);

/* ---- direct lighting and HDR output -------------------------------------- */

/* Renderer::ApplyDirectLighting (rendering/renderer.cpp:827-849,
 * shader/directdeferredlighting.frag): one spot light with a normal-offset
 * PCF shadow lookup (LINEAR compare-GREATER sampler, border depth 0) into
 * its shadow map; normal_offset_bias / shadow_bias are Light::
 * normalOffsetShadowBias / shadowBias (0.01 / 0.0001, scene/light.hpp:13).
 * Background pixels receive 0. */
int drv_direct_lighting(drv_ctx* ctx, const drv_gbuffer* gbuffer, const uint8_t* roughness_metallic,
                        const drv_spot_light* light, const float* shadow_map, int32_t shadow_resolution,
                        float normal_offset_bias, float shadow_bias, float* out_rgb, int32_t accumulate);
/* shader/tonemapping.frag: out = log2(hdr * exposure + 1) / log2(lmax + 1)
 * per channel (Renderer defaults: exposure 1, LMax 1.2; renderer.cpp:38-39);
 * device buffers of width x height x 3 floats. */
int drv_tonemap(drv_ctx* ctx, const float* hdr_rgb, float exposure, float lmax, float* out_rgb);
/* Renderer::SaveToPFM (rendering/renderer.cpp:1229-1235, hdrimage.cpp:6-30):
 * writes the device HDR image (width x height x 3 floats) as a PFM file. */
int drv_save_pfm(drv_ctx* ctx, const float* hdr_rgb, const char* path);

/* ---- brute-force RSM reference ----------------------------------------- */

/* Renderer::ApplyRSMsBruteForce (rendering/renderer.cpp:851-874,
 * shader/bruteforcersm.frag): every pixel sums the light of every VAL of the
 * RSM with the BRDF of shader/lightingfunctions.glsl:46-58, unshadowed - the
 * reference's ground truth for the cache path.  roughness_metallic is the
 * RG8 G-buffer texture (NULL: roughness 1, metallic 0); diffuse_only = 1
 * keeps only the diffuse term (what the caches represent).  Background
 * pixels (depth < 1e-5) receive 0. */
int drv_bruteforce_rsm(drv_ctx* ctx, const drv_gbuffer* gbuffer, const uint8_t* roughness_metallic,
                       const drv_rsm* rsm, const drv_spot_light* light, int32_t diffuse_only, float* out_rgb,
                       int32_t accumulate);

/* ---- surrogate visibility -------------------------------------------- */

/* A 32-64-64-64-1 MLP (ReLU hidden layers, sigmoid output; BF16 operands,
 * FP32 accumulation) that predicts the cone-traced shadowing of a (cache,
 * shadow block) pair from 32 features: cache and VAL geometry and 12
 * voxel-mip occupancy samples along the segment.  It has no counterpart in
 * the reference; when enabled it replaces the cone trace of
 * shader/cacheLightingRSM.comp:167-231 inside drv_light_caches. */
#define DRV_MLP_PARAMS 10497 /* W1[32][64], b1[64], W2[64][64], b2, W3[64][64], b3, w4[64], b4 */
#define DRV_MLP_FEATURES 32

typedef struct drv_surrogate {
  int32_t enabled;          /* 1: drv_light_caches predicts visibility with the MLP */
  float spot_fraction;      /* fraction of caches cone-traced each light as a spot check (0 = off) */
  float fallback_threshold; /* spot-check SH relative L2 error above which the light is re-gathered with cone tracing */
} drv_surrogate;

typedef struct drv_surrogate_stats {
  float spot_error;     /* SH relative L2 error of the MLP on the spot-check caches (last light) */
  int32_t fell_back;    /* 1 if the last light fell back to cone tracing */
  uint32_t spot_caches; /* caches in the spot check */
} drv_surrogate_stats;

/* Network parameters (FP32 master copy, layout of DRV_MLP_PARAMS).  Setting
 * them resets the optimizer state. */
int drv_mlp_set_params(drv_ctx* ctx, const float* host_params, uint32_t count);
int drv_mlp_get_params(drv_ctx* ctx, float* host_params, uint32_t count);
int drv_set_surrogate(drv_ctx* ctx, const drv_surrogate* settings);
int drv_get_surrogate_stats(drv_ctx* ctx, drv_surrogate_stats* out);
/* Training samples from the last cone-traced drv_light_caches call: n random
 * (cache, shadow block) pairs -> features x_dev[n][32] (16-byte aligned) and
 * the cone-traced visibility y_dev[n] (device pointers). */
int drv_mlp_collect(drv_ctx* ctx, uint32_t n, uint32_t seed, float* x_dev, float* y_dev);
/* Features of explicit (cache, block) pairs (pairs_dev[n][2], device) of the
 * last drv_light_caches call into x_dev[n][32] (16-byte aligned); a pair
 * naming no live cache or block gets NaN features. */
int drv_mlp_features(drv_ctx* ctx, const uint32_t* pairs_dev, uint32_t n, float* x_dev);
/* Network output for features x_dev[n][32] -> p_dev[n] (device). */
int drv_mlp_forward(drv_ctx* ctx, const float* x_dev, uint32_t n, float* p_dev);
/* Gradient of the mean binary cross-entropy over all n samples, computed by
 * the training kernels: grad_host[DRV_MLP_PARAMS + 1], the last entry being
 * the summed loss. */
int drv_mlp_gradient(drv_ctx* ctx, const float* x_dev, const float* y_dev, uint32_t n, float* grad_host);
/* `steps` Adam steps (beta 0.9 / 0.999) on random minibatches of `batch`
 * samples; mean_loss_host (optional) receives the last step's mean loss. */
int drv_mlp_train(drv_ctx* ctx, const float* x_dev, const float* y_dev, uint32_t n, uint32_t batch, uint32_t steps,
                  float lr, uint32_t seed, float* mean_loss_host);

/* Online fine-tuning: collects n_samples fresh cone-traced samples from the
 * last drv_light_caches call (its spot-check caches when the surrogate was
 * on) into library-owned buffers and runs drv_mlp_train on them.  A render
 * loop calls it every few frames (the bench and tests use 16). */
int drv_mlp_finetune(drv_ctx* ctx, uint32_t n_samples, uint32_t batch, uint32_t steps, float lr, uint32_t seed,
                     float* mean_loss_host);

/* ---- launch tuning ---------------------------------------------------- */

/* Launch shapes of the gather stage; the defaults suit one 1080p view, and
 * the batched path's autotuner picks them per workload.  No counterpart in
 * the reference (its dispatch is fixed at 64 caches per workgroup,
 * shader/cacheLightingRSM.comp:7). */
typedef struct drv_tuning {
  int32_t visibility_kernel; /* 0: lane per cache, walking its item's blocks (default); 1: refilling lane pool */
  int32_t blocks_per_item;   /* shadow blocks per visibility work item, 1..64 (default 1); the kernel
                                uses the largest count up to this that leaves two items per wave */
  int32_t gather_max_splits; /* largest VPL split count of the gather, power of two 1..256 (default 256) */
  int32_t waves_per_cu;      /* launch grid of the light pass kernels, 4..1024 waves per CU (default 256:
                                more waves than fit at once, so the hardware deals out the last items) */
  int32_t refill_min_lanes;  /* lane pool: idle lanes of a wave that trigger a refill, 1..64 (default 16) */
  int32_t steps_per_batch;   /* cone-trace steps planned and loaded together: 1-4 or 8 (default 1) */
} drv_tuning;

int drv_set_tuning(drv_ctx* ctx, const drv_tuning* tuning);
int drv_get_tuning(const drv_ctx* ctx, drv_tuning* out);

/* ---- batched views ----------------------------------------------------- */

/* A batch lights the caches of several views (one drv_ctx each, e.g. views
 * of different scenes, resolutions and VPL counts) with one persistent
 * work-queue kernel instead of a pass per view: drv_batch_light_caches is
 * drv_light_caches for every view at once (same results up to the order of
 * the VPL sums).  All contexts must share sh_bands, indirect_shadow and their
 * HIP stream.  A batch with a view that has indirect_specular lights every
 * view with its own passes (caches per wave 0 below; the fused kernel writes
 * SH only).  No counterpart in the reference (one view per frame,
 * rendering/renderer.cpp:539-594). */
typedef struct drv_batch drv_batch;

/* Caches-per-wave value of drv_batch_set_tuning that selects the split queue:
 * the single-view pass's stages (VPL records, cone-traced visibility, SH
 * gather, split reduction) each run as one persistent launch over the work
 * items of every view, and the gather's VPL splits are planned over all views
 * together. */
#define DRV_BATCH_SPLIT_QUEUE (-1)
/* The split queue with the cone-traced visibility of each view in a launch of
 * its own (the single-view kernel) and the other stages queued over all
 * views. */
#define DRV_BATCH_SPLIT_GATHER (-2)

/* A batch keeps pointers to its views (all on one stream): destroy it before
 * any of them. */
int drv_batch_create(drv_ctx* const* views, int32_t n_views, drv_batch** out);
int drv_batch_destroy(drv_batch* batch);
const char* drv_batch_last_error(const drv_batch* batch);
/* Work decomposition: caches per wavefront (64, 32 or 16; lane groups of a
 * wave split the shadow blocks of a cache when < 64), VPL tile length (VPLs
 * per work item, a power of two >= the shadow block size; a view with more
 * VPLs is split into partial sums), persistent grid size in waves per CU.
 * Caches per wavefront 0 selects per-view passes instead of the fused queue:
 * each view's own cone trace, SH gather and reduction (drv_light_caches)
 * back to back on the shared stream; DRV_BATCH_SPLIT_QUEUE selects the split
 * queue and DRV_BATCH_SPLIT_GATHER the split queue with per-view visibility
 * launches (the VPL tile is then unused). */
int drv_batch_set_tuning(drv_batch* batch, int32_t caches_per_wave, int32_t vpl_tile, int32_t waves_per_cu);
int drv_batch_get_tuning(const drv_batch* batch, int32_t* caches_per_wave, int32_t* vpl_tile, int32_t* waves_per_cu);
/* rsms[i], lights[i] belong to view i. */
int drv_batch_light_caches(drv_batch* batch, const drv_rsm* rsms, const drv_spot_light* lights);
/* Times drv_batch_light_caches over the candidate decompositions (median of
 * `reps` runs each), keeps the fastest and returns it; the views' cache
 * entries are restored afterwards. */
int drv_batch_autotune(drv_batch* batch, const drv_rsm* rsms, const drv_spot_light* lights, int32_t reps,
                       int32_t* caches_per_wave, int32_t* vpl_tile, float* ms);

/* Surrogate visibility for every view of a batch: host_params (optional,
 * DRV_MLP_PARAMS floats) become every view's network, settings every view's
 * drv_set_surrogate.  With the network enabled the batch lights its views
 * with their own passes (caches per wave 0): each view runs the network, its
 * own spot check and its own device-gated fallback (drv_get_surrogate_stats
 * of the view reports them). */
int drv_batch_set_surrogate(drv_batch* batch, const float* host_params, const drv_surrogate* settings);
/* Online fine-tuning of the batch's network: n_samples fresh cone-traced
 * samples shared out over the views that have labels from the last light
 * pass (their spot sets when the network lit them), `steps` Adam steps on
 * view 0's network, whose parameters then go to every view. */
int drv_batch_mlp_finetune(drv_batch* batch, uint32_t n_samples, uint32_t batch_size, uint32_t steps, float lr,
                           uint32_t seed, float* mean_loss_host);

/* ---- HIP graph of a frame ----------------------------------------------- */

/* Captures everything the context's stream receives between begin and end
 * (any entry points of this or other contexts on the same stream) into a
 * HIP graph, replayed by drv_graph_launch.  The stream must not be the null
 * stream; calls that read back to the host (drv_get_cache_count, drv_read_*)
 * and first-time buffer growth are not allowed while capturing, so run the
 * frame once before capturing it.  Kernel arguments (matrices, volumes) are
 * those of the captured calls. */
int drv_graph_begin(drv_ctx* ctx);
int drv_graph_end(drv_ctx* ctx);
int drv_graph_launch(drv_ctx* ctx);

/* ---- profiling -------------------------------------------------------- */

/* When on, each stage is bracketed by HIP events on the context stream;
 * level 2 also counts cone-march steps (one atomic per wave). */
int drv_set_profiling(drv_ctx* ctx, int32_t enabled);
int drv_get_stage_times(drv_ctx* ctx, drv_stage_times* out);

#ifdef __cplusplus
}
#endif

#endif /* DRV_H */
