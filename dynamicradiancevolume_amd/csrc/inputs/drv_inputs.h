/*
 * drv_inputs.h - seeded procedural scenes and a CPU G-buffer / RSM producer.
 *
 * The reference renders its pipeline inputs with GL raster passes
 * (shader/fillgbuffer.frag, shader/fillrsm.frag, shader/downsamplersm.frag,
 * rendering/renderer.cpp:760-811, 1279-1339) over scenes loaded from model
 * files.  Those passes sit outside the indirect-lighting path; this library
 * produces inputs in exactly their formats (D32F depth, RG16I packed normals,
 * SRGB8 albedo; RGB16F flux, RG16I normals, RG16F depth/depth^2 plus the mip
 * pyramid) from seeded procedural scenes, so the HIP path and the oracle can
 * be fed identical bytes.
 */
#ifndef DRV_INPUTS_H
#define DRV_INPUTS_H

#include <stdint.h>

#include "../../../include/drv.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DRVIN_SCENE_CORNELL 0 /* 5 walls (open front) + 2 boxes */
#define DRVIN_SCENE_BOXES 1   /* ground plane + n random boxes and ramps */

/* Builds a scene as a non-indexed triangle soup.  positions: float[9 * cap],
 * albedo: linear RGB per triangle float[3 * cap], normals: unit face normal
 * per triangle float[3 * cap], bbox: float[6] = min xyz, max xyz.  Any output
 * pointer may be NULL.  Returns the triangle count (also when it exceeds
 * `cap`, in which case nothing beyond `cap` is written), or < 0 on error. */
int64_t drvin_scene_build(int32_t kind, uint64_t seed, int32_t n_objects, int32_t tess,
                          float* positions, float* albedo, float* normals, int64_t cap, float* bbox);

/* Camera matrices: right-handed look-at + reverse-Z perspective with
 * horizontal field of view `hfov` (radians).  Outputs row-major float[16]. */
int drvin_camera(const float eye[3], const float target[3], float hfov, float aspect, float near_plane,
                 float far_plane, float view_proj[16], float inverse_view_proj[16]);

/* Rasterizes the G-buffer: depth float[W*H] (0 = empty), normal int16[2*W*H],
 * albedo uint8[3*W*H] sRGB-encoded.  Row 0 is NDC y = -1. */
int drvin_raster_gbuffer(const float* positions, const float* albedo, const float* normals, int64_t n_tri,
                         const float view_proj[16], int32_t width, int32_t height, float* depth,
                         int16_t* normal, uint8_t* albedo_srgb, int32_t n_threads);

/* Rasterizes the RSM of one spot light at its render resolution (as
 * shader/fillrsm.frag), res x h texels (h = rsm_render_height: res for a
 * square RSM, res / 2 for a 2:1 one): flux RGB16F uint16[3*res*h], normal
 * int16[2*res*h], depth_lin_sq RG16F uint16[2*res*h]. */
int drvin_raster_rsm(const float* positions, const float* albedo, const float* normals, int64_t n_tri,
                     const drv_spot_light* light, uint16_t* flux, int16_t* normal,
                     uint16_t* depth_lin_sq, int32_t n_threads);

/* One RSM mip step res x res_h -> res/2 x res_h/2 (as
 * shader/downsamplersm.frag): flux is the 2x2 sum, the normal the
 * renormalized mean of the four unpacked normals, (d, d^2) the 2x2 mean. */
int drvin_rsm_downsample(int32_t res, int32_t res_h, const uint16_t* flux, const int16_t* normal,
                         const uint16_t* depth_lin_sq, uint16_t* flux_out, int16_t* normal_out,
                         uint16_t* depth_lin_sq_out);

/* Helpers exposed for tests. */
uint16_t drvin_f32_to_f16(float f);
float drvin_f16_to_f32(uint16_t h);
void drvin_pack_normal16i(const float n[3], int16_t out[2]);

#ifdef __cplusplus
}
#endif

#endif
