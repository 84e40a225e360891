// C ABI of the pipeline: the host side of Renderer (rendering/renderer.cpp)
// and Voxelization (rendering/voxelization.cpp) re-expressed over HIP device
// buffers and one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/drv.h"
#include "drv_internal.hpp"
#include "mat4.hpp"
#include "split_plan.hpp"

using drvk::Counter;

struct drv_ctx {
  drv_config cfg{};
  hipStream_t stream = nullptr;
  // Stage B (voxelization) runs on its own stream, forked from and joined
  // back into `stream`, so it overlaps stage A (allocation).
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, vox_done_ev = nullptr;
  // With the voxel overlap on (drv_set_voxel_overlap), recorded on `stream`
  // after the last enqueued work that reads the voxel volume (valid only
  // outside graph capture): the next voxelization waits for it instead of
  // for everything on `stream`, so it overlaps the rest of the frame (SH
  // gather, reduction, apply).
  hipEvent_t vox_free_ev = nullptr;
  bool vox_free_valid = false;
  bool vox_overlap = false;
  // The event the next voxelization waits for when vox_free_valid: this
  // context's vox_free_ev, or the one a batch records once for all its views.
  hipEvent_t vox_free_wait = nullptr;
  // Set by a batch around its views' own light passes: their voxel-free
  // point is recorded once by the batch instead.
  bool defer_vox_read = false;
  // The surrogate's spot check also runs on `aux`, beside the network pass.
  hipEvent_t spot_fork_ev = nullptr, spot_done_ev = nullptr;
  bool vox_pending = false;
  int device = 0;
  int cu_count = 256;
  std::string err;

  drv_sh_constants shc{};
  drv_frame frame{};
  drv_volume vol{};
  bool have_frame = false;
  bool have_volume = false;
  bool allocated = false;  // drv_allocate_caches ran since creation

  float* caches = nullptr;
  Counter* counter = nullptr;
  float4* phantoms = nullptr;        // allocation: positions of out-of-volume appends [max_caches]
  uint32_t* alloc_counts = nullptr;  // allocation: claimed cells per chunk, then of the last call [kAllocChunks + 1]
  uint32_t* alloc_claimed = nullptr; // allocation: CAV index of each claimed cell [cav_cells]
  uint32_t* cav = nullptr;
  size_t cav_cells = 0;
  uint8_t* vox_target = nullptr;
  uint8_t* vox_mips = nullptr;
  drvk::VoxelCorner* vox_corners = nullptr;
  drvk::VoxelLayout vox;
  float* srgb_lut = nullptr;

  drvk::Vpl* vpls = nullptr;
  size_t vpl_cap = 0;
  drvk::ShadowBlock* blocks = nullptr;
  drvk::MarchStep* march = nullptr;  // kMarchSteps per block
  size_t block_cap = 0;
  float* visibility = nullptr;
  size_t vis_blocks_cap = 0;  // blocks the visibility buffer holds (pitch = max_caches)
  float* partial = nullptr;
  uint32_t last_n_blocks = 0;
  bool counted_steps = false;

  // Surrogate visibility network.
  float* mlp_params = nullptr;  // FP32 master [kMlpParams]
  float* mlp_m = nullptr;       // Adam moments
  float* mlp_v = nullptr;
  float* mlp_grad = nullptr;  // reduced gradient [kMlpParams + 1] (last: loss)
  void* mlp_pack = nullptr;   // inference operands
  void* mlp_train_w = nullptr;  // training LDS images
  float* mlp_partial = nullptr;
  size_t mlp_partial_waves = 0;
  uint32_t* mlp_idx = nullptr;
  size_t mlp_idx_cap = 0;
  uint32_t* mlp_pairs = nullptr;
  size_t mlp_pairs_cap = 0;
  uint32_t* mlp_n = nullptr;
  float* ft_x = nullptr;  // fine-tuning samples
  float* ft_y = nullptr;
  size_t ft_x_cap = 0, ft_y_cap = 0;
  uint32_t adam_t = 0;
  bool mlp_ready = false;
  drv_surrogate sur{0, 0.01f, 0.05f};
  // Spot check: two compact cache sets, their visibility, counters
  // [traced set, network set, fallback gate] and the device-side stats.
  float* spot_caches = nullptr;
  size_t spot_caches_size = 0;  // floats
  size_t spot_cap = 0;          // caches per set
  float* spot_vis = nullptr;
  size_t spot_vis_size = 0;
  Counter* spot_ctr = nullptr;
  uint32_t* spot_stats = nullptr;
  // Where drv_mlp_collect finds cone-traced labels of the last light pass.
  const float* label_caches = nullptr;
  const Counter* label_counter = nullptr;
  const float* label_vis = nullptr;
  uint32_t label_pitch = 0;

  drvk::Tuning tune;
  // The split partial buffer: 27 rows of `partial_cols` floats (a gather
  // with s splits at pitch p needs ceil(s / 4) x p columns).
  size_t partial_cols = 0;

  // Indirect specular: atlas, work sums, spill; spec_first is set by the
  // allocation pass so that the frame's first light overwrites the sums.
  drvk::SpecAtlas spec;
  bool spec_first = true;

  // Device-rendered inputs: raster scratch, G-buffer, RSM mip chain.
  unsigned long long* rz = nullptr;
  uint32_t* rtri = nullptr;
  uint32_t* rbig = nullptr;  // large-triangle list: count + indices
  size_t raster_cap = 0, raster_tri_cap = 0, rbig_cap = 0;
  float* g_depth = nullptr;
  int16_t* g_normal = nullptr;
  uint8_t* g_albedo = nullptr;
  uint16_t* rsm_flux = nullptr;
  int16_t* rsm_normal = nullptr;
  uint16_t* rsm_dls = nullptr;
  size_t rsm_cap = 0, rsm_normal_cap = 0, rsm_dls_cap = 0;  // texels of the chain
  float* rsm_shadow = nullptr;  // the light's depth buffer at render resolution
  size_t rsm_shadow_cap = 0;
  int rsm_res = 0;                                           // render resolution of the last RSM

  bool profiling = false;
  int profiling_level = 0;
  uint64_t* step_counter = nullptr;  // [traces, reference steps, executed steps] when counting
  hipEvent_t ev[16] = {};
  uint32_t recorded = 0;  // events recorded at least once
  hipGraphExec_t graph_exec = nullptr;
  drv_stage_times times{};
};

namespace {


int fail(drv_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int check_hip(drv_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return DRV_OK;
  return fail(ctx, DRV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define DRV_HIP(ctx, expr)                                   \
  do {                                                       \
    int _rc = check_hip((ctx), (expr), #expr);               \
    if (_rc != DRV_OK) return _rc;                           \
  } while (0)

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int log2i(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) ++l;
  return l;
}

template <typename T>
int dev_alloc(drv_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return DRV_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) return fail(ctx, DRV_ERR_NO_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  return DRV_OK;
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Host matrices of PrepareLights (renderer.cpp:687-692): camera at the light,
// reverse-Z perspective with fov = 2 * halfAngle and aspect 1.
void light_matrices(const drv_spot_params& p, float* vp_out, float* ivp_out) {
  double eye[3] = {p.position[0], p.position[1], p.position[2]};
  double tgt[3] = {eye[0] + p.direction[0], eye[1] + p.direction[1], eye[2] + p.direction[2]};
  double up[3] = {0.0, 1.0, 0.0};
  double view[16], proj[16], vp[16], ivp[16];
  drvm::look_at(eye, tgt, up, view);
  drvm::perspective_reverse_z(2.0 * p.half_angle, 1.0, p.near_plane, p.far_plane, proj);
  drvm::mat4_mul(proj, view, vp);
  drvm::mat4_invert(vp, ivp);
  drvm::to_float(vp, vp_out);
  drvm::to_float(ivp, ivp_out);
}

void record(drv_ctx* ctx, int i, hipStream_t s) {
  if (ctx->profiling && hipEventRecord(ctx->ev[i], s) == hipSuccess) ctx->recorded |= 1u << i;
}

void record(drv_ctx* ctx, int i) { record(ctx, i, ctx->stream); }

// Makes `stream` wait for a voxelization enqueued on the side stream.
void join_voxels(drv_ctx* ctx) {
  if (!ctx->vox_pending) return;
  (void)hipStreamWaitEvent(ctx->stream, ctx->vox_done_ev, 0);
  ctx->vox_pending = false;
}

// The same before work that reads the voxels: until that work's voxels_read
// records a new voxel-free point, the last one precedes reads already
// enqueued, so a call that fails in between leaves none (the next
// voxelization then waits for everything on the stream).
void join_voxels_for_read(drv_ctx* ctx) {
  join_voxels(ctx);
  ctx->vox_free_valid = false;
}

bool is_capturing(hipStream_t s);

// Every call that enqueues reads of the voxel volume on `stream` calls this
// after them: with the voxel overlap on, the next voxelization may start once
// they have run.  Under graph capture the event would belong to the graph, so
// the next voxelization then waits for the whole stream instead.
void voxels_read(drv_ctx* ctx) {
  if (ctx->defer_vox_read) return;
  ctx->vox_free_valid = ctx->vox_overlap && !is_capturing(ctx->stream) &&
                        hipEventRecord(ctx->vox_free_ev, ctx->stream) == hipSuccess;
  ctx->vox_free_wait = ctx->vox_free_ev;
}

// Time between two recorded events; 0 for a pair that was never recorded.
// A failed query must not leave an error behind for the next entry point's
// hipGetLastError check.
float elapsed(drv_ctx* ctx, int a, int b) {
  const uint32_t need = (1u << a) | (1u << b);
  if ((ctx->recorded & need) != need) return 0.0f;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.0f;
  }
  return ms;
}

drvk::VoxelView voxel_view(const drv_ctx* ctx) {
  drvk::VoxelView v{};
  v.corners = ctx->vox_corners;
  v.corner_bytes = static_cast<uint32_t>(ctx->vox.corner_count * sizeof(drvk::VoxelCorner));
  v.res = ctx->vox.res;
  v.levels = ctx->vox.levels;
  for (int l = 0; l < ctx->vox.levels && l < 16; ++l) v.corner_offset[l] = static_cast<uint32_t>(ctx->vox.corner_offset[l]);
  return v;
}

// The light pass grid: waves_per_cu waves per CU in 256-thread workgroups.
int persistent_grid(const drv_ctx* ctx) { return std::max(1, ctx->cu_count * ctx->tune.waves_per_cu / 4); }

// The network's kernels keep their weights in registers, loaded once per
// workgroup: a persistent grid of 64 waves per CU, whatever the light pass
// grid is.
int network_grid(const drv_ctx* ctx) { return std::max(1, ctx->cu_count * 64 / 4); }

template <typename T>
int ensure(drv_ctx* ctx, T** p, size_t* cap, size_t need) {
  if (need <= *cap && *p) return DRV_OK;
  dev_free(*p);
  *cap = 0;
  int rc = dev_alloc(ctx, p, need);
  if (rc == DRV_OK) *cap = need;
  return rc;
}

// Every k-th cache is spot-checked, k = round(1 / spot fraction) clamped to
// [1, capacity] (a fraction below 1 / capacity checks one cache).
uint32_t spot_divisor(const drv_ctx* ctx) {
  const double cap = static_cast<double>(ctx->cfg.max_caches);
  const double k = std::round(1.0 / static_cast<double>(ctx->sur.spot_fraction));
  return static_cast<uint32_t>(std::min(cap, std::max(1.0, k)));
}

// Caches per spot set (set a and set b each hold this many entries).
size_t spot_set_capacity(const drv_ctx* ctx) {
  const size_t cap = static_cast<size_t>(ctx->cfg.max_caches);
  const size_t k = spot_divisor(ctx);
  return (cap + k - 1) / k;
}

// Visibility of every (live cache, shadow block) from the network, checked on
// a spot set of caches against the cone trace; the full cone trace runs
// (device-gated) when the spot check fails.  The spot check (a few latency-
// bound launches over ~1 % of the caches) runs on the side stream while the
// network pass over every cache fills the GPU; the gated trace waits for it.
int surrogate_visibility(drv_ctx* ctx, uint32_t n_blocks, uint32_t n_vpl, uint32_t interval, int grid) {
  const size_t cap = static_cast<size_t>(ctx->cfg.max_caches);
  const int stride = drvk::entry_stride(ctx->cfg.sh_bands);
  const drvk::VoxelView vox = voxel_view(ctx);
  const drvk::FeatureArgsHost fh = drvk::make_feature_args(ctx->vol, ctx->frame.camera_position, vox);
  const bool spot = ctx->sur.spot_fraction > 0.0f;
  int rc;
  if (spot) {
    const uint32_t k = spot_divisor(ctx);
    const size_t spot_cap = spot_set_capacity(ctx);
    if ((rc = ensure(ctx, &ctx->spot_caches, &ctx->spot_caches_size, 2 * spot_cap * stride)) != DRV_OK) return rc;
    if ((rc = ensure(ctx, &ctx->spot_vis, &ctx->spot_vis_size, 2 * spot_cap * n_blocks)) != DRV_OK) return rc;
    if (!ctx->spot_ctr && (rc = dev_alloc(ctx, &ctx->spot_ctr, 4)) != DRV_OK) return rc;
    if (!ctx->spot_stats && (rc = dev_alloc(ctx, &ctx->spot_stats, 4)) != DRV_OK) return rc;
    ctx->spot_cap = spot_cap;
    const uint32_t sc = static_cast<uint32_t>(spot_cap);
    // One list of both sets: set a (traced) at entries [0, n), set b (network)
    // at [n, 2n), n = spot_ctr[0].live on the device; one visibility array of
    // pitch 2 sc, so one SH gather and one reduction light both sets.
    const uint32_t pitch2 = 2u * sc;
    float* spot_list = ctx->spot_caches;
    hipStream_t ss = ctx->aux;
    DRV_HIP(ctx, hipEventRecord(ctx->spot_fork_ev, ctx->stream));
    DRV_HIP(ctx, hipStreamWaitEvent(ss, ctx->spot_fork_ev, 0));
    drvk::launch_spot_select(ctx->counter, k, ctx->caches, stride, spot_list, sc, ctx->spot_ctr, ss);
    drvk::launch_visibility(ctx->vol, vox, spot_list, stride, &ctx->spot_ctr[0], ctx->blocks, n_blocks, ctx->spot_vis,
                            pitch2, nullptr, ctx->tune, grid, ss, ctx->march);
    drvk::launch_visibility_mlp(fh, ctx->mlp_pack, ctx->mlp_params, spot_list, stride, &ctx->spot_ctr[1],
                                &ctx->spot_ctr[0], ctx->blocks, n_blocks, ctx->spot_vis, pitch2, network_grid(ctx), ss);
    // A small list fits more splits into the partial buffer than the main
    // list: the most (a power of two) whose ceil(s / 4) x pitch2 columns fit.
    // drv_light_caches sized the buffer to hold at least one group.
    const uint32_t spot_splits = drvk::max_fitting_splits(ctx->partial_cols, pitch2, drvk::kMaxSplits);
    if (spot_splits == 0u) return fail(ctx, DRV_ERR_STATE, "light_caches: spot-check partials do not fit");
    drvk::launch_gather(ctx->shc, ctx->cfg.sh_bands, spot_list, stride, &ctx->spot_ctr[3], ctx->vpls, n_vpl,
                        ctx->spot_vis, pitch2, interval, 1, ctx->partial, pitch2, spot_splits, grid, true, ss);
    drvk::launch_reduce_sh(ctx->shc, ctx->cfg.sh_bands, spot_list, stride, &ctx->spot_ctr[3], ctx->partial, pitch2,
                           ctx->cu_count * 8, ss);
    drvk::launch_spot_error(spot_list, sc, stride, ctx->cfg.sh_bands, ctx->spot_ctr, ctx->sur.fallback_threshold,
                            ctx->counter, ctx->spot_stats, &ctx->spot_ctr[2], ss);
    DRV_HIP(ctx, hipEventRecord(ctx->spot_done_ev, ss));
    ctx->label_caches = spot_list;
    ctx->label_counter = &ctx->spot_ctr[0];
    ctx->label_vis = ctx->spot_vis;
    ctx->label_pitch = pitch2;
  } else {
    ctx->label_caches = nullptr;
  }
  drvk::launch_visibility_mlp(fh, ctx->mlp_pack, ctx->mlp_params, ctx->caches, stride, ctx->counter, nullptr,
                              ctx->blocks, n_blocks, ctx->visibility, static_cast<uint32_t>(cap), network_grid(ctx),
                              ctx->stream);
  if (spot) {  // gate counter: live = 0 unless the spot check failed
    DRV_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->spot_done_ev, 0));
    drvk::launch_visibility(ctx->vol, vox, ctx->caches, stride, &ctx->spot_ctr[2], ctx->blocks, n_blocks,
                            ctx->visibility, static_cast<uint32_t>(cap), nullptr, ctx->tune, grid, ctx->stream,
                            ctx->march);
  }
  return DRV_OK;
}

}  // namespace

extern "C" {

int drv_create(const drv_config* config, drv_ctx** out_ctx) {
  if (!config || !out_ctx) return DRV_ERR_INVALID_ARG;
  *out_ctx = nullptr;
  const drv_config& c = *config;
  if (c.width <= 0 || c.height <= 0 || c.max_caches <= 0 || c.cav_resolution <= 0 || c.cav_cascades <= 0 ||
      c.cav_cascades > DRV_MAX_CASCADES || !is_pow2(c.voxel_resolution) || c.voxel_resolution < 2 ||
      c.voxel_resolution > 512 || (c.sh_bands != 1 && c.sh_bands != 2) || c.cav_cascade0_world_size <= 0.0f)
    return DRV_ERR_INVALID_ARG;
  // Address arithmetic of the CAV is 32-bit (cacheGather.comp:25-28).
  if (static_cast<double>(c.cav_cascades) * c.cav_resolution * c.cav_resolution * c.cav_resolution > 2147483647.0)
    return DRV_ERR_INVALID_ARG;
  if (c.indirect_specular) {
    const int size = c.specular_envmap_size == 0 ? 16 : c.specular_envmap_size;
    if (!is_pow2(size) || size > 32 || c.specular_hole_fill_level < 0) return DRV_ERR_INVALID_ARG;
  }
  drv_ctx* ctx = new (std::nothrow) drv_ctx();
  if (!ctx) return DRV_ERR_NO_MEMORY;
  ctx->cfg = c;
  if (c.indirect_specular) {
    drv_config& k = ctx->cfg;
    k.indirect_specular = 1;
    if (k.specular_envmap_size == 0) k.specular_envmap_size = 16;
    k.specular_hole_fill_level = std::min(k.specular_hole_fill_level, log2i(k.specular_envmap_size));
  } else {
    ctx->cfg.indirect_specular = 0;
  }
  int rc = DRV_OK;
  auto bail = [&](int code) {
    drv_destroy(ctx);
    return code;
  };
  if (hipGetDevice(&ctx->device) != hipSuccess) return bail(DRV_ERR_HIP);
  // Device address of the allocation-free light pass's scratch array on this
  // device (resolved once per device, with it current).
  if (drvk::static_scratch_init(ctx->device) != hipSuccess) return bail(DRV_ERR_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->cu_count = prop.multiProcessorCount;

  // UpdateConstantUBO SH factors (renderer.cpp:296-306).
  const float pi = 3.14159265358979f;
  drv_sh_constants& s = ctx->shc;
  s.cos_lobe0 = std::sqrt(pi) / 2.0f;
  s.cos_lobe1 = std::sqrt(pi / 3.0f);
  s.cos_lobe2n2_p1_n1 = -std::sqrt(15.0f * pi) / 8.0f;
  s.cos_lobe20 = std::sqrt(5.0f * pi) / 16.0f;
  s.cos_lobe2p2 = std::sqrt(15.0f * pi) / 16.0f;
  s.eva0 = 1.0f / (2.0f * std::sqrt(pi));
  s.eva1 = std::sqrt(3.0f) / (2.0f * std::sqrt(pi));
  s.eva2n2_p1_n1 = std::sqrt(15.0f / (4.0f * pi));
  s.eva20 = std::sqrt(5.0f / (16.0f * pi));
  s.eva2p2 = std::sqrt(15.0f / (16.0f * pi));

  // AllocateCacheData (renderer.cpp:262-276): the buffer holds max_caches
  // entries of the active layout.
  const size_t stride = static_cast<size_t>(drvk::entry_stride(c.sh_bands));
  if ((rc = dev_alloc(ctx, &ctx->caches, stride * c.max_caches)) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->counter, 1)) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->phantoms, static_cast<size_t>(c.max_caches))) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->alloc_counts, drvk::kAllocChunks + 1)) != DRV_OK) return bail(rc);
  // SetCAVCascades: R32UI atlas of (C * R) x R x R (renderer.cpp:1179).
  ctx->cav_cells = static_cast<size_t>(c.cav_cascades) * c.cav_resolution * c.cav_resolution * c.cav_resolution;
  if ((rc = dev_alloc(ctx, &ctx->cav, ctx->cav_cells)) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->alloc_claimed, ctx->cav_cells)) != DRV_OK) return bail(rc);

  // Voxelization::SetResolution (voxelization.cpp:56-60): R8 target plus an
  // R8 volume with its full mip chain, plus the corner-expanded copy.
  drvk::VoxelLayout& L = ctx->vox;
  L.res = c.voxel_resolution;
  L.levels = log2i(L.res) + 1;
  L.mip_offset.resize(L.levels);
  L.corner_offset.resize(L.levels);
  for (int l = 0; l < L.levels; ++l) {
    const size_t w = static_cast<size_t>(L.res >> l);
    L.mip_offset[l] = L.mip_bytes;
    L.mip_bytes += w * w * w;
    L.corner_offset[l] = L.corner_count;
    L.corner_count += (w + 1) * (w + 1) * (w + 1);
  }
  const size_t res3 = static_cast<size_t>(L.res) * L.res * L.res;
  if ((rc = dev_alloc(ctx, &ctx->vox_target, res3)) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->vox_mips, L.mip_bytes)) != DRV_OK) return bail(rc);
  if ((rc = dev_alloc(ctx, &ctx->vox_corners, L.corner_count)) != DRV_OK) return bail(rc);

  // SRGB8 decode table.
  float lut[256];
  for (int i = 0; i < 256; ++i) {
    const double v = i / 255.0;
    lut[i] = static_cast<float>(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
  }
  if ((rc = dev_alloc(ctx, &ctx->srgb_lut, 256)) != DRV_OK) return bail(rc);
  if (hipMemcpy(ctx->srgb_lut, lut, sizeof(lut), hipMemcpyHostToDevice) != hipSuccess) return bail(DRV_ERR_HIP);

  // Specular atlas (renderer.cpp:253-286): side 2^ceil(log2(ceil(sqrt(max
  // caches)) * S)), log2(S) + 1 levels; work sums and spill per cache.
  if (ctx->cfg.indirect_specular) {
    drvk::SpecAtlas& A = ctx->spec;
    A.S = ctx->cfg.specular_envmap_size;
    A.levels = log2i(A.S) + 1;
    A.fill = ctx->cfg.specular_hole_fill_level;
    const double side = std::ceil(std::sqrt(static_cast<double>(c.max_caches))) * A.S;
    const double w = std::pow(2.0, std::ceil(std::log2(side)));
    if (w > 32768.0) return bail(DRV_ERR_INVALID_ARG);  // texel indices of a level stay 32-bit
    A.width = static_cast<uint32_t>(w);
    A.per_row = A.width / static_cast<uint32_t>(A.S);
    uint64_t off = 0;
    for (int l = 0; l < A.levels; ++l) {
      A.level_offset[l] = off;
      const uint64_t wl = A.width >> l;
      off += wl * wl;
    }
    A.texels = off;
    const size_t n = static_cast<size_t>(c.max_caches);
    if ((rc = dev_alloc(ctx, &A.atlas, A.texels)) != DRV_OK) return bail(rc);
    if ((rc = dev_alloc(ctx, &A.work, n * A.S * A.S * 3)) != DRV_OK) return bail(rc);
    if ((rc = dev_alloc(ctx, &A.spill, n * (2 * A.S + 1) * 3)) != DRV_OK) return bail(rc);
    drvk::launch_fill_u32(A.atlas, A.texels, 0u, nullptr);
  }

  // Zero-initialize the state a frame may read before writing it.
  drvk::launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->counter), sizeof(Counter) / 4, 0u, nullptr);
  drvk::launch_fill_u32(ctx->cav, ctx->cav_cells, 0u, nullptr);
  drvk::launch_fill_u32(ctx->alloc_counts, drvk::kAllocChunks + 1, 0u, nullptr);
  drvk::launch_fill_u8(ctx->vox_target, res3, 0, nullptr);
  drvk::launch_fill_u8(ctx->vox_mips, L.mip_bytes, 0, nullptr);
  drvk::launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->vox_corners), L.corner_count * 4, 0u, nullptr);
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(DRV_ERR_HIP);
  if (hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->vox_done_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->vox_free_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->spot_fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->spot_done_ev, hipEventDisableTiming) != hipSuccess)
    return bail(DRV_ERR_HIP);
  if (hipDeviceSynchronize() != hipSuccess) return bail(DRV_ERR_HIP);
  *out_ctx = ctx;
  return DRV_OK;
}

int drv_destroy(drv_ctx* ctx) {
  if (!ctx) return DRV_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  if (ctx->aux) {
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipStreamDestroy(ctx->aux);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->vox_done_ev) (void)hipEventDestroy(ctx->vox_done_ev);
  if (ctx->vox_free_ev) (void)hipEventDestroy(ctx->vox_free_ev);
  if (ctx->spot_fork_ev) (void)hipEventDestroy(ctx->spot_fork_ev);
  if (ctx->spot_done_ev) (void)hipEventDestroy(ctx->spot_done_ev);
  dev_free(ctx->caches);
  dev_free(ctx->counter);
  dev_free(ctx->phantoms);
  dev_free(ctx->alloc_counts);
  dev_free(ctx->alloc_claimed);
  dev_free(ctx->cav);
  dev_free(ctx->spec.atlas);
  dev_free(ctx->spec.vpls_lanes);
  dev_free(ctx->spec.work);
  dev_free(ctx->spec.spill);
  dev_free(ctx->vox_target);
  dev_free(ctx->vox_mips);
  dev_free(ctx->vox_corners);
  dev_free(ctx->srgb_lut);
  dev_free(ctx->vpls);
  dev_free(ctx->blocks);
  dev_free(ctx->march);
  dev_free(ctx->visibility);
  dev_free(ctx->partial);
  dev_free(ctx->step_counter);
  dev_free(ctx->mlp_params);
  dev_free(ctx->mlp_m);
  dev_free(ctx->mlp_v);
  dev_free(ctx->mlp_grad);
  dev_free(ctx->mlp_pack);
  dev_free(ctx->mlp_train_w);
  dev_free(ctx->mlp_partial);
  dev_free(ctx->mlp_idx);
  dev_free(ctx->mlp_pairs);
  dev_free(ctx->mlp_n);
  dev_free(ctx->ft_x);
  dev_free(ctx->ft_y);
  dev_free(ctx->spot_caches);
  dev_free(ctx->spot_vis);
  dev_free(ctx->spot_ctr);
  dev_free(ctx->spot_stats);
  dev_free(ctx->rz);
  dev_free(ctx->rtri);
  dev_free(ctx->rbig);
  dev_free(ctx->g_depth);
  dev_free(ctx->g_normal);
  dev_free(ctx->g_albedo);
  dev_free(ctx->rsm_flux);
  dev_free(ctx->rsm_normal);
  dev_free(ctx->rsm_dls);
  dev_free(ctx->rsm_shadow);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  delete ctx;
  return DRV_OK;
}

const char* drv_last_error(const drv_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int drv_set_stream(drv_ctx* ctx, void* hip_stream) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  return DRV_OK;
}

int drv_synchronize(drv_ctx* ctx) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  join_voxels(ctx);
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_get_sh_constants(const drv_ctx* ctx, drv_sh_constants* out) {
  if (!ctx || !out) return DRV_ERR_INVALID_ARG;
  *out = ctx->shc;
  return DRV_OK;
}

int drv_set_frame(drv_ctx* ctx, const drv_frame* frame) {
  if (!ctx || !frame) return DRV_ERR_INVALID_ARG;
  ctx->frame = *frame;
  ctx->have_frame = true;
  return DRV_OK;
}

int drv_update_volume(drv_ctx* ctx, const float cam[3], const float smin[3], const float smax[3]) {
  if (!ctx || !cam || !smin || !smax) return DRV_ERR_INVALID_ARG;
  const drv_config& cfg = ctx->cfg;
  drv_volume& out = ctx->vol;
  std::memset(&out, 0, sizeof(out));
  // renderer.cpp:350-360: cubic voxel volume around the scene bounds.
  float vmin[3], vmax[3], ext[3];
  for (int i = 0; i < 3; ++i) {
    vmin[i] = smin[i] - 0.001f;
    vmax[i] = smax[i] + 0.001f;
    ext[i] = vmax[i] - vmin[i];
  }
  const float largest = std::max(std::max(ext[0], ext[1]), ext[2]);
  for (int i = 0; i < 3; ++i) {
    vmax[i] = vmax[i] + (largest - ext[i]);
    out.volume_world_min[i] = vmin[i];
    out.volume_world_max[i] = vmax[i];
  }
  out.voxel_size_in_world = (vmax[0] - vmin[0]) / static_cast<float>(cfg.voxel_resolution);
  out.cav_transition_zone_size = cfg.cav_transition_size;
  // renderer.cpp:366-428: camera-centred cascades, snapped to their cell
  // size; the decision box is unsnapped and shrunk by 1.5 cells.
  float size = cfg.cav_cascade0_world_size;
  for (int c = 0; c < cfg.cav_cascades; ++c) {
    const float cell = size / static_cast<float>(cfg.cav_resolution);
    drv_cascade& k = out.cascades[c];
    for (int i = 0; i < 3; ++i) {
      const float snapped = std::round(cam[i] / cell) * cell;
      k.min[i] = snapped - size * 0.5f;
      k.max[i] = snapped + size * 0.5f;
      k.decision_min[i] = cam[i] - size * 0.5f + cell * 1.5f;
      k.decision_max[i] = cam[i] + size * 0.5f - cell * 1.5f;
    }
    k.world_voxel_size = cell;
    size = size * 2.0f;
  }
  ctx->have_volume = true;
  return DRV_OK;
}

int drv_set_volume(drv_ctx* ctx, const drv_volume* volume) {
  if (!ctx || !volume) return DRV_ERR_INVALID_ARG;
  ctx->vol = *volume;
  ctx->have_volume = true;
  return DRV_OK;
}

int drv_get_volume(const drv_ctx* ctx, drv_volume* out) {
  if (!ctx || !out) return DRV_ERR_INVALID_ARG;
  *out = ctx->vol;
  return DRV_OK;
}

int drv_prepare_spot_light(const drv_spot_params* p, drv_spot_light* o) {
  if (!p || !o || p->rsm_resolution == 0 || !(p->half_angle > 0.0f) || !(p->near_plane > 0.0f) ||
      !(p->far_plane > p->near_plane))
    return DRV_ERR_INVALID_ARG;
  const float dl = std::sqrt(p->direction[0] * p->direction[0] + p->direction[1] * p->direction[1] +
                             p->direction[2] * p->direction[2]);
  if (!(dl > 0.0f)) return DRV_ERR_INVALID_ARG;
  std::memset(o, 0, sizeof(*o));
  for (int i = 0; i < 3; ++i) {
    o->light_intensity[i] = p->intensity[i];
    o->light_position[i] = p->position[i];
    o->light_direction[i] = p->direction[i] / dl;
  }
  o->light_cos_half_angle = std::cos(p->half_angle);
  light_matrices(*p, o->light_view_projection, o->inverse_light_view_projection);
  // renderer.cpp:694-719
  const int res_pow2 = 1 << static_cast<int>(std::ceil(std::log2(static_cast<double>(p->rsm_resolution))));
  o->rsm_render_resolution = res_pow2;
  const int read_res = static_cast<int>(res_pow2 / std::pow(2.0, static_cast<double>(p->rsm_read_lod)));
  if (read_res < 1) return DRV_ERR_INVALID_ARG;
  o->rsm_read_resolution = read_res;
  // A 2:1 RSM keeps the square light frustum and halves the rows.
  const bool wide = p->rsm_wide != 0;
  o->rsm_render_height = wide ? res_pow2 / 2 : res_pow2;
  const int read_h = wide ? read_res / 2 : read_res;
  if (read_h < 1) return DRV_ERR_INVALID_ARG;
  o->rsm_read_height = read_h;
  // The solid angle of one read texel at unit distance (renderer.cpp:700-705
  // for a square read: the clip width squared over the texel count).
  const float clip_w = std::sin(p->half_angle) * p->near_plane * 2.0f;
  const float vaf = clip_w * clip_w / (p->near_plane * p->near_plane * read_res * read_h);
  o->val_area_factor = vaf;
  o->indirect_shadow_lod = static_cast<float>(p->indirect_shadow_lod);
  const float block = static_cast<float>(1 << p->indirect_shadow_lod);
  if (block > static_cast<float>(read_h)) return DRV_ERR_INVALID_ARG;  // renderer.cpp:712
  o->indirect_shadow_block_size = block;
  o->indirect_shadow_sample_interval = static_cast<int>(block * block);
  o->indirect_shadow_super_val_width = std::sqrt(vaf) * block;
  o->indirect_shadow_sampling_offset = 0.5f + std::sqrt(2.0f) * block / 2.0f;
  return DRV_OK;
}

// ---- stage B ----------------------------------------------------------

int drv_set_voxel_overlap(drv_ctx* ctx, int32_t on) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  ctx->vox_overlap = on != 0;
  // An event recorded while the overlap was on must not start a
  // voxelization after it is turned off.
  if (!ctx->vox_overlap) ctx->vox_free_valid = false;
  return DRV_OK;
}

int drv_voxelize_begin(drv_ctx* ctx) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  // Fork: the side stream starts after everything already on `stream`, or,
  // with the voxel overlap on, after the last enqueued work that reads the
  // voxels (the previous frame's cone trace) when that point is known (not
  // under capture, a reader since the last voxelization).
  if (ctx->vox_free_valid && !is_capturing(ctx->stream)) {
    DRV_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->vox_free_wait, 0));
  } else {
    DRV_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
    DRV_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->fork_ev, 0));
  }
  ctx->vox_free_valid = false;
  record(ctx, 0, ctx->aux);
  const size_t res3 = static_cast<size_t>(ctx->vox.res) * ctx->vox.res * ctx->vox.res;
  drvk::launch_fill_u8(ctx->vox_target, res3, 0, ctx->aux);  // voxelization.cpp:105
  ctx->vox_pending = true;
  return check_hip(ctx, hipGetLastError(), "voxelize_begin");
}

int drv_voxelize_mesh(drv_ctx* ctx, const float* positions, const uint32_t* indices, uint32_t n_tri,
                      const float world[16]) {
  if (!ctx || (!positions && n_tri) || !world) return DRV_ERR_INVALID_ARG;
  if (!ctx->have_volume) return fail(ctx, DRV_ERR_STATE, "voxelize: volume not set");
  if (!ctx->vox_pending) return fail(ctx, DRV_ERR_STATE, "voxelize_mesh: drv_voxelize_begin first");
  drvk::launch_voxelize(ctx->vol, ctx->vox.res, positions, indices, n_tri, world, ctx->vox_target, ctx->aux);
  return check_hip(ctx, hipGetLastError(), "voxelize_mesh");
}

int drv_voxelize_end(drv_ctx* ctx, float adaption_factor) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  if (!ctx->vox_pending) return fail(ctx, DRV_ERR_STATE, "voxelize_end: drv_voxelize_begin first");
  record(ctx, 1, ctx->aux);
  // voxelization.cpp:100: nothing is blended or mipmapped when the floored
  // adaption is zero.
  if (adaption_factor > 0.0f) {
    drvk::launch_voxel_blend_mips(ctx->vox, ctx->vox_target, ctx->vox_mips, adaption_factor, ctx->aux);
    drvk::launch_voxel_corners(ctx->vox, ctx->vox_mips, ctx->vox_corners, ctx->aux);
  }
  record(ctx, 2, ctx->aux);
  // Join point: the next pass that reads the voxels waits for this event.
  DRV_HIP(ctx, hipEventRecord(ctx->vox_done_ev, ctx->aux));
  return check_hip(ctx, hipGetLastError(), "voxelize_end");
}

// ---- stage A ----------------------------------------------------------

int drv_allocate_caches(drv_ctx* ctx, const drv_gbuffer* gb) {
  if (!ctx || !gb || !gb->depth) return DRV_ERR_INVALID_ARG;
  if (gb->width != ctx->cfg.width || gb->height != ctx->cfg.height)
    return fail(ctx, DRV_ERR_INVALID_ARG, "allocate: G-buffer size differs from the configured backbuffer");
  if (!ctx->have_frame || !ctx->have_volume) return fail(ctx, DRV_ERR_STATE, "allocate: frame or volume not set");
  record(ctx, 3);
  // renderer.cpp:969-970 (the clear of the counter and the address volume)
  // and cacheGather.comp, cachePrepareLighting.comp.
  drvk::launch_allocate(ctx->cfg, ctx->vol, ctx->frame, gb->depth, ctx->cav, ctx->caches, ctx->counter,
                        ctx->phantoms, ctx->alloc_counts, ctx->alloc_claimed, ctx->stream);
  record(ctx, 4);
  ctx->allocated = true;
  ctx->spec_first = true;
  // A new cache list: visibility and training labels of earlier passes no
  // longer refer to these caches.
  ctx->last_n_blocks = 0;
  ctx->label_caches = nullptr;
  return check_hip(ctx, hipGetLastError(), "allocate");
}

// ---- stage C ----------------------------------------------------------

}  // extern "C"

namespace {

// Validation of one light's RSM (cacheLightingRSM.comp:80-163 needs a square
// power-of-two read resolution; the 64-VPL chunking needs R >= 8).
int validate_light(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light, uint32_t* n_vpl_out,
                   uint32_t* interval_out, uint32_t* n_blocks_out) {
  if (!ctx || !rsm || !light || !rsm->flux || !rsm->normal || !rsm->depth_lin_sq) return DRV_ERR_INVALID_ARG;
  if (!ctx->allocated) return fail(ctx, DRV_ERR_STATE, "light_caches: no allocation yet");
  const int R = rsm->read_resolution;
  const int H = rsm->read_height > 0 ? rsm->read_height : R;
  const int LH = light->rsm_read_height > 0 ? light->rsm_read_height : light->rsm_read_resolution;
  if (!is_pow2(R) || R < 8 || R > 65536 || R != light->rsm_read_resolution)
    return fail(ctx, DRV_ERR_INVALID_ARG, "light_caches: read resolution must be a power of two >= 8 matching the light");
  // Square reads (cacheLightingRSM.comp:142) or the 2:1 extension, whose
  // Morton decode covers the 2h x h texels once.
  if ((H != R && 2 * H != R) || H != LH || R * static_cast<int64_t>(H) < 64)
    return fail(ctx, DRV_ERR_INVALID_ARG, "light_caches: read height must equal the width or half of it, matching the light");
  const uint32_t n_vpl = static_cast<uint32_t>(R) * static_cast<uint32_t>(H);
  const bool shadow = ctx->cfg.indirect_shadow != 0;
  const uint32_t interval = static_cast<uint32_t>(light->indirect_shadow_sample_interval);
  const int lod = static_cast<int>(light->indirect_shadow_lod);
  if (shadow) {
    if (!rsm->depth_lin_sq_shadow) return fail(ctx, DRV_ERR_INVALID_ARG, "light_caches: shadow LOD level missing");
    if (interval == 0 || interval != (1u << (2 * lod)) || (H >> lod) < 1)
      return fail(ctx, DRV_ERR_INVALID_ARG, "light_caches: inconsistent indirect shadow LOD");
  }
  const uint32_t n_blocks = shadow ? n_vpl / interval : 0u;
  *n_vpl_out = n_vpl;
  *interval_out = shadow ? interval : 64u;
  *n_blocks_out = n_blocks;
  return DRV_OK;
}

// The visibility buffer (n_blocks rows of capacity floats) and the split
// partials: one partial of 27 sums per cache and group of 4 splits; with the
// surrogate's spot check the partials of its list (pitch 2 x set capacity)
// must fit at least one group too.
int ensure_light_buffers(drv_ctx* ctx, uint32_t n_blocks, bool use_net) {
  const size_t cap = static_cast<size_t>(ctx->cfg.max_caches);
  int rc;
  if (ctx->cfg.indirect_shadow && n_blocks > ctx->vis_blocks_cap) {
    dev_free(ctx->visibility);
    ctx->vis_blocks_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->visibility, cap * n_blocks)) != DRV_OK) return rc;
    ctx->vis_blocks_cap = n_blocks;
  }
  size_t cols = drvk::gather_partial_columns(static_cast<uint32_t>(ctx->tune.gather_max_splits),
                                             static_cast<uint32_t>(cap));
  if (use_net && ctx->sur.spot_fraction > 0.0f) cols = std::max(cols, 2 * spot_set_capacity(ctx));
  if (ctx->partial_cols < cols) {
    dev_free(ctx->partial);
    ctx->partial_cols = 0;
    if ((rc = dev_alloc(ctx, &ctx->partial, cols * 27)) != DRV_OK) return rc;
    ctx->partial_cols = cols;
  }
  return DRV_OK;
}

// validate_light plus the VPL / shadow block buffers the light needs.
int prepare_light(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light, uint32_t* n_vpl_out,
                  uint32_t* interval_out, uint32_t* n_blocks_out) {
  int rc = validate_light(ctx, rsm, light, n_vpl_out, interval_out, n_blocks_out);
  if (rc != DRV_OK) return rc;
  const uint32_t n_vpl = *n_vpl_out, n_blocks = *n_blocks_out;
  const bool shadow = ctx->cfg.indirect_shadow != 0;
  if (n_vpl > ctx->vpl_cap) {
    dev_free(ctx->vpls);
    if ((rc = dev_alloc(ctx, &ctx->vpls, n_vpl)) != DRV_OK) return rc;
    ctx->vpl_cap = n_vpl;
  }
  if (shadow && n_blocks > ctx->block_cap) {
    dev_free(ctx->blocks);
    dev_free(ctx->march);
    ctx->block_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->blocks, n_blocks)) != DRV_OK) return rc;
    if ((rc = dev_alloc(ctx, &ctx->march, static_cast<size_t>(n_blocks) * drvk::kMarchSteps)) != DRV_OK) return rc;
    ctx->block_cap = n_blocks;
  }
  return DRV_OK;
}

}  // namespace

extern "C" {

int drv_light_caches(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light) {
  uint32_t n_vpl = 0, interval = 0, n_blocks = 0;
  int rc = prepare_light(ctx, rsm, light, &n_vpl, &interval, &n_blocks);
  if (rc != DRV_OK) return rc;
  // The VPL preparation reads no voxels, so it runs while a voxelization on
  // the side stream finishes; the cone trace joins it.  With stage timing on
  // the join comes first, so no stage's time includes the wait.
  if (ctx->profiling) join_voxels_for_read(ctx);
  const bool shadow = ctx->cfg.indirect_shadow != 0;
  const size_t cap = static_cast<size_t>(ctx->cfg.max_caches);
  const bool use_net = shadow && ctx->sur.enabled && ctx->mlp_ready;
  if ((rc = ensure_light_buffers(ctx, n_blocks, use_net)) != DRV_OK) return rc;

  const int stride = drvk::entry_stride(ctx->cfg.sh_bands);
  const int grid = persistent_grid(ctx);
  uint64_t* steps = nullptr;
  if (ctx->profiling && ctx->profiling_level >= 2) {
    if (!ctx->step_counter && (rc = dev_alloc(ctx, &ctx->step_counter, 3)) != DRV_OK) return rc;
    drvk::launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->step_counter), 6, 0u, ctx->stream);
    steps = ctx->step_counter;
  }
  record(ctx, 5);
  const int specular = ctx->cfg.indirect_specular;
  if (specular && n_vpl > ctx->spec.vpls_lanes_cap) {
    dev_free(ctx->spec.vpls_lanes);
    ctx->spec.vpls_lanes_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->spec.vpls_lanes, n_vpl)) != DRV_OK) return rc;
    ctx->spec.vpls_lanes_cap = n_vpl;
  }
  // The gather's split plan rides on the VPL preparation (one launch less).
  const drvk::GatherPlan plan = drvk::gather_plan(ctx->counter, interval,
                                                  static_cast<uint32_t>(ctx->tune.gather_max_splits), grid);
  const drvk::VoxelView vview = voxel_view(ctx);
  drvk::launch_prepare_vpls(*rsm, *light, ctx->vpls, shadow ? ctx->blocks : nullptr, specular,
                            specular ? ctx->spec.vpls_lanes : nullptr, &plan, ctx->stream,
                            shadow ? ctx->march : nullptr, &vview);
  record(ctx, 6);
  join_voxels_for_read(ctx);
  if (use_net) {
    if ((rc = surrogate_visibility(ctx, n_blocks, n_vpl, interval, grid)) != DRV_OK) return rc;
    steps = nullptr;
  } else if (shadow) {
    drvk::launch_visibility(ctx->vol, vview, ctx->caches, stride, ctx->counter, ctx->blocks, n_blocks,
                            ctx->visibility, static_cast<uint32_t>(cap), steps, ctx->tune, grid, ctx->stream,
                            ctx->march);
    ctx->label_caches = ctx->caches;
    ctx->label_counter = ctx->counter;
    ctx->label_vis = ctx->visibility;
    ctx->label_pitch = static_cast<uint32_t>(cap);
  }
  record(ctx, 7);
  voxels_read(ctx);  // after the stage's end event, outside the visibility interval
  drvk::launch_gather(ctx->shc, ctx->cfg.sh_bands, ctx->caches, stride, ctx->counter, ctx->vpls, n_vpl,
                      ctx->visibility, static_cast<uint32_t>(cap), interval, shadow ? 1 : 0,
                      ctx->partial, static_cast<uint32_t>(cap), static_cast<uint32_t>(ctx->tune.gather_max_splits),
                      grid, false, ctx->stream);  // planned by the VPL preparation
  record(ctx, 8);
  drvk::launch_reduce_sh(ctx->shc, ctx->cfg.sh_bands, ctx->caches, stride, ctx->counter, ctx->partial,
                         static_cast<uint32_t>(cap), ctx->cu_count * 8, ctx->stream);
  record(ctx, 9);
  if (specular) {
    // renderer.cpp:553-558: LightCachesRSM with the env-map writes, then
    // PrepareSpecularEnvmaps.
    drvk::launch_specular(ctx->frame.camera_position, ctx->caches, stride, ctx->counter, ctx->spec.vpls_lanes, n_vpl,
                          shadow ? ctx->visibility : nullptr, static_cast<uint32_t>(cap), interval, ctx->spec,
                          ctx->spec_first ? 1 : 0, ctx->cu_count * 8, ctx->stream);
    ctx->spec_first = false;
    record(ctx, 12);
  }
  ctx->counted_steps = steps != nullptr;
  ctx->last_n_blocks = n_blocks;
  return check_hip(ctx, hipGetLastError(), "light_caches");
}

// drv_light_caches without any device allocation: the fused work-queue kernel
// with its records, descriptor and partials in the library's static scratch
// array of this device (see k_static.hip).
int drv_light_caches_noalloc(drv_ctx* ctx, const drv_rsm* rsm, const drv_spot_light* light) {
  uint32_t n_vpl = 0, interval = 0, n_blocks = 0;
  int rc = validate_light(ctx, rsm, light, &n_vpl, &interval, &n_blocks);
  if (rc != DRV_OK) return rc;
  if (n_vpl > drvk::kStaticMaxVpls)
    return fail(ctx, DRV_ERR_INVALID_ARG,
                "light_caches_noalloc: " + std::to_string(n_vpl) + " VPLs exceed the static scratch limit of " +
                    std::to_string(drvk::kStaticMaxVpls) + " (RSM read resolution <= 256)");
  if (ctx->cfg.indirect_specular)
    return fail(ctx, DRV_ERR_STATE, "light_caches_noalloc: indirect specular needs drv_light_caches");
  if (ctx->sur.enabled && ctx->mlp_ready && ctx->cfg.indirect_shadow)
    return fail(ctx, DRV_ERR_STATE, "light_caches_noalloc: the surrogate needs drv_light_caches (this path cone-traces)");
  int cur = -1;
  DRV_HIP(ctx, hipGetDevice(&cur));
  if (cur != ctx->device) return fail(ctx, DRV_ERR_STATE, "light_caches_noalloc: the context's device is not current");
  unsigned char* base = drvk::static_scratch(ctx->device);
  if (!base) return fail(ctx, DRV_ERR_STATE, "light_caches_noalloc: static scratch not resolved on this device");
  const bool shadow = ctx->cfg.indirect_shadow != 0;
  drvk::BatchView v;
  std::memset(&v, 0, sizeof(v));
  v.caches = ctx->caches;
  v.counter = ctx->counter;
  v.vox = voxel_view(ctx);
  for (int i = 0; i < 3; ++i) v.vmin[i] = ctx->vol.volume_world_min[i];
  v.voxel_size = ctx->vol.voxel_size_in_world;
  v.n_vpl = n_vpl;
  v.interval = interval;
  join_voxels_for_read(ctx);
  drvk::launch_light_static(base, *rsm, *light, v, drvk::entry_stride(ctx->cfg.sh_bands), ctx->cfg.sh_bands,
                            shadow ? 1 : 0, ctx->shc, std::max(1, ctx->cu_count * 64 / 4), ctx->stream);
  voxels_read(ctx);
  // No per-pair visibility or training labels are left behind.
  ctx->last_n_blocks = 0;
  ctx->label_caches = nullptr;
  ctx->counted_steps = false;
  return check_hip(ctx, hipGetLastError(), "light_caches_noalloc");
}

// ---- stage D ----------------------------------------------------------

int drv_apply_caches(drv_ctx* ctx, const drv_gbuffer* gb, float* out_rgb, int32_t accumulate) {
  if (!ctx || !gb || !gb->depth || !gb->normal || !gb->albedo || !out_rgb) return DRV_ERR_INVALID_ARG;
  if (gb->width != ctx->cfg.width || gb->height != ctx->cfg.height)
    return fail(ctx, DRV_ERR_INVALID_ARG, "apply: G-buffer size differs from the configured backbuffer");
  if (!ctx->allocated) return fail(ctx, DRV_ERR_STATE, "apply: no allocation yet");
  record(ctx, 10);
  drvk::launch_apply(ctx->cfg, ctx->vol, ctx->frame, ctx->shc, *gb, ctx->cav, ctx->caches, ctx->counter,
                     ctx->srgb_lut, ctx->cfg.indirect_specular ? &ctx->spec : nullptr, out_rgb, accumulate,
                     ctx->stream);
  record(ctx, 11);
  return check_hip(ctx, hipGetLastError(), "apply");
}

// ---- specular atlas -----------------------------------------------------

int drv_get_specular_atlas(drv_ctx* ctx, int32_t* width, int32_t* caches_per_row, int32_t* levels) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  if (!ctx->cfg.indirect_specular) return fail(ctx, DRV_ERR_STATE, "specular atlas: indirect_specular is off");
  if (width) *width = static_cast<int32_t>(ctx->spec.width);
  if (caches_per_row) *caches_per_row = static_cast<int32_t>(ctx->spec.per_row);
  if (levels) *levels = ctx->spec.levels;
  return DRV_OK;
}

int drv_read_specular_atlas(drv_ctx* ctx, int32_t level, uint32_t* host_out) {
  if (!ctx || !host_out) return DRV_ERR_INVALID_ARG;
  if (!ctx->cfg.indirect_specular) return fail(ctx, DRV_ERR_STATE, "specular atlas: indirect_specular is off");
  if (level < 0 || level >= ctx->spec.levels) return fail(ctx, DRV_ERR_INVALID_ARG, "specular atlas: no such level");
  const uint64_t wl = ctx->spec.width >> level;
  DRV_HIP(ctx, hipMemcpyAsync(host_out, ctx->spec.atlas + ctx->spec.level_offset[level], wl * wl * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

// ---- introspection ------------------------------------------------------

int drv_get_cache_count(drv_ctx* ctx, uint32_t* demanded, uint32_t* live) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  Counter c{};
  DRV_HIP(ctx, hipMemcpyAsync(&c, ctx->counter, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (demanded) *demanded = c.total;
  if (live) *live = std::min<uint32_t>(c.total, static_cast<uint32_t>(ctx->cfg.max_caches));
  return DRV_OK;
}

int drv_read_caches(drv_ctx* ctx, float* host_out, uint32_t count) {
  if (!ctx || (!host_out && count)) return DRV_ERR_INVALID_ARG;
  if (count > static_cast<uint32_t>(ctx->cfg.max_caches)) return fail(ctx, DRV_ERR_INVALID_ARG, "read_caches: count > capacity");
  const size_t bytes = sizeof(float) * drvk::entry_stride(ctx->cfg.sh_bands) * count;
  if (bytes) DRV_HIP(ctx, hipMemcpyAsync(host_out, ctx->caches, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_read_cav(drv_ctx* ctx, uint32_t* host_out) {
  if (!ctx || !host_out) return DRV_ERR_INVALID_ARG;
  DRV_HIP(ctx, hipMemcpyAsync(host_out, ctx->cav, ctx->cav_cells * 4, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_read_voxels(drv_ctx* ctx, int32_t level, uint8_t* host_out) {
  if (!ctx || !host_out) return DRV_ERR_INVALID_ARG;
  if (level < -1 || level >= ctx->vox.levels) return fail(ctx, DRV_ERR_INVALID_ARG, "read_voxels: bad level");
  // level -1 reads the voxelization target (before the blend).
  join_voxels_for_read(ctx);
  const uint8_t* src = level < 0 ? ctx->vox_target : ctx->vox_mips + ctx->vox.mip_offset[level];
  const size_t w = static_cast<size_t>(ctx->vox.res >> (level < 0 ? 0 : level));
  DRV_HIP(ctx, hipMemcpyAsync(host_out, src, w * w * w, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_read_visibility(drv_ctx* ctx, float* host_out, uint64_t count) {
  if (!ctx || (!host_out && count)) return DRV_ERR_INVALID_ARG;
  uint32_t total = 0, live = 0;
  int rc = drv_get_cache_count(ctx, &total, &live);
  if (rc != DRV_OK) return rc;
  const uint64_t nb = ctx->last_n_blocks;
  if (count > static_cast<uint64_t>(live) * nb) return fail(ctx, DRV_ERR_INVALID_ARG, "read_visibility: count too large");
  if (nb == 0 || count == 0) return DRV_OK;
  // Device layout is block-major with pitch max_caches; the host view is
  // cache-major [live][n_blocks].
  const size_t cap = static_cast<size_t>(ctx->cfg.max_caches);
  std::vector<float> buf(static_cast<size_t>(nb) * live);
  DRV_HIP(ctx, hipMemcpy2DAsync(buf.data(), sizeof(float) * live, ctx->visibility, sizeof(float) * cap,
                                sizeof(float) * live, nb, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t c = i / nb, b = i % nb;
    host_out[i] = buf[b * live + c];
  }
  return DRV_OK;
}

// ---- inputs rendered on the device -------------------------------------------

namespace {

size_t rsm_chain_texels(int res) {
  size_t n = 0;
  for (int r = res; r >= 1; r /= 2) n += static_cast<size_t>(r) * r;
  return n;
}

int ensure_raster(drv_ctx* ctx, size_t pixels, size_t n_tri) {
  int rc;
  if ((rc = ensure(ctx, &ctx->rz, &ctx->raster_cap, pixels)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->rbig, &ctx->rbig_cap, n_tri + 1)) != DRV_OK) return rc;
  return ensure(ctx, &ctx->rtri, &ctx->raster_tri_cap, pixels);
}

bool valid_scene(const drv_scene* s) {
  return s && (s->n_tri == 0 || (s->positions && s->albedo && s->normals));
}

}  // namespace

int drv_render_gbuffer(drv_ctx* ctx, const drv_scene* scene, const float view_projection[16], drv_gbuffer* out) {
  if (!ctx || !valid_scene(scene) || !view_projection || !out) return DRV_ERR_INVALID_ARG;
  const int w = ctx->cfg.width, h = ctx->cfg.height;
  const size_t n = static_cast<size_t>(w) * h;
  int rc;
  if ((rc = ensure_raster(ctx, n, scene->n_tri)) != DRV_OK) return rc;
  if (!ctx->g_depth) {
    if ((rc = dev_alloc(ctx, &ctx->g_depth, n)) != DRV_OK) return rc;
    if ((rc = dev_alloc(ctx, &ctx->g_normal, 2 * n)) != DRV_OK) return rc;
    if ((rc = dev_alloc(ctx, &ctx->g_albedo, 3 * n)) != DRV_OK) return rc;
  }
  drvk::launch_render_gbuffer(scene->positions, scene->albedo, scene->normals, scene->n_tri, view_projection, w, h,
                              ctx->rz, ctx->rtri, ctx->rbig, ctx->g_depth, ctx->g_normal, ctx->g_albedo, ctx->stream);
  out->depth = ctx->g_depth;
  out->normal = ctx->g_normal;
  out->albedo = ctx->g_albedo;
  out->width = w;
  out->height = h;
  out->roughness_metallic = nullptr;  // the scenes carry no material maps
  return check_hip(ctx, hipGetLastError(), "render_gbuffer");
}

int drv_render_rsm(drv_ctx* ctx, const drv_scene* scene, const drv_spot_light* light, drv_rsm* out) {
  if (!ctx || !valid_scene(scene) || !light || !out) return DRV_ERR_INVALID_ARG;
  const int res = light->rsm_render_resolution;
  if (!is_pow2(res) || res > 8192 || light->rsm_read_resolution < 1 || res % light->rsm_read_resolution != 0)
    return fail(ctx, DRV_ERR_INVALID_ARG, "render_rsm: render resolution must be a power of two <= 8192");
  if (light->rsm_render_height > 0 && light->rsm_render_height != res)
    return fail(ctx, DRV_ERR_INVALID_ARG, "render_rsm: the device RSM producer renders square RSMs (rsm_wide 0)");
  const size_t chain = rsm_chain_texels(res);
  int rc;
  if ((rc = ensure_raster(ctx, static_cast<size_t>(res) * res, scene->n_tri)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->rsm_flux, &ctx->rsm_cap, 3 * chain)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->rsm_normal, &ctx->rsm_normal_cap, 2 * chain)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->rsm_dls, &ctx->rsm_dls_cap, 2 * chain)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->rsm_shadow, &ctx->rsm_shadow_cap, static_cast<size_t>(res) * res)) != DRV_OK) return rc;
  drvk::launch_render_rsm(scene->positions, scene->albedo, scene->normals, scene->n_tri, *light, ctx->rz, ctx->rtri,
                          ctx->rbig, ctx->rsm_flux, ctx->rsm_normal, ctx->rsm_dls, ctx->rsm_shadow, ctx->stream);
  ctx->rsm_res = res;
  // Level l starts after the texels of levels 0..l-1.
  const int read_lod = log2i(res / light->rsm_read_resolution);
  const int shadow_lod = std::min(read_lod + static_cast<int>(light->indirect_shadow_lod), log2i(res));
  const size_t o_read = rsm_chain_texels(res) - rsm_chain_texels(res >> read_lod);
  const size_t o_shadow = rsm_chain_texels(res) - rsm_chain_texels(res >> shadow_lod);
  out->flux = ctx->rsm_flux + 3 * o_read;
  out->normal = ctx->rsm_normal + 2 * o_read;
  out->depth_lin_sq = ctx->rsm_dls + 2 * o_read;
  out->depth_lin_sq_shadow = ctx->rsm_dls + 2 * o_shadow;
  out->read_resolution = light->rsm_read_resolution;
  out->read_height = light->rsm_read_resolution;
  return check_hip(ctx, hipGetLastError(), "render_rsm");
}

int drv_read_gbuffer(drv_ctx* ctx, float* depth, int16_t* normal, uint8_t* albedo) {
  if (!ctx || !depth || !normal || !albedo) return DRV_ERR_INVALID_ARG;
  if (!ctx->g_depth) return fail(ctx, DRV_ERR_STATE, "read_gbuffer: no G-buffer rendered yet");
  const size_t n = static_cast<size_t>(ctx->cfg.width) * ctx->cfg.height;
  DRV_HIP(ctx, hipMemcpyAsync(depth, ctx->g_depth, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipMemcpyAsync(normal, ctx->g_normal, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipMemcpyAsync(albedo, ctx->g_albedo, 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_read_rsm(drv_ctx* ctx, int32_t level, uint16_t* flux, int16_t* normal, uint16_t* depth_lin_sq) {
  if (!ctx || !flux || !normal || !depth_lin_sq) return DRV_ERR_INVALID_ARG;
  if (ctx->rsm_res == 0) return fail(ctx, DRV_ERR_STATE, "read_rsm: no RSM rendered yet");
  if (level < 0 || (ctx->rsm_res >> level) < 1) return fail(ctx, DRV_ERR_INVALID_ARG, "read_rsm: bad level");
  const size_t off = rsm_chain_texels(ctx->rsm_res) - rsm_chain_texels(ctx->rsm_res >> level);
  const size_t n = static_cast<size_t>(ctx->rsm_res >> level) * static_cast<size_t>(ctx->rsm_res >> level);
  DRV_HIP(ctx, hipMemcpyAsync(flux, ctx->rsm_flux + 3 * off, 6 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipMemcpyAsync(normal, ctx->rsm_normal + 2 * off, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipMemcpyAsync(depth_lin_sq, ctx->rsm_dls + 2 * off, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_get_shadow_map(drv_ctx* ctx, const float** shadow_map, int32_t* resolution) {
  if (!ctx || !shadow_map || !resolution) return DRV_ERR_INVALID_ARG;
  if (ctx->rsm_res == 0) return fail(ctx, DRV_ERR_STATE, "get_shadow_map: no RSM rendered yet");
  *shadow_map = ctx->rsm_shadow;
  *resolution = ctx->rsm_res;
  return DRV_OK;
}

int drv_read_shadow_map(drv_ctx* ctx, float* host_out) {
  if (!ctx || !host_out) return DRV_ERR_INVALID_ARG;
  if (ctx->rsm_res == 0) return fail(ctx, DRV_ERR_STATE, "read_shadow_map: no RSM rendered yet");
  const size_t n = static_cast<size_t>(ctx->rsm_res) * ctx->rsm_res;
  DRV_HIP(ctx, hipMemcpyAsync(host_out, ctx->rsm_shadow, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

// ---- direct lighting and HDR output ---------------------------------------------

int drv_direct_lighting(drv_ctx* ctx, const drv_gbuffer* gb, const uint8_t* roughness_metallic,
                        const drv_spot_light* light, const float* shadow_map, int32_t shadow_resolution,
                        float normal_offset_bias, float shadow_bias, float* out_rgb, int32_t accumulate) {
  if (!ctx || !gb || !gb->depth || !gb->normal || !gb->albedo || !light || !shadow_map || shadow_resolution <= 0 ||
      !out_rgb)
    return DRV_ERR_INVALID_ARG;
  if (gb->width != ctx->cfg.width || gb->height != ctx->cfg.height)
    return fail(ctx, DRV_ERR_INVALID_ARG, "direct_lighting: G-buffer size differs from the configured backbuffer");
  if (!ctx->have_frame) return fail(ctx, DRV_ERR_STATE, "direct_lighting: frame not set");
  drvk::launch_direct(ctx->cfg, ctx->frame, *light, *gb, roughness_metallic, ctx->srgb_lut, shadow_map,
                      shadow_resolution, normal_offset_bias, shadow_bias, out_rgb, accumulate, ctx->stream);
  return check_hip(ctx, hipGetLastError(), "direct_lighting");
}

int drv_tonemap(drv_ctx* ctx, const float* hdr_rgb, float exposure, float lmax, float* out_rgb) {
  if (!ctx || !hdr_rgb || !out_rgb || !(lmax > 0.0f)) return DRV_ERR_INVALID_ARG;
  const size_t n = static_cast<size_t>(ctx->cfg.width) * ctx->cfg.height * 3;
  // Renderer::SetTonemapLMax (renderer.cpp:1223-1227): DragoDivider = log2(LMax + 1)
  drvk::launch_tonemap(hdr_rgb, n, exposure, std::log2(lmax + 1.0f), out_rgb, ctx->stream);
  return check_hip(ctx, hipGetLastError(), "tonemap");
}

int drv_save_pfm(drv_ctx* ctx, const float* hdr_rgb, const char* path) {
  if (!ctx || !hdr_rgb || !path) return DRV_ERR_INVALID_ARG;
  const int w = ctx->cfg.width, h = ctx->cfg.height;
  std::vector<float> img(static_cast<size_t>(w) * h * 3);
  DRV_HIP(ctx, hipMemcpyAsync(img.data(), hdr_rgb, img.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // WritePfm (rendering/hdrimage.cpp:6-30): "PF", size, scale -1 (little
  // endian), then RGB floats row by row from the bottom row up.
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ctx, DRV_ERR_INVALID_ARG, std::string("save_pfm: cannot open ") + path);
  std::fprintf(f, "PF\n%d %d\n-1.000000\n", w, h);
  const size_t wrote = std::fwrite(img.data(), sizeof(float), img.size(), f);
  const bool ok = std::fclose(f) == 0 && wrote == img.size();
  return ok ? DRV_OK : fail(ctx, DRV_ERR_HIP, std::string("save_pfm: write failed: ") + path);
}

// ---- brute-force RSM reference ------------------------------------------------

int drv_bruteforce_rsm(drv_ctx* ctx, const drv_gbuffer* gb, const uint8_t* roughness_metallic, const drv_rsm* rsm,
                       const drv_spot_light* light, int32_t diffuse_only, float* out_rgb, int32_t accumulate) {
  if (!ctx || !gb || !gb->depth || !gb->normal || !gb->albedo || !rsm || !rsm->flux || !rsm->normal ||
      !rsm->depth_lin_sq || !light || !out_rgb)
    return DRV_ERR_INVALID_ARG;
  if (gb->width != ctx->cfg.width || gb->height != ctx->cfg.height)
    return fail(ctx, DRV_ERR_INVALID_ARG, "bruteforce_rsm: G-buffer size differs from the configured backbuffer");
  if (!ctx->have_frame) return fail(ctx, DRV_ERR_STATE, "bruteforce_rsm: frame not set");
  const int R = rsm->read_resolution;
  const int H = rsm->read_height > 0 ? rsm->read_height : R;
  const int LH = light->rsm_read_height > 0 ? light->rsm_read_height : light->rsm_read_resolution;
  if (!is_pow2(R) || R > 65536 || R != light->rsm_read_resolution || (H != R && 2 * H != R) || H != LH)
    return fail(ctx, DRV_ERR_INVALID_ARG, "bruteforce_rsm: read resolution must be a power of two matching the light");
  const uint32_t n = static_cast<uint32_t>(R) * static_cast<uint32_t>(H);
  int rc;
  if (n > ctx->vpl_cap) {
    dev_free(ctx->vpls);
    ctx->vpl_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->vpls, n)) != DRV_OK) return rc;
    ctx->vpl_cap = n;
  }
  // VAL records (position, flux, normal) of the read level; the VPL order
  // does not matter to a plain sum.
  drvk::launch_prepare_vpls(*rsm, *light, ctx->vpls, nullptr, 0, nullptr, nullptr, ctx->stream);
  drvk::launch_bruteforce(ctx->cfg, ctx->frame, *light, *gb, roughness_metallic, ctx->srgb_lut, ctx->vpls, n,
                          diffuse_only, out_rgb, accumulate, ctx->stream);
  return check_hip(ctx, hipGetLastError(), "bruteforce_rsm");
}

// ---- surrogate visibility network -------------------------------------------

namespace {

int mlp_alloc(drv_ctx* ctx) {
  if (ctx->mlp_params) return DRV_OK;
  int rc;
  if ((rc = dev_alloc(ctx, &ctx->mlp_params, drvk::kMlpParams)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, &ctx->mlp_m, drvk::kMlpParams)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, &ctx->mlp_v, drvk::kMlpParams)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, &ctx->mlp_grad, drvk::kMlpParams + 1)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, reinterpret_cast<uint8_t**>(&ctx->mlp_pack), drvk::kMlpPackBytes)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, reinterpret_cast<uint16_t**>(&ctx->mlp_train_w), drvk::kMlpTrainW)) != DRV_OK) return rc;
  if ((rc = dev_alloc(ctx, &ctx->mlp_n, 1)) != DRV_OK) return rc;
  return DRV_OK;
}

int require_net(drv_ctx* ctx) {
  if (!ctx->mlp_ready) return fail(ctx, DRV_ERR_STATE, "surrogate: network parameters not set");
  return DRV_OK;
}

// Features (and labels) need the shadow blocks and voxels of the last light pass.
int require_features(drv_ctx* ctx) {
  if (ctx->last_n_blocks == 0) return fail(ctx, DRV_ERR_STATE, "surrogate: no shadowed light pass yet");
  return DRV_OK;
}

// Gradient of the mean loss over the `batch` samples listed in mlp_idx: the
// training kernel's per-wave partials, reduced in a fixed order into mlp_grad
// (last entry: summed loss), and for lr > 0 one Adam step plus a repack.
int mlp_step(drv_ctx* ctx, const float* x, const float* y, uint32_t batch, float lr) {
  const uint32_t waves = drvk::mlp_train_waves(batch, ctx->cu_count);
  int rc;
  if ((rc = ensure(ctx, &ctx->mlp_partial, &ctx->mlp_partial_waves, static_cast<size_t>(waves) * (drvk::kMlpParams + 1))) !=
      DRV_OK)
    return rc;
  drvk::launch_mlp_train(x, y, ctx->mlp_idx, batch, ctx->mlp_params, ctx->mlp_train_w, ctx->mlp_partial, waves,
                         ctx->stream);
  float bc1 = 1.0f, bc2 = 1.0f;
  if (lr > 0.0f) {
    ctx->adam_t += 1;
    bc1 = 1.0f - static_cast<float>(std::pow(0.9, ctx->adam_t));
    bc2 = 1.0f - static_cast<float>(std::pow(0.999, ctx->adam_t));
  }
  drvk::launch_mlp_adam(ctx->mlp_partial, waves, ctx->mlp_params, ctx->mlp_m, ctx->mlp_v, ctx->mlp_grad, lr, 0.9f,
                        0.999f, 1e-8f, bc1, bc2, ctx->stream);
  if (lr > 0.0f) drvk::launch_mlp_pack(ctx->mlp_params, ctx->mlp_pack, ctx->mlp_train_w, ctx->stream);
  return check_hip(ctx, hipGetLastError(), "mlp_step");
}

}  // namespace

int drv_mlp_set_params(drv_ctx* ctx, const float* host_params, uint32_t count) {
  if (!ctx || !host_params || count != static_cast<uint32_t>(drvk::kMlpParams)) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = mlp_alloc(ctx)) != DRV_OK) return rc;
  DRV_HIP(ctx, hipMemcpyAsync(ctx->mlp_params, host_params, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
  drvk::launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->mlp_m), drvk::kMlpParams, 0u, ctx->stream);
  drvk::launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->mlp_v), drvk::kMlpParams, 0u, ctx->stream);
  drvk::launch_mlp_pack(ctx->mlp_params, ctx->mlp_pack, ctx->mlp_train_w, ctx->stream);
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->adam_t = 0;
  ctx->mlp_ready = true;
  return check_hip(ctx, hipGetLastError(), "mlp_set_params");
}

int drv_mlp_get_params(drv_ctx* ctx, float* host_params, uint32_t count) {
  if (!ctx || !host_params || count != static_cast<uint32_t>(drvk::kMlpParams)) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = require_net(ctx)) != DRV_OK) return rc;
  DRV_HIP(ctx, hipMemcpyAsync(host_params, ctx->mlp_params, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_set_surrogate(drv_ctx* ctx, const drv_surrogate* settings) {
  if (!ctx || !settings) return DRV_ERR_INVALID_ARG;
  if (!(settings->spot_fraction >= 0.0f && settings->spot_fraction <= 1.0f) || !(settings->fallback_threshold >= 0.0f))
    return fail(ctx, DRV_ERR_INVALID_ARG, "set_surrogate: spot fraction must lie in [0, 1], threshold >= 0");
  if (settings->enabled) {
    int rc;
    if ((rc = require_net(ctx)) != DRV_OK) return rc;
  }
  ctx->sur = *settings;
  return DRV_OK;
}

int drv_get_surrogate_stats(drv_ctx* ctx, drv_surrogate_stats* out) {
  if (!ctx || !out) return DRV_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  if (!ctx->spot_stats) return DRV_OK;
  uint32_t st[4] = {0, 0, 0, 0};
  DRV_HIP(ctx, hipMemcpyAsync(st, ctx->spot_stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(&out->spot_error, &st[0], 4);
  out->fell_back = static_cast<int32_t>(st[1]);
  out->spot_caches = st[2];
  return DRV_OK;
}

int drv_mlp_collect(drv_ctx* ctx, uint32_t n, uint32_t seed, float* x_dev, float* y_dev) {
  if (!ctx || (n && (!x_dev || !y_dev))) return DRV_ERR_INVALID_ARG;
  if (reinterpret_cast<uintptr_t>(x_dev) % 16 != 0) return fail(ctx, DRV_ERR_INVALID_ARG, "mlp_collect: x must be 16-byte aligned");
  int rc;
  if ((rc = require_features(ctx)) != DRV_OK) return rc;
  if (!ctx->label_caches) return fail(ctx, DRV_ERR_STATE, "mlp_collect: the last light pass has no cone-traced labels");
  join_voxels_for_read(ctx);
  if (n == 0) return DRV_OK;
  // Pairs are drawn over the label set's live caches (read on the device).
  Counter c{};
  DRV_HIP(ctx, hipMemcpyAsync(&c, ctx->label_counter, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (c.live == 0) return fail(ctx, DRV_ERR_STATE, "mlp_collect: no live caches");
  if ((rc = ensure(ctx, &ctx->mlp_pairs, &ctx->mlp_pairs_cap, 2 * static_cast<size_t>(n))) != DRV_OK) return rc;
  if (!ctx->mlp_n && (rc = dev_alloc(ctx, &ctx->mlp_n, 1)) != DRV_OK) return rc;
  drvk::launch_mlp_sample_pairs(ctx->label_counter, ctx->last_n_blocks, n, seed, ctx->mlp_pairs, ctx->mlp_n,
                                ctx->stream);
  const drvk::FeatureArgsHost fh = drvk::make_feature_args(ctx->vol, ctx->frame.camera_position, voxel_view(ctx));
  drvk::launch_mlp_features(fh, ctx->label_caches, drvk::entry_stride(ctx->cfg.sh_bands), c.live, ctx->blocks,
                            ctx->last_n_blocks, ctx->mlp_pairs, n, x_dev, ctx->label_vis, ctx->label_pitch, y_dev,
                            ctx->stream);
  voxels_read(ctx);
  return check_hip(ctx, hipGetLastError(), "mlp_collect");
}

int drv_mlp_features(drv_ctx* ctx, const uint32_t* pairs_dev, uint32_t n, float* x_dev) {
  if (!ctx || (n && (!pairs_dev || !x_dev))) return DRV_ERR_INVALID_ARG;
  if (reinterpret_cast<uintptr_t>(x_dev) % 16 != 0) return fail(ctx, DRV_ERR_INVALID_ARG, "mlp_features: x must be 16-byte aligned");
  int rc;
  if ((rc = require_features(ctx)) != DRV_OK) return rc;
  const drvk::FeatureArgsHost fh = drvk::make_feature_args(ctx->vol, ctx->frame.camera_position, voxel_view(ctx));
  join_voxels_for_read(ctx);
  // Pairs name caches of the current list (live count read on the host).
  uint32_t total = 0, live = 0;
  if ((rc = drv_get_cache_count(ctx, &total, &live)) != DRV_OK) return rc;
  drvk::launch_mlp_features(fh, ctx->caches, drvk::entry_stride(ctx->cfg.sh_bands), live, ctx->blocks,
                            ctx->last_n_blocks, pairs_dev, n, x_dev, nullptr, 0, nullptr, ctx->stream);
  voxels_read(ctx);
  return check_hip(ctx, hipGetLastError(), "mlp_features");
}

int drv_mlp_forward(drv_ctx* ctx, const float* x_dev, uint32_t n, float* p_dev) {
  if (!ctx || (n && (!x_dev || !p_dev))) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = require_net(ctx)) != DRV_OK) return rc;
  drvk::launch_mlp_forward(ctx->mlp_pack, ctx->mlp_params, x_dev, n, p_dev, network_grid(ctx), ctx->stream);
  return check_hip(ctx, hipGetLastError(), "mlp_forward");
}

int drv_mlp_gradient(drv_ctx* ctx, const float* x_dev, const float* y_dev, uint32_t n, float* grad_host) {
  if (!ctx || !x_dev || !y_dev || !grad_host || n == 0) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = require_net(ctx)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->mlp_idx, &ctx->mlp_idx_cap, n)) != DRV_OK) return rc;
  drvk::launch_mlp_batch_indices(n, n, 0u, ctx->mlp_idx, true, ctx->stream);
  if ((rc = mlp_step(ctx, x_dev, y_dev, n, 0.0f)) != DRV_OK) return rc;
  DRV_HIP(ctx, hipMemcpyAsync(grad_host, ctx->mlp_grad, sizeof(float) * (drvk::kMlpParams + 1), hipMemcpyDeviceToHost,
                              ctx->stream));
  DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DRV_OK;
}

int drv_mlp_train(drv_ctx* ctx, const float* x_dev, const float* y_dev, uint32_t n, uint32_t batch, uint32_t steps,
                  float lr, uint32_t seed, float* mean_loss_host) {
  if (!ctx || !x_dev || !y_dev || n == 0 || batch == 0 || !(lr > 0.0f)) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = require_net(ctx)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->mlp_idx, &ctx->mlp_idx_cap, batch)) != DRV_OK) return rc;
  for (uint32_t t = 0; t < steps; ++t) {
    drvk::launch_mlp_batch_indices(n, batch, seed * 0x9e3779b1u + t, ctx->mlp_idx, false, ctx->stream);
    if ((rc = mlp_step(ctx, x_dev, y_dev, batch, lr)) != DRV_OK) return rc;
  }
  if (mean_loss_host) {
    float loss = 0.0f;
    if (steps > 0) {
      DRV_HIP(ctx, hipMemcpyAsync(&loss, ctx->mlp_grad + drvk::kMlpParams, sizeof(float), hipMemcpyDeviceToHost,
                                  ctx->stream));
      DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    *mean_loss_host = loss / static_cast<float>(batch);
  }
  return check_hip(ctx, hipGetLastError(), "mlp_train");
}

int drv_mlp_finetune(drv_ctx* ctx, uint32_t n_samples, uint32_t batch, uint32_t steps, float lr, uint32_t seed,
                     float* mean_loss_host) {
  if (!ctx || n_samples == 0) return DRV_ERR_INVALID_ARG;
  int rc;
  if ((rc = require_net(ctx)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->ft_x, &ctx->ft_x_cap, static_cast<size_t>(n_samples) * 32)) != DRV_OK) return rc;
  if ((rc = ensure(ctx, &ctx->ft_y, &ctx->ft_y_cap, static_cast<size_t>(n_samples))) != DRV_OK) return rc;
  if ((rc = drv_mlp_collect(ctx, n_samples, seed, ctx->ft_x, ctx->ft_y)) != DRV_OK) return rc;
  return drv_mlp_train(ctx, ctx->ft_x, ctx->ft_y, n_samples, batch, steps, lr, seed ^ 0x5bd1e995u, mean_loss_host);
}

// ---- batched views ----------------------------------------------------------

}  // extern "C"

struct drv_batch {
  std::vector<drv_ctx*> views;
  hipStream_t stream = nullptr;
  // Recorded once after a batched light pass's last voxel read, for the
  // views with the voxel overlap on (one event instead of one per view).
  hipEvent_t vox_free_ev = nullptr;
  int stride = 16;
  int sh_bands = 1;
  int shadow = 0;
  int cu_count = 256;
  int32_t cpw = 64;  // caches per wave of the fused queue; 0 per-view passes; DRV_BATCH_SPLIT_QUEUE or
                     // DRV_BATCH_SPLIT_GATHER
  uint32_t vpl_tile = 4096;
  int waves_per_cu = 64;
  bool specular = false;  // some view keeps env-maps: per-view passes only
  drvk::BatchView* d_views = nullptr;
  std::vector<drvk::BatchView> h_views;  // as last uploaded
  drvk::BatchQueue* d_queue = nullptr;
  float* partial = nullptr;
  size_t partial_size = 0;  // floats
  // Split queue: device view descriptors (and their last upload) and the
  // item totals the plan kernel writes.
  void* d_split = nullptr;
  std::vector<uint8_t> h_split;
  drvk::SplitTotals* d_totals = nullptr;
  // Fine-tuning samples collected from every view's spot set.
  float* ft_x = nullptr;
  float* ft_y = nullptr;
  size_t ft_x_cap = 0, ft_y_cap = 0;
  std::string err;
};

namespace {

constexpr size_t kBatchPartialFloats = size_t(1) << 28;  // 1 GiB of split partials at most

int bfail(drv_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

// The fused queue cone-traces every pair; views that predict visibility
// with the surrogate need their own passes (caches per wave 0).
bool any_surrogate(const drv_batch* b) {
  for (const drv_ctx* v : b->views)
    if (v->sur.enabled && v->mlp_ready && v->cfg.indirect_shadow) return true;
  return false;
}

bool is_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

// The voxel-free point of every view after a batched light pass: one event
// on the shared stream after the last cone trace of all views.
void batch_voxels_read(drv_batch* b) {
  bool any = false;
  for (const drv_ctx* c : b->views) any = any || c->vox_overlap;
  const bool ok = any && !is_capturing(b->stream) && hipEventRecord(b->vox_free_ev, b->stream) == hipSuccess;
  for (drv_ctx* c : b->views) {
    c->vox_free_valid = ok && c->vox_overlap;
    if (c->vox_free_valid) c->vox_free_wait = b->vox_free_ev;
  }
}

// The split queue: every view's buffers as its own light pass sizes them,
// the device descriptors (uploaded when they change), then one launch per
// stage over the items of all views.  Visibility and training labels stay
// behind as after drv_light_caches.
int batch_split(drv_batch* b, const drv_rsm* rsms, const drv_spot_light* lights) {
  const size_t n = b->views.size();
  const bool capturing = is_capturing(b->stream);
  std::vector<drvk::SplitViewHost> hv(n);
  uint32_t max_vpl = 64, max_march = 0;
  for (size_t v = 0; v < n; ++v) {
    drv_ctx* ctx = b->views[v];
    uint32_t n_vpl = 0, interval = 0, n_blocks = 0;
    const size_t vpl_had = ctx->vpl_cap, blk_had = ctx->block_cap, vis_had = ctx->vis_blocks_cap,
                 cols_had = ctx->partial_cols;
    int rc = prepare_light(ctx, &rsms[v], &lights[v], &n_vpl, &interval, &n_blocks);
    if (rc == DRV_OK) rc = ensure_light_buffers(ctx, n_blocks, false);
    if (rc != DRV_OK) return bfail(b, rc, "batch view " + std::to_string(v) + ": " + ctx->err);
    if (capturing && (ctx->vpl_cap != vpl_had || ctx->block_cap != blk_had || ctx->vis_blocks_cap != vis_had ||
                      ctx->partial_cols != cols_had))
      return bfail(b, DRV_ERR_STATE, "batch_light_caches: buffers grew during graph capture");
    const uint32_t cap = static_cast<uint32_t>(ctx->cfg.max_caches);
    drvk::SplitViewHost& d = hv[v];
    std::memset(&d, 0, sizeof(d));
    d.caches = ctx->caches;
    d.counter = ctx->counter;
    d.vpls = ctx->vpls;
    d.blocks = b->shadow ? ctx->blocks : nullptr;
    d.march = b->shadow ? ctx->march : nullptr;
    d.vis = b->shadow ? ctx->visibility : nullptr;
    d.vis_pitch = cap;
    d.partial = ctx->partial;
    d.partial_pitch = cap;
    d.max_splits = std::max<uint32_t>(1u, drvk::max_fitting_splits(ctx->partial_cols, cap,
                                                                   static_cast<uint32_t>(ctx->tune.gather_max_splits)));
    d.vox = voxel_view(ctx);
    for (int i = 0; i < 3; ++i) d.vmin[i] = ctx->vol.volume_world_min[i];
    d.voxel_size = ctx->vol.voxel_size_in_world;
    d.n_vpl = n_vpl;
    d.interval = interval;
    d.n_blocks = n_blocks;
    d.rsm = rsms[v];
    d.light = lights[v];
    max_vpl = std::max(max_vpl, n_vpl);
    if (b->shadow) max_march = std::max(max_march, n_blocks * static_cast<uint32_t>(drvk::kMarchSteps));
  }
  const size_t bytes = drvk::split_desc_bytes() * n;
  std::vector<uint8_t> img(bytes);
  drvk::build_split_descs(hv.data(), static_cast<uint32_t>(n), img.data());
  if (!b->d_split || !b->d_totals || img != b->h_split) {
    if (capturing) return bfail(b, DRV_ERR_STATE, "batch_light_caches: view descriptors changed during graph capture");
    if (!b->d_split && hipMalloc(&b->d_split, bytes) != hipSuccess)
      return bfail(b, DRV_ERR_NO_MEMORY, "batch_light_caches: descriptors do not fit");
    if (!b->d_totals && hipMalloc(reinterpret_cast<void**>(&b->d_totals), sizeof(drvk::SplitTotals)) != hipSuccess)
      return bfail(b, DRV_ERR_NO_MEMORY, "batch_light_caches: descriptors do not fit");
    // Ordered on the batch stream before this frame's kernels read them;
    // earlier frames' kernels that read the old descriptors are ordered
    // before the copy on the same stream.
    if (hipMemcpyAsync(b->d_split, img.data(), bytes, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return bfail(b, DRV_ERR_HIP, "batch_light_caches: descriptor upload failed");
    b->h_split.swap(img);
  }
  for (size_t v = 0; v < n; ++v) join_voxels_for_read(b->views[v]);
  drv_ctx* v0 = b->views[0];
  const int grid = persistent_grid(v0);
  const uint32_t nv = static_cast<uint32_t>(n);
  drvk::launch_batch_split_prepare(b->d_split, nv, b->d_totals, b->sh_bands, b->shadow, max_vpl, max_march,
                                   static_cast<uint32_t>(grid) * 4u, b->stream);
  if (b->shadow && b->cpw == DRV_BATCH_SPLIT_GATHER) {
    // Each view's visibility with the single-view kernel and launch shape.
    for (size_t v = 0; v < n; ++v) {
      drv_ctx* ctx = b->views[v];
      drvk::launch_visibility(ctx->vol, hv[v].vox, ctx->caches, b->stride, ctx->counter, ctx->blocks,
                              hv[v].n_blocks, ctx->visibility, hv[v].vis_pitch, nullptr, ctx->tune,
                              persistent_grid(ctx), b->stream, ctx->march);
    }
  } else if (b->shadow) {
    drvk::launch_batch_split_visibility(b->d_split, nv, b->d_totals, b->stride, grid, b->stream);
  }
  batch_voxels_read(b);
  drvk::launch_batch_split_gather(b->d_split, nv, b->d_totals, b->stride, b->sh_bands, b->shadow, v0->shc, grid,
                                  b->stream);
  for (size_t v = 0; v < n; ++v) {
    drv_ctx* ctx = b->views[v];
    ctx->last_n_blocks = hv[v].n_blocks;
    ctx->counted_steps = false;
    if (b->shadow) {
      ctx->label_caches = ctx->caches;
      ctx->label_counter = ctx->counter;
      ctx->label_vis = ctx->visibility;
      ctx->label_pitch = static_cast<uint32_t>(ctx->cfg.max_caches);
    } else {
      ctx->label_caches = nullptr;
    }
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, DRV_ERR_HIP, std::string("batch_light_caches: ") + hipGetErrorString(e));
  return DRV_OK;
}

}  // namespace

extern "C" {

int drv_batch_create(drv_ctx* const* views, int32_t n_views, drv_batch** out) {
  if (!views || n_views <= 0 || n_views > 65535 || !out) return DRV_ERR_INVALID_ARG;
  *out = nullptr;
  for (int i = 0; i < n_views; ++i) {
    if (!views[i]) return DRV_ERR_INVALID_ARG;
    if (views[i]->cfg.sh_bands != views[0]->cfg.sh_bands ||
        (views[i]->cfg.indirect_shadow != 0) != (views[0]->cfg.indirect_shadow != 0) ||
        views[i]->stream != views[0]->stream)
      return DRV_ERR_INVALID_ARG;
  }
  drv_batch* b = new (std::nothrow) drv_batch();
  if (!b) return DRV_ERR_NO_MEMORY;
  b->views.assign(views, views + n_views);
  b->stream = views[0]->stream;
  b->sh_bands = views[0]->cfg.sh_bands;
  b->stride = drvk::entry_stride(b->sh_bands);
  b->shadow = views[0]->cfg.indirect_shadow != 0 ? 1 : 0;
  b->cu_count = views[0]->cu_count;
  // The fused kernel writes SH only; a batch with env-map views lights every
  // view with its own passes.
  for (int i = 0; i < n_views; ++i) b->specular = b->specular || views[i]->cfg.indirect_specular != 0;
  if (b->specular) b->cpw = 0;
  b->h_views.resize(n_views);
  std::memset(b->h_views.data(), 0, sizeof(drvk::BatchView) * n_views);
  if (hipMalloc(reinterpret_cast<void**>(&b->d_views), sizeof(drvk::BatchView) * n_views) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_queue), sizeof(drvk::BatchQueue)) != hipSuccess) {
    drv_batch_destroy(b);
    return DRV_ERR_NO_MEMORY;
  }
  if (hipEventCreateWithFlags(&b->vox_free_ev, hipEventDisableTiming) != hipSuccess) {
    b->vox_free_ev = nullptr;
    drv_batch_destroy(b);
    return DRV_ERR_HIP;
  }
  *out = b;
  return DRV_OK;
}

int drv_batch_destroy(drv_batch* b) {
  if (!b) return DRV_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  else (void)hipDeviceSynchronize();
  // Views that would wait for the batch's voxel-free event fork from their
  // stream at their next voxelization instead.
  for (drv_ctx* c : b->views) {
    if (c->vox_free_wait == b->vox_free_ev) {
      c->vox_free_valid = false;
      c->vox_free_wait = c->vox_free_ev;
    }
  }
  if (b->vox_free_ev) (void)hipEventDestroy(b->vox_free_ev);
  dev_free(b->d_views);
  dev_free(b->d_queue);
  dev_free(b->partial);
  dev_free(b->ft_x);
  dev_free(b->ft_y);
  if (b->d_split) (void)hipFree(b->d_split);
  dev_free(b->d_totals);
  delete b;
  return DRV_OK;
}

const char* drv_batch_last_error(const drv_batch* b) { return b ? b->err.c_str() : "null batch"; }

int drv_batch_set_tuning(drv_batch* b, int32_t cpw, int32_t vpl_tile, int32_t waves_per_cu) {
  if (!b) return DRV_ERR_INVALID_ARG;
  if ((cpw != 64 && cpw != 32 && cpw != 16 && cpw != 0 && cpw != DRV_BATCH_SPLIT_QUEUE &&
       cpw != DRV_BATCH_SPLIT_GATHER) ||
      !is_pow2(vpl_tile) || vpl_tile < 16 || waves_per_cu < 4 || waves_per_cu > 64)
    return bfail(b, DRV_ERR_INVALID_ARG,
                 "batch_set_tuning: caches per wave 64/32/16 (0: per-view passes, -1: split queue, -2: split queue "
                 "with per-view visibility), VPL tile a power of two >= 16");
  if (b->specular && cpw != 0)
    return bfail(b, DRV_ERR_INVALID_ARG, "batch_set_tuning: views with indirect_specular need per-view passes (0)");
  if (any_surrogate(b) && cpw != 0)
    return bfail(b, DRV_ERR_STATE, "batch_set_tuning: views with the surrogate enabled need per-view passes (0)");
  b->cpw = cpw;
  b->vpl_tile = static_cast<uint32_t>(vpl_tile);
  b->waves_per_cu = waves_per_cu;
  return DRV_OK;
}

int drv_batch_get_tuning(const drv_batch* b, int32_t* cpw, int32_t* vpl_tile, int32_t* waves_per_cu) {
  if (!b) return DRV_ERR_INVALID_ARG;
  if (cpw) *cpw = b->cpw;
  if (vpl_tile) *vpl_tile = static_cast<int32_t>(b->vpl_tile);
  if (waves_per_cu) *waves_per_cu = b->waves_per_cu;
  return DRV_OK;
}

int drv_batch_light_caches(drv_batch* b, const drv_rsm* rsms, const drv_spot_light* lights) {
  if (!b || !rsms || !lights) return DRV_ERR_INVALID_ARG;
  const size_t n = b->views.size();
  if (b->cpw == 0) {
    // Per-view passes: each view's own cone trace, SH gather and reduction,
    // back to back on the shared stream (views large enough to fill the GPU
    // on their own gain nothing from the fused queue).
    struct Defer {  // the views' own voxel-free records, for the duration of the loop
      drv_batch* b;
      explicit Defer(drv_batch* bb) : b(bb) {
        for (drv_ctx* c : b->views) c->defer_vox_read = true;
      }
      ~Defer() {
        for (drv_ctx* c : b->views) c->defer_vox_read = false;
      }
    };
    {
      Defer defer(b);
      for (size_t v = 0; v < n; ++v) {
        const int rc = drv_light_caches(b->views[v], &rsms[v], &lights[v]);
        if (rc != DRV_OK) {
          // The views lit so far read voxels after their last voxel-free
          // point: their next voxelization waits for the whole stream.
          for (drv_ctx* c : b->views) c->vox_free_valid = false;
          return bfail(b, rc, "batch view " + std::to_string(v) + ": " + b->views[v]->err);
        }
      }
    }
    batch_voxels_read(b);
    return DRV_OK;
  }
  if (any_surrogate(b))
    return bfail(b, DRV_ERR_STATE,
                 "batch_light_caches: a view has the surrogate enabled; the fused and split queues only cone-trace "
                 "(select caches per wave 0)");
  if (b->cpw == DRV_BATCH_SPLIT_QUEUE || b->cpw == DRV_BATCH_SPLIT_GATHER) return batch_split(b, rsms, lights);
  const bool capturing = is_capturing(b->stream);
  size_t partial_need = 1;
  std::vector<drvk::BatchView> hv(n);
  std::memset(hv.data(), 0, sizeof(drvk::BatchView) * n);
  for (size_t v = 0; v < n; ++v) {
    drv_ctx* ctx = b->views[v];
    uint32_t n_vpl = 0, interval = 0, n_blocks = 0;
    const size_t vpl_had = ctx->vpl_cap, blk_had = ctx->block_cap;
    int rc = prepare_light(ctx, &rsms[v], &lights[v], &n_vpl, &interval, &n_blocks);
    if (rc != DRV_OK) return bfail(b, rc, "batch view " + std::to_string(v) + ": " + ctx->err);
    if (capturing && (ctx->vpl_cap != vpl_had || ctx->block_cap != blk_had))
      return bfail(b, DRV_ERR_STATE, "batch_light_caches: buffers grew during graph capture");
    if (interval > b->vpl_tile)
      return bfail(b, DRV_ERR_INVALID_ARG, "batch_light_caches: VPL tile shorter than a shadow block");
    drvk::BatchView& d = hv[v];
    d.caches = ctx->caches;
    d.counter = ctx->counter;
    d.vpls = ctx->vpls;
    d.blocks = b->shadow ? ctx->blocks : nullptr;
    d.vox = voxel_view(ctx);
    for (int i = 0; i < 3; ++i) d.vmin[i] = ctx->vol.volume_world_min[i];
    d.voxel_size = ctx->vol.voxel_size_in_world;
    d.n_vpl = n_vpl;
    d.interval = interval;
    const size_t splits = n_vpl / std::min<uint32_t>(b->vpl_tile, n_vpl);
    const size_t cpw = static_cast<size_t>(b->cpw);
    const size_t tiles = (static_cast<size_t>(ctx->cfg.max_caches) + cpw - 1) / cpw;
    if (splits > 1) partial_need += tiles * splits * 27 * cpw;
  }
  // Split partials are bounded (the plan kernel lengthens the VPL tile
  // when the live caches would need more).
  partial_need = std::min(partial_need, kBatchPartialFloats);
  if (partial_need > b->partial_size) {
    if (capturing) return bfail(b, DRV_ERR_STATE, "batch_light_caches: buffers grew during graph capture");
    dev_free(b->partial);
    b->partial_size = 0;
    if (hipMalloc(reinterpret_cast<void**>(&b->partial), partial_need * sizeof(float)) != hipSuccess)
      return bfail(b, DRV_ERR_NO_MEMORY, "batch_light_caches: partial sums do not fit");
    b->partial_size = partial_need;
  }
  if (std::memcmp(hv.data(), b->h_views.data(), sizeof(drvk::BatchView) * n) != 0) {
    if (capturing) return bfail(b, DRV_ERR_STATE, "batch_light_caches: view descriptors changed during graph capture");
    if (hipMemcpy(b->d_views, hv.data(), sizeof(drvk::BatchView) * n, hipMemcpyHostToDevice) != hipSuccess)
      return bfail(b, DRV_ERR_HIP, "batch_light_caches: descriptor upload failed");
    b->h_views = hv;
  }
  for (size_t v = 0; v < n; ++v) {
    drv_ctx* ctx = b->views[v];
    join_voxels_for_read(ctx);  // the batch runs on the views' shared stream
    drvk::launch_prepare_vpls(rsms[v], lights[v], ctx->vpls, b->shadow ? ctx->blocks : nullptr, 0, nullptr,
                              nullptr, b->stream);
    // The batch leaves no per-pair visibility behind.
    ctx->last_n_blocks = 0;
    ctx->label_caches = nullptr;
    ctx->counted_steps = false;
  }
  const int grid = std::max(1, b->cu_count * b->waves_per_cu / 4);
  drvk::launch_batch(b->d_views, static_cast<uint32_t>(n), b->d_queue, b->stride, b->sh_bands, b->shadow,
                     static_cast<uint32_t>(b->cpw),
                     b->vpl_tile, b->partial, b->partial_size, b->views[0]->shc, grid, b->stream);
  batch_voxels_read(b);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, DRV_ERR_HIP, std::string("batch_light_caches: ") + hipGetErrorString(e));
  return DRV_OK;
}

int drv_batch_autotune(drv_batch* b, const drv_rsm* rsms, const drv_spot_light* lights, int32_t reps,
                       int32_t* cpw_out, int32_t* tile_out, float* ms_out) {
  if (!b || !rsms || !lights || reps < 1) return DRV_ERR_INVALID_ARG;
  if (is_capturing(b->stream)) return bfail(b, DRV_ERR_STATE, "batch_autotune: not while capturing");
  const size_t n = b->views.size();
  // Snapshot of every view's entries (restored at the end).
  std::vector<float*> saved(n, nullptr);
  std::vector<size_t> bytes(n, 0);
  int rc = DRV_OK;
  auto cleanup = [&]() {
    for (float* p : saved)
      if (p) (void)hipFree(p);
  };
  for (size_t v = 0; v < n; ++v) {
    bytes[v] = sizeof(float) * b->stride * static_cast<size_t>(b->views[v]->cfg.max_caches);
    if (hipMalloc(reinterpret_cast<void**>(&saved[v]), bytes[v]) != hipSuccess) {
      cleanup();
      return bfail(b, DRV_ERR_NO_MEMORY, "batch_autotune: snapshot does not fit");
    }
    (void)hipMemcpyAsync(saved[v], b->views[v]->caches, bytes[v], hipMemcpyDeviceToDevice, b->stream);
  }
  uint32_t max_vpl = 16;
  for (size_t v = 0; v < n; ++v) {
    const uint32_t r = static_cast<uint32_t>(rsms[v].read_resolution);
    const uint32_t h = rsms[v].read_height > 0 ? static_cast<uint32_t>(rsms[v].read_height) : r;
    max_vpl = std::max(max_vpl, r * h);
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int32_t keep_cpw = b->cpw;
  const uint32_t keep_tile = b->vpl_tile;
  float best = 1e30f;
  int32_t best_cpw = keep_cpw;
  uint32_t best_tile = keep_tile;
  const int32_t cpws[6] = {0, DRV_BATCH_SPLIT_QUEUE, DRV_BATCH_SPLIT_GATHER, 64, 32, 16};
  const bool per_view_only = b->specular || any_surrogate(b);
  for (int32_t cpw : cpws) {
    if (per_view_only && cpw != 0) continue;  // env-map or surrogate views: per-view passes only
    // Per-view passes and the split queue have no VPL tile: one candidate.
    for (uint32_t tile = cpw <= 0 ? max_vpl : 256; tile <= max_vpl; tile *= 2) {
      b->cpw = cpw;
      b->vpl_tile = tile;
      std::vector<float> t;
      for (int r = 0; r <= reps && rc == DRV_OK; ++r) {  // first run warms up (buffer growth)
        (void)hipEventRecord(e0, b->stream);
        rc = drv_batch_light_caches(b, rsms, lights);
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) t.push_back(ms);
      }
      if (rc != DRV_OK) break;
      std::sort(t.begin(), t.end());
      const float med = t[t.size() / 2];
      if (med < best) {
        best = med;
        best_cpw = cpw;
        best_tile = tile;
      }
    }
    if (rc != DRV_OK) break;
  }
  b->cpw = rc == DRV_OK ? best_cpw : keep_cpw;
  b->vpl_tile = rc == DRV_OK ? best_tile : keep_tile;
  for (size_t v = 0; v < n; ++v)
    (void)hipMemcpyAsync(b->views[v]->caches, saved[v], bytes[v], hipMemcpyDeviceToDevice, b->stream);
  (void)hipStreamSynchronize(b->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cleanup();
  if (rc != DRV_OK) return rc;
  if (cpw_out) *cpw_out = b->cpw;
  if (tile_out) *tile_out = static_cast<int32_t>(b->vpl_tile);
  if (ms_out) *ms_out = best;
  return DRV_OK;
}

int drv_batch_set_surrogate(drv_batch* b, const float* host_params, const drv_surrogate* settings) {
  if (!b || !settings) return DRV_ERR_INVALID_ARG;
  for (size_t v = 0; v < b->views.size(); ++v) {
    drv_ctx* ctx = b->views[v];
    int rc;
    if (host_params && (rc = drv_mlp_set_params(ctx, host_params, static_cast<uint32_t>(drvk::kMlpParams))) != DRV_OK)
      return bfail(b, rc, "batch_set_surrogate: view " + std::to_string(v) + ": " + ctx->err);
    if ((rc = drv_set_surrogate(ctx, settings)) != DRV_OK)
      return bfail(b, rc, "batch_set_surrogate: view " + std::to_string(v) + ": " + ctx->err);
  }
  // The fused queue only cone-traces: views with the network run their own
  // passes (each with its spot check and device-gated fallback).
  if (settings->enabled) b->cpw = 0;
  return DRV_OK;
}

int drv_batch_mlp_finetune(drv_batch* b, uint32_t n_samples, uint32_t batch, uint32_t steps, float lr, uint32_t seed,
                           float* mean_loss_host) {
  if (!b || n_samples == 0 || batch == 0 || !(lr > 0.0f)) return DRV_ERR_INVALID_ARG;
  const size_t n_views = b->views.size();
  drv_ctx* owner = b->views[0];
  int rc;
  if ((rc = require_net(owner)) != DRV_OK) return bfail(b, rc, "batch_mlp_finetune: view 0: " + owner->err);
  if ((rc = ensure(owner, &b->ft_x, &b->ft_x_cap, static_cast<size_t>(n_samples) * 32)) != DRV_OK ||
      (rc = ensure(owner, &b->ft_y, &b->ft_y_cap, static_cast<size_t>(n_samples))) != DRV_OK)
    return bfail(b, rc, "batch_mlp_finetune: sample buffers do not fit");
  // Fresh cone-traced samples from every view that has labels (its spot set
  // when the network lit it, else its whole cache list), an equal share each.
  std::vector<size_t> labelled;
  for (size_t v = 0; v < n_views; ++v)
    if (b->views[v]->label_caches && b->views[v]->last_n_blocks > 0) labelled.push_back(v);
  if (labelled.empty()) return bfail(b, DRV_ERR_STATE, "batch_mlp_finetune: no view has cone-traced labels");
  uint32_t got = 0;
  for (size_t i = 0; i < labelled.size(); ++i) {
    const uint32_t share = n_samples / static_cast<uint32_t>(labelled.size()) +
                           (i < n_samples % labelled.size() ? 1u : 0u);
    if (share == 0) continue;
    drv_ctx* ctx = b->views[labelled[i]];
    if ((rc = drv_mlp_collect(ctx, share, seed * 131u + static_cast<uint32_t>(labelled[i]), b->ft_x + size_t(got) * 32,
                              b->ft_y + got)) != DRV_OK)
      return bfail(b, rc, "batch_mlp_finetune: view " + std::to_string(labelled[i]) + ": " + ctx->err);
    got += share;
  }
  if ((rc = drv_mlp_train(owner, b->ft_x, b->ft_y, got, std::min(batch, got), steps, lr, seed ^ 0x5bd1e995u,
                          mean_loss_host)) != DRV_OK)
    return bfail(b, rc, "batch_mlp_finetune: " + owner->err);
  // One network for the batch: view 0's trained parameters go to every view.
  for (size_t v = 1; v < n_views; ++v) {
    drv_ctx* ctx = b->views[v];
    if ((rc = mlp_alloc(ctx)) != DRV_OK) return bfail(b, rc, "batch_mlp_finetune: view " + std::to_string(v));
    if (hipMemcpyAsync(ctx->mlp_params, owner->mlp_params, sizeof(float) * drvk::kMlpParams, hipMemcpyDeviceToDevice,
                       b->stream) != hipSuccess)
      return bfail(b, DRV_ERR_HIP, "batch_mlp_finetune: parameter copy failed");
    drvk::launch_mlp_pack(ctx->mlp_params, ctx->mlp_pack, ctx->mlp_train_w, b->stream);
    ctx->mlp_ready = true;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, DRV_ERR_HIP, std::string("batch_mlp_finetune: ") + hipGetErrorString(e));
  return DRV_OK;
}

// ---- HIP graph -------------------------------------------------------------

int drv_graph_begin(drv_ctx* ctx) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  if (!ctx->stream) return fail(ctx, DRV_ERR_STATE, "graph_begin: the null stream cannot be captured");
  if (ctx->profiling) return fail(ctx, DRV_ERR_STATE, "graph_begin: turn profiling off before capturing");
  DRV_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  return DRV_OK;
}

int drv_graph_end(drv_ctx* ctx) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  join_voxels(ctx);  // a capture must end with every forked stream joined
  hipGraph_t g = nullptr;
  DRV_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
  if (ctx->graph_exec) {
    (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
  }
  const hipError_t e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    ctx->graph_exec = nullptr;
    return check_hip(ctx, e, "hipGraphInstantiate");
  }
  return DRV_OK;
}

int drv_graph_launch(drv_ctx* ctx) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  if (!ctx->graph_exec) return fail(ctx, DRV_ERR_STATE, "graph_launch: nothing captured");
  DRV_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return DRV_OK;
}

// ---- tuning ---------------------------------------------------------------

int drv_set_tuning(drv_ctx* ctx, const drv_tuning* t) {
  if (!ctx || !t) return DRV_ERR_INVALID_ARG;
  if ((t->visibility_kernel != 0 && t->visibility_kernel != 1) || t->blocks_per_item < 1 || t->blocks_per_item > 64 ||
      t->gather_max_splits < 1 || t->gather_max_splits > static_cast<int32_t>(drvk::kMaxSplits) || !is_pow2(t->gather_max_splits) ||
      t->waves_per_cu < 4 || t->waves_per_cu > 1024 || t->refill_min_lanes < 1 || t->refill_min_lanes > 64 ||
      ((t->steps_per_batch < 1 || t->steps_per_batch > 4) && t->steps_per_batch != 8))
    return fail(ctx, DRV_ERR_INVALID_ARG, "set_tuning: value out of range");
  ctx->tune.visibility_kernel = t->visibility_kernel;
  ctx->tune.blocks_per_item = t->blocks_per_item;
  ctx->tune.gather_max_splits = t->gather_max_splits;
  ctx->tune.waves_per_cu = t->waves_per_cu;
  ctx->tune.refill_min_lanes = t->refill_min_lanes;
  ctx->tune.steps_per_batch = t->steps_per_batch;
  return DRV_OK;
}

int drv_get_tuning(const drv_ctx* ctx, drv_tuning* out) {
  if (!ctx || !out) return DRV_ERR_INVALID_ARG;
  out->visibility_kernel = ctx->tune.visibility_kernel;
  out->blocks_per_item = ctx->tune.blocks_per_item;
  out->gather_max_splits = ctx->tune.gather_max_splits;
  out->waves_per_cu = ctx->tune.waves_per_cu;
  out->refill_min_lanes = ctx->tune.refill_min_lanes;
  out->steps_per_batch = ctx->tune.steps_per_batch;
  return DRV_OK;
}

// ---- profiling ------------------------------------------------------------

int drv_set_profiling(drv_ctx* ctx, int32_t enabled) {
  if (!ctx) return DRV_ERR_INVALID_ARG;
  ctx->profiling = enabled != 0;
  ctx->profiling_level = enabled;
  return DRV_OK;
}

int drv_get_stage_times(drv_ctx* ctx, drv_stage_times* out) {
  if (!ctx || !out) return DRV_ERR_INVALID_ARG;
  if (ctx->profiling) {
    // Events of the last call of each stage (elapsed() yields 0 for a pair
    // that was never recorded).
    DRV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->times.voxelize_ms = elapsed(ctx, 0, 1);
    ctx->times.voxel_mips_ms = elapsed(ctx, 1, 2);
    ctx->times.allocate_ms = elapsed(ctx, 3, 4);
    ctx->times.light_caches_ms = elapsed(ctx, 5, ctx->cfg.indirect_specular ? 12 : 9);
    ctx->times.specular_ms = ctx->cfg.indirect_specular ? elapsed(ctx, 9, 12) : 0.0f;
    ctx->times.prepare_ms = elapsed(ctx, 5, 6);
    ctx->times.visibility_ms = elapsed(ctx, 6, 7);
    ctx->times.gather_ms = elapsed(ctx, 7, 8);
    ctx->times.reduce_ms = elapsed(ctx, 8, 9);
    ctx->times.apply_ms = elapsed(ctx, 10, 11);
    if (ctx->counted_steps && ctx->step_counter) {
      uint64_t st[3] = {0, 0, 0};
      DRV_HIP(ctx, hipMemcpy(st, ctx->step_counter, sizeof(st), hipMemcpyDeviceToHost));
      ctx->times.cone_traces = st[0];
      ctx->times.cone_steps = st[1];
      ctx->times.cone_steps_executed = st[2];
    }
  }
  *out = ctx->times;
  return DRV_OK;
}

}  // extern "C"
