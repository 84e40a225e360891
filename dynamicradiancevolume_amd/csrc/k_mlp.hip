// Surrogate visibility: a 32-64-64-64-1 MLP (ReLU, sigmoid output) that
// stands in for the per-(cache, shadow block) cone trace of
// shader/cacheLightingRSM.comp:167-231.  BF16 operands, FP32 accumulation, on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layout.  Every layer is computed transposed, Z^T = W^T X^T: the MFMA's M
// dimension is the 16 output units of a tile, N the 16 samples of a chain,
// K the inputs.  An MFMA result leaves lane (j, g) = (lane % 16, lane / 16)
// holding units 16t + 4g + r (r = 0..3) of sample j for each unit tile t, and
// that lane is also the one that supplies sample j's inputs to the next
// layer.  So the next layer's K slots of lane group g are filled from the
// lane's own registers: slot (kb, g, e) <- unit 16(2kb + e/4) + 4g + e%4, and
// the packed weights use the same slot -> unit map.  No LDS is needed between
// layers.  The slot -> k assignment of the hardware is irrelevant: A and B
// use the same one.
//
// Training (k_mlp_train) stages operands in LDS and runs the forward pass,
// the backward pass and the weight-gradient GEMMs (K = samples) with the same
// MFMA; every wave writes its partial gradient and k_mlp_adam reduces them in
// a fixed order and applies Adam to the FP32 master weights.
#include <algorithm>

#include "drv_internal.hpp"
#include "k_common.hpp"
#include "k_sample.hpp"

namespace drvk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ __bf16 to_bf16(float v) { return static_cast<__bf16>(v); }
__device__ __forceinline__ float from_bf16(__bf16 v) { return static_cast<float>(v); }

// ---- features ---------------------------------------------------------------

// The 32 features of a (cache, block) sample in slot order: slot 8g + e holds
// geometry feature G[5g + e] for e < 5 and voxel sample S[3g + e - 5] else.
//   G: 0-2 cache position in the voxel volume [0,1]^3; 3-5 unit direction to
//      the camera (caches carry no normal); 6-8 averaged VAL position in the
//      volume; 9-11 mean VPL normal of the block; 12-14 unit direction to the
//      VAL; 15 distance / 64 voxels; 16 cone aperture * 16; 17 march goal / 64
//      voxels; 18 cos(VPL normal, -direction); 19 cos(camera dir, direction).
//   S: 0-7 occupancy at (i + 0.5) / 8 of the march segment, at the level of
//      the cone radius there; 8-11 occupancy at (i + 0.5) / 4 on mip level 3.
struct FeatureCtx {
  float vmin[3];
  float voxel_size;
  float cam[3];
  VoxelView vox;
};

// The features that depend on the cache alone (computed once per cache, not
// once per pair): position in the volume and unit direction to the camera.
struct CacheFeatures {
  f3 wp;
  float vpx, vpy, vpz;
  float cx, cy, cz;
};

__device__ __forceinline__ CacheFeatures cache_features(const FeatureCtx& fc, f3 wp) {
  const float res = static_cast<float>(fc.vox.res);
  const float vsz = fc.voxel_size * res;
  CacheFeatures k;
  k.wp = wp;
  k.vpx = (wp.x - fc.vmin[0]) / vsz;
  k.vpy = (wp.y - fc.vmin[1]) / vsz;
  k.vpz = (wp.z - fc.vmin[2]) / vsz;
  float cx = fc.cam[0] - wp.x, cy = fc.cam[1] - wp.y, cz = fc.cam[2] - wp.z;
  const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
  const float ci = cl > 0.0f ? 1.0f / cl : 0.0f;
  k.cx = cx * ci;
  k.cy = cy * ci;
  k.cz = cz * ci;
  return k;
}

__device__ __forceinline__ void features_pair(const FeatureCtx& fc, const LevelInfo* lvi, const CacheFeatures& k,
                                              const ShadowBlock& b, float* out) {
  const float res = static_cast<float>(fc.vox.res);
  const float vsz = fc.voxel_size * res;
  const f3 wp = k.wp;
  const float vpx = k.vpx, vpy = k.vpy, vpz = k.vpz;
  const float cx = k.cx, cy = k.cy, cz = k.cz;
  float tx = b.x - wp.x, ty = b.y - wp.y, tz = b.z - wp.z;
  const float ld = sqrtf(tx * tx + ty * ty + tz * tz);
  const float li = ld > 0.0f ? 1.0f / ld : 0.0f;
  tx *= li;
  ty *= li;
  tz *= li;
  const float dist = ld / fc.voxel_size;
  const float goal = dist - 2.0f;
  const float G[20] = {vpx, vpy, vpz, cx, cy, cz,
                       (b.x - fc.vmin[0]) / vsz, (b.y - fc.vmin[1]) / vsz, (b.z - fc.vmin[2]) / vsz,
                       b.nx, b.ny, b.nz, tx, ty, tz,
                       dist * (1.0f / 64.0f), b.rho * 16.0f, goal * (1.0f / 64.0f),
                       -(b.nx * tx + b.ny * ty + b.nz * tz), cx * tx + cy * ty + cz * tz};
  const int q = fc.vox.levels - 1;
  const float span = fmaxf(goal, 0.0f);
  const float p0x = vpx * res, p0y = vpy * res, p0z = vpz * res;
  // All 12 fetch addresses first, then the loads, then the blends.
  Footprint fp[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    float d;
    int lvl;
    if (k < 8) {
      d = 2.0f + span * ((static_cast<float>(k) + 0.5f) * 0.125f);
      const float lod = __builtin_amdgcn_logf(fmaxf(d * b.rho, 1.0f));
      lvl = min(static_cast<int>(lod), q);
    } else {
      d = 2.0f + span * ((static_cast<float>(k - 8) + 0.5f) * 0.25f);
      lvl = min(3, q);
    }
    fp[k] = footprint(lvi, lvl, p0x + tx * d, p0y + ty * d, p0z + tz * d);
  }
  // Two rounds of six 16-byte corner loads in flight (twelve would take 48
  // registers at once).
  float occ[12];
  const __amdgpu_buffer_rsrc_t cr = corner_rsrc(fc.vox);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    corner_t c[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = corner_load(cr, fp[6 * h + k].offset);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const Footprint& f = fp[6 * h + k];
      occ[6 * h + k] = blend_corner(c[k], f.ax, f.ay, f.az) * (1.0f / 255.0f);
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
#pragma unroll
    for (int e = 0; e < 5; ++e) out[8 * g + e] = G[5 * g + e];
#pragma unroll
    for (int e = 0; e < 3; ++e) out[8 * g + 5 + e] = occ[3 * g + e];
  }
}

__device__ __forceinline__ void features_all(const FeatureCtx& fc, const LevelInfo* lvi, f3 wp, const ShadowBlock& b,
                                             float* out) {
  features_pair(fc, lvi, cache_features(fc, wp), b, out);
}

// ---- inference with the network in LDS --------------------------------------

// The network as a workgroup holds it in LDS: the packed A operands in the
// layout of the global pack (layer 1 [t][lane], layers 2 and 3 [t][kb][lane],
// one bf16x8 per lane, so a wave's operand read is one 1 KiB row without bank
// conflicts), then the biases and the output weights.  Keeping the 80
// operand registers out of the VGPR file lets a SIMD hold several waves, whose
// voxel taps then overlap each other's matrix work.
struct NetLds {
  bf16x8 a[1280];
  float b1[64], b2[64], b3[64], w4[64];
  float b4;
};

// Whole workgroup: copies the network into LDS (the caller synchronizes).
__device__ __forceinline__ void stage_net(const bf16x8* pack, const float* params, NetLds& n) {
  for (uint32_t i = threadIdx.x; i < 1280u; i += blockDim.x) n.a[i] = pack[i];
  for (uint32_t i = threadIdx.x; i < 64u; i += blockDim.x) {
    n.b1[i] = params[kMlpB1 + i];
    n.b2[i] = params[kMlpB2 + i];
    n.b3[i] = params[kMlpB3 + i];
    n.w4[i] = params[kMlpW4 + i];
  }
  if (threadIdx.x == 0) n.b4 = params[kMlpB4];
}

// Bias (or w4) values of the units a lane's D registers hold: 16t + 4g + r.
__device__ __forceinline__ f32x4 unit4(const float* v, int t, int g) {
  const float4 q = *reinterpret_cast<const float4*>(v + 16 * t + 4 * g);
  return f32x4{q.x, q.y, q.z, q.w};
}

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// relu of two values, rounded to bf16 and packed (a in the low half).  The
// ReLU acts on the rounded values as 16-bit integers: a negative bf16 (sign
// bit set) is a negative integer and goes to +0, a non-negative one is kept,
// which is relu-then-round for every input (one packed integer max instead of
// two quieting float maxes per value).
__device__ __forceinline__ uint32_t relu2(float a, float b) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const bf16x2 v = __builtin_convertvector((f32x2{a, b}), bf16x2);
  const s16x2 r = __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), s16x2{0, 0});
  return __builtin_bit_cast(uint32_t, r);
}

// relu(acc) of tiles (2kb, 2kb + 1) -> bf16 slots (kb, e) of the next layer.
__device__ __forceinline__ bf16x8 relu_pack(f32x4 lo, f32x4 hi) {
  const u32x4 w = {relu2(lo[0], lo[1]), relu2(lo[2], lo[3]), relu2(hi[0], hi[1]), relu2(hi[2], hi[3])};
  return __builtin_bit_cast(bf16x8, w);
}

// The two bf16 halves of a packed word as floats.
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

constexpr int kChains = 4;  // 16-sample chains per forward call (64 samples)

// Forward pass of four 16-sample chains; chain c's 8 feature slots per lane
// are in x[c].  Each weight operand is read from LDS once and used by all
// four chains.  Returns in p[c] the visibility of sample 16c + lane % 16, in
// every lane group.
__device__ __forceinline__ void mlp_forward64(const NetLds& n, const bf16x8 (&x)[kChains], int lane, int g,
                                              float (&p)[kChains]) {
  bf16x8 x2[kChains][2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    f32x4 h[2][kChains];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = 2 * kb + u;
      const bf16x8 w = n.a[t * 64 + lane];
      const f32x4 bias = unit4(n.b1, t, g);
#pragma unroll
      for (int c = 0; c < kChains; ++c) h[u][c] = mfma(w, x[c], bias);
    }
#pragma unroll
    for (int c = 0; c < kChains; ++c) x2[c][kb] = relu_pack(h[0][c], h[1][c]);
  }
  bf16x8 x3[kChains][2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    f32x4 h[2][kChains];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = 2 * kb + u;
      const f32x4 bias = unit4(n.b2, t, g);
#pragma unroll
      for (int c = 0; c < kChains; ++c) h[u][c] = bias;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const bf16x8 w = n.a[256 + (t * 2 + k) * 64 + lane];
#pragma unroll
        for (int c = 0; c < kChains; ++c) h[u][c] = mfma(w, x2[c][k], h[u][c]);
      }
    }
#pragma unroll
    for (int c = 0; c < kChains; ++c) x3[c][kb] = relu_pack(h[0][c], h[1][c]);
  }
  // Output unit: h3 rounded to bf16 like every hidden activation, dot with
  // the FP32 w4, summed over the four lane groups of the sample.
  float z[kChains] = {};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const f32x4 bias = unit4(n.b3, t, g);
    f32x4 h[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) h[c] = bias;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const bf16x8 w = n.a[768 + (t * 2 + k) * 64 + lane];
#pragma unroll
      for (int c = 0; c < kChains; ++c) h[c] = mfma(w, x3[c][k], h[c]);
    }
    const f32x4 wv = unit4(n.w4, t, g);
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      const uint32_t w01 = relu2(h[c][0], h[c][1]), w23 = relu2(h[c][2], h[c][3]);
      z[c] = __builtin_fmaf(wv[0], bf16_lo(w01), z[c]);
      z[c] = __builtin_fmaf(wv[1], bf16_hi(w01), z[c]);
      z[c] = __builtin_fmaf(wv[2], bf16_lo(w23), z[c]);
      z[c] = __builtin_fmaf(wv[3], bf16_hi(w23), z[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    float s = z[c];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    s += n.b4;
    p[c] = 1.0f / (1.0f + __expf(-s));
  }
}

struct InferArgs {
  FeatureCtx fc;
  const bf16x8* pack;
  const float* params;
  const float* caches;
  int stride;
  const Counter* counter;
  const Counter* first;  // optional: the list's caches start at entry first->live
  const ShadowBlock* blocks;
  uint32_t n_blocks;
  float* vis;
  uint32_t pitch;
};

constexpr uint32_t kMlpBlocksPerItem = 8;
// Staged feature rows: 32 bf16 (four 16-byte chunks) per row, chunk q of
// row r stored at chunk position q ^ ((r >> 2) & 3), so that the 16 rows a
// chain's lanes read (and the 16 lanes of a row-store) fall on distinct banks.
__device__ __forceinline__ int stage_chunk(int row, int q) { return row * 32 + 8 * (q ^ ((row >> 2) & 3)); }
constexpr int kInferWaves = 4;  // waves per inference workgroup

// Surrogate visibility for every (live cache, shadow block): waves over
// (64-cache tile, 8-block group) items, one workgroup's waves sharing one LDS
// copy of the network.  Per block, each lane builds the features of its own
// cache (the block record is wave-uniform) and stages them in LDS as bf16;
// the four 16-sample MFMA chains then evaluate the network together, and
// chain c's result for sample 16c + j is taken by lane (j, g = c), the lane
// that owns that cache, so the store is one coalesced row.
__global__ __launch_bounds__(64 * kInferWaves) void k_visibility_mlp(InferArgs a) {
  __shared__ NetLds net;
  __shared__ __align__(16) __bf16 stage[kInferWaves][64 * 32];
  __shared__ LevelInfo lvi[16];
  const uint32_t live = a.counter->live;
  const uint32_t first = a.first ? a.first->live : 0u;
  const uint32_t tiles = (live + 63u) / 64u;
  const uint32_t n_waves = gridDim.x * kInferWaves;
  // Blocks per item: the largest of 8, 4, 2 that still gives every wave
  // two items or more (the cache record and its cache-only features are
  // loaded once per item); a few caches (a spot-check set) get one.
  uint32_t bpi = kMlpBlocksPerItem;
  while (bpi > 1u && tiles * ((a.n_blocks + bpi - 1) / bpi) < 2u * n_waves) bpi >>= 1;
  const uint32_t groups = (a.n_blocks + bpi - 1) / bpi;
  const uint32_t items = tiles * groups;
  if (blockIdx.x * kInferWaves >= items) return;  // workgroup-uniform: no wave of it has an item
  stage_levels(a.fc.vox, lvi);
  stage_net(a.pack, a.params, net);
  __syncthreads();
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const int j = lane & 15, g = lane >> 4;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kInferWaves + (threadIdx.x >> 6));
  if (wave >= items) return;
  __bf16* st = stage[threadIdx.x >> 6];
  for (uint32_t item = wave; item < items; item += n_waves) {
    const uint32_t tile = item / groups, group = item % groups;
    const uint32_t cache = tile * 64u + static_cast<uint32_t>(lane);
    const bool valid = cache < live;
    const float* e = a.caches + static_cast<size_t>(first + (valid ? cache : 0u)) * a.stride;
    const CacheFeatures cf = cache_features(a.fc, make3(e[0], e[1], e[2]));
    const uint32_t b0 = group * bpi;
    const uint32_t b1 = min(a.n_blocks, b0 + bpi);
    for (uint32_t b = b0; b < b1; ++b) {
      float f[32];
      features_pair(a.fc, lvi, cf, a.blocks[b], f);
      bf16x8 row[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 8; ++k) row[q][k] = to_bf16(f[8 * q + k]);
      // the wave's own rows: all lanes store, then read other lanes' rows
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 4; ++q) *reinterpret_cast<bf16x8*>(st + stage_chunk(lane, q)) = row[q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bf16x8 x[kChains];
#pragma unroll
      for (int c = 0; c < kChains; ++c) x[c] = *reinterpret_cast<const bf16x8*>(st + stage_chunk(16 * c + j, g));
      float p[kChains];
      mlp_forward64(net, x, lane, g, p);
      const float mine = g == 0 ? p[0] : g == 1 ? p[1] : g == 2 ? p[2] : p[3];
      if (valid) a.vis[static_cast<size_t>(b) * a.pitch + first + cache] = mine;
    }
  }
}

// Features of explicit (cache, block) pairs, X[n][32] in slot order.
struct FeatArgs {
  FeatureCtx fc;
  const float* caches;
  int stride;
  const ShadowBlock* blocks;
  const uint32_t* pairs;  // [n][2] = (cache, block)
  uint32_t n;
  uint32_t n_caches;  // valid cache indices (entries in `caches`)
  uint32_t n_blocks;
  float* x;
  const float* vis;  // optional labels source (block-major, pitch)
  uint32_t pitch;
  float* y;
};

__global__ __launch_bounds__(256) void k_mlp_features(FeatArgs a) {
  __shared__ LevelInfo lvi[16];
  stage_levels(a.fc.vox, lvi);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t cache = a.pairs[2 * i], b = a.pairs[2 * i + 1];
  if (cache >= a.n_caches || b >= a.n_blocks) {  // out-of-range pair: NaN features, no label
    for (int k = 0; k < 32; ++k) a.x[static_cast<size_t>(i) * 32 + k] = __builtin_nanf("");
    if (a.y) a.y[i] = __builtin_nanf("");
    return;
  }
  const float* e = a.caches + static_cast<size_t>(cache) * a.stride;
  const f3 wp = make3(e[0], e[1], e[2]);
  float f[32];
  features_all(a.fc, lvi, wp, a.blocks[b], f);
  float4* xo = reinterpret_cast<float4*>(a.x + static_cast<size_t>(i) * 32);
#pragma unroll
  for (int k = 0; k < 8; ++k) xo[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
  if (a.y) a.y[i] = a.vis[static_cast<size_t>(b) * a.pitch + cache];
}

// Random (cache, block) pairs from a counter-based hash: pair i draws its
// cache and block independently (uniform over live caches x blocks).
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void k_mlp_sample_pairs(const Counter* counter, uint32_t n_blocks, uint32_t n, uint32_t seed,
                                   uint32_t* pairs, uint32_t* n_out) {
  const uint32_t live = counter->live;
  const uint32_t m = (live == 0 || n_blocks == 0) ? 0u : n;
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = m;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t h1 = hash32(i * 2u + 0x9e3779b9u * seed + 1u);
    const uint32_t h2 = hash32(i * 2u + 1u + 0x85ebca6bu * seed);
    pairs[2 * i] = live ? h1 % live : 0u;
    pairs[2 * i + 1] = n_blocks ? h2 % n_blocks : 0u;
  }
}

// Batch inference on given features (the same LDS-resident forward pass,
// 64 samples per wave pass).
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_forward(const bf16x8* pack, const float* params,
                                                                  const float* x, uint32_t n, float* p_out) {
  __shared__ NetLds net;
  const uint32_t n_waves = gridDim.x * kInferWaves;
  const uint32_t groups = (n + 63u) / 64u;
  if (blockIdx.x * kInferWaves >= groups) return;  // workgroup-uniform
  stage_net(pack, params, net);
  __syncthreads();
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const int j = lane & 15, g = lane >> 4;
  const uint32_t wave = blockIdx.x * kInferWaves + (threadIdx.x >> 6);
  for (uint32_t q = wave; q < groups; q += n_waves) {
    bf16x8 xv[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      const uint32_t s = q * 64u + 16u * c + static_cast<uint32_t>(j);
      const bool valid = s < n;
#pragma unroll
      for (int k = 0; k < 8; ++k) xv[c][k] = to_bf16(valid ? x[static_cast<size_t>(s) * 32 + 8 * g + k] : 0.0f);
    }
    float p[kChains];
    mlp_forward64(net, xv, lane, g, p);
    const uint32_t s = q * 64u + 16u * g + static_cast<uint32_t>(j);
    const float mine = g == 0 ? p[0] : g == 1 ? p[1] : g == 2 ? p[2] : p[3];
    if (s < n) p_out[s] = mine;
  }
}

// ---- packing ----------------------------------------------------------------

// Hidden unit held in slot (kb, g, e) of a lane: 16(2kb + e/4) + 4g + e%4.
__device__ __forceinline__ int slot_unit(int kb, int g, int e) { return 16 * (2 * kb + e / 4) + 4 * g + (e & 3); }

// From the FP32 master parameters: (1) the inference A operands, (2) the
// training LDS images W1T [64][40], W2T/W3T [64][72] ([out][in], padded), and
// W2/W3 [64][72] ([in][out]).
__global__ void k_mlp_pack(const float* params, bf16x8* pack, __bf16* train_w) {
  const float* W1 = params + kMlpW1;
  const float* W2 = params + kMlpW2;
  const float* W3 = params + kMlpW3;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < 1280) {
    const int lane = tid & 63;
    const int row = lane & 15, g = lane >> 4;
    bf16x8 v;
    if (tid < 256) {
      const int t = tid >> 6;
      for (int e = 0; e < 8; ++e) v[e] = to_bf16(W1[(8 * g + e) * 64 + 16 * t + row]);
    } else {
      const int q = (tid - 256) >> 6;  // 0..15: layer 2 (0..7) or 3 (8..15)
      const int layer = q >> 3, t = (q & 7) >> 1, kb = q & 1;
      const float* W = layer == 0 ? W2 : W3;
      for (int e = 0; e < 8; ++e) v[e] = to_bf16(W[slot_unit(kb, g, e) * 64 + 16 * t + row]);
    }
    pack[tid] = v;
  }
  // Training images.
  for (int i = tid; i < kMlpTrainW; i += gridDim.x * blockDim.x) {
    float v = 0.0f;
    if (i < 64 * 40) {
      const int o = i / 40, c = i % 40;
      v = c < 32 ? W1[c * 64 + o] : 0.0f;
    } else {
      const int r = i - 64 * 40;
      const int m = r / (64 * 72);  // 0: W2T, 1: W3T, 2: W2, 3: W3
      const int rr = r % (64 * 72);
      const int a = rr / 72, c = rr % 72;
      if (c < 64) {
        const float* W = (m == 0 || m == 2) ? W2 : W3;
        v = (m < 2) ? W[c * 64 + a] : W[a * 64 + c];
      }
    }
    train_w[i] = to_bf16(v);
  }
}

// ---- training -----------------------------------------------------------------

constexpr int kTrainChain = 32;  // samples per wave per chain
constexpr int kTrainWaves = 2;   // waves per workgroup
constexpr int LD32 = 40, LD64 = 72;

struct TrainLds {
  __bf16 w1t[64 * LD32];
  __bf16 w2t[64 * LD64];
  __bf16 w3t[64 * LD64];
  __bf16 w2[64 * LD64];
  __bf16 w3[64 * LD64];
  struct Wave {
    __bf16 x[kTrainChain * LD32];    // [s][f]
    __bf16 xt[32 * LD32];            // [f][s]
    __bf16 h[3][kTrainChain * LD64];   // [s][u]
    __bf16 ht[3][64 * LD32];           // [u][s]
    __bf16 dz[kTrainChain * LD64];   // [s][u]
    __bf16 dzt[64 * LD32];           // [u][s]
    float y[kTrainChain];
    float wgt[kTrainChain];
    float dz4[kTrainChain];
  } wv[kTrainWaves];
};

// acc[mt][nt] += A[16mt.., :] . B[:, 16nt..] with A row-major [M][K] (lda)
// and B given as B^T row-major [N][K] (ldbt), K = 32 * KB.
template <int MT, int NT, int KB>
__device__ __forceinline__ void wave_gemm(f32x4 (&acc)[MT][NT], const __bf16* A, int lda, const __bf16* Bt, int ldbt,
                                          int lane) {
  const int row = lane & 15, kg = (lane >> 4) * 8;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    bf16x8 a[MT], b[NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) a[mt] = *reinterpret_cast<const bf16x8*>(A + (16 * mt + row) * lda + 32 * kb + kg);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      b[nt] = *reinterpret_cast<const bf16x8*>(Bt + (16 * nt + row) * ldbt + 32 * kb + kg);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(a[mt], b[nt], acc[mt][nt]);
  }
}

template <int MT, int NT>
__device__ __forceinline__ void zero(f32x4 (&acc)[MT][NT]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
}

// Writes relu(acc + bias) of a [64 units] x [32 samples] D tile set as
// bf16 [s][u] and [u][s].
__device__ __forceinline__ void store_hidden(const f32x4 (&acc)[4][2], const float* bias, __bf16* h, __bf16* ht,
                                             int lane) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const int s = 16 * nt + j;
      bf16x4 v;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int u = 16 * mt + 4 * g + r;
        const __bf16 q = to_bf16(fmaxf(acc[mt][nt][r] + bias[u], 0.0f));
        v[r] = q;
        ht[u * LD32 + s] = q;
      }
      *reinterpret_cast<bf16x4*>(h + s * LD64 + 16 * mt + 4 * g) = v;
    }
}

struct TrainArgs {
  const float* x;        // [n][32]
  const float* y;        // [n]
  const uint32_t* idx;   // minibatch sample indices [n_batch]
  uint32_t n_batch;
  const float* params;   // FP32 master
  const __bf16* train_w; // packed LDS images
  float* partial;        // [n_waves][kMlpParams + 1] (last: loss)
  float inv_batch;
  uint32_t chains_per_wave;
};

__global__ __launch_bounds__(128) void k_mlp_train(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  TrainLds& L = *reinterpret_cast<TrainLds*>(smem);
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const int wid = static_cast<int>(threadIdx.x >> 6);
  const int j = lane & 15, g = lane >> 4;
  const uint32_t gwave = blockIdx.x * kTrainWaves + wid;
  const uint32_t n_waves = gridDim.x * kTrainWaves;
  // weights -> LDS
  {
    const uint4* src = reinterpret_cast<const uint4*>(a.train_w);
    uint4* dst = reinterpret_cast<uint4*>(L.w1t);
    for (int i = threadIdx.x; i < kMlpTrainW / 8; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  auto& W = L.wv[wid];
  const float* b1 = a.params + kMlpB1;
  const float* b2 = a.params + kMlpB2;
  const float* b3 = a.params + kMlpB3;
  const float* w4 = a.params + kMlpW4;
  const float b4 = a.params[kMlpB4];
  f32x4 gw1[4][2], gw2[4][4], gw3[4][4];
  zero(gw1);
  zero(gw2);
  zero(gw3);
  float gb1[4][4], gb2[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) gb1[mt][r] = gb2[mt][r] = 0.0f;
  float gb3 = 0.0f, gw4 = 0.0f, gb4 = 0.0f, loss = 0.0f;

  for (uint32_t c = 0; c < a.chains_per_wave; ++c) {
    const uint32_t chain = gwave + c * n_waves;
    // 1. samples -> LDS (lanes 0..31: one sample each)
    if (lane < kTrainChain) {
      const uint32_t bi = chain * kTrainChain + lane;
      const bool ok = bi < a.n_batch;
      const uint32_t si = ok ? a.idx[bi] : 0u;
      for (int f = 0; f < 32; ++f) {
        const __bf16 q = to_bf16(ok ? a.x[static_cast<size_t>(si) * 32 + f] : 0.0f);
        W.x[lane * LD32 + f] = q;
        W.xt[f * LD32 + lane] = q;
      }
      W.y[lane] = ok ? a.y[si] : 0.0f;
      W.wgt[lane] = ok ? 1.0f : 0.0f;
    }
    __syncthreads();
    // 2-4. forward
    {
      f32x4 acc[4][2];
      zero(acc);
      wave_gemm<4, 2, 1>(acc, L.w1t, LD32, W.x, LD32, lane);
      store_hidden(acc, b1, W.h[0], W.ht[0], lane);
    }
    __syncthreads();
    {
      f32x4 acc[4][2];
      zero(acc);
      wave_gemm<4, 2, 2>(acc, L.w2t, LD64, W.h[0], LD64, lane);
      store_hidden(acc, b2, W.h[1], W.ht[1], lane);
    }
    __syncthreads();
    {
      f32x4 acc[4][2];
      zero(acc);
      wave_gemm<4, 2, 2>(acc, L.w3t, LD64, W.h[1], LD64, lane);
      store_hidden(acc, b3, W.h[2], W.ht[2], lane);
    }
    __syncthreads();
    // 5. output unit, loss, dz4: lane = (sample s = lane % 32, half = lane / 32)
    {
      const int s = lane & 31, half = lane >> 5;
      float z = 0.0f;
      for (int u = 32 * half; u < 32 * half + 32; ++u) z = __builtin_fmaf(w4[u], from_bf16(W.h[2][s * LD64 + u]), z);
      z += __shfl_xor(z, 32);
      z += b4;
      const float p = 1.0f / (1.0f + __expf(-z));
      if (half == 0) {
        const float yv = W.y[s], wv = W.wgt[s];
        const float pc = fminf(fmaxf(p, 1e-7f), 1.0f - 1e-7f);
        loss += -wv * (yv * __logf(pc) + (1.0f - yv) * __logf(1.0f - pc));
        W.dz4[s] = wv * (p - yv) * a.inv_batch;
      }
    }
    __syncthreads();
    // 6. lane u: dW4, db3 and dZ3 = (w4 dz4) * (h3 > 0)
    {
      const int u = lane;
      float acc4 = 0.0f, accb = 0.0f;
      for (int s = 0; s < kTrainChain; ++s) {
        const float d4 = W.dz4[s];
        const float h = from_bf16(W.ht[2][u * LD32 + s]);
        acc4 = __builtin_fmaf(d4, h, acc4);
        const float dz = h > 0.0f ? w4[u] * d4 : 0.0f;
        accb += dz;
        const __bf16 q = to_bf16(dz);
        W.dzt[u * LD32 + s] = q;
        W.dz[s * LD64 + u] = q;
      }
      gw4 += acc4;
      gb3 += accb;
      if (lane == 0)
        for (int s = 0; s < kTrainChain; ++s) gb4 += W.dz4[s];
    }
    __syncthreads();
    // 7. dW3^T += dZ3^T . H2
    wave_gemm<4, 4, 1>(gw3, W.dzt, LD32, W.ht[1], LD32, lane);
    // 8. dH2^T = W3 . dZ3^T  ([in][out] x [out][s])
    f32x4 dh[4][2];
    zero(dh);
    wave_gemm<4, 2, 2>(dh, L.w3, LD64, W.dz, LD64, lane);
    __syncthreads();
    // 9. dZ2 = dH2 * (h2 > 0), db2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const int s = 16 * nt + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          const float dz = from_bf16(W.h[1][s * LD64 + u]) > 0.0f ? dh[mt][nt][r] : 0.0f;
          gb2[mt][r] += dz;
          const __bf16 q = to_bf16(dz);
          W.dzt[u * LD32 + s] = q;
          W.dz[s * LD64 + u] = q;
        }
      }
    __syncthreads();
    // 10. dW2^T += dZ2^T . H1 ; 11. dH1^T = W2 . dZ2^T
    wave_gemm<4, 4, 1>(gw2, W.dzt, LD32, W.ht[0], LD32, lane);
    zero(dh);
    wave_gemm<4, 2, 2>(dh, L.w2, LD64, W.dz, LD64, lane);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const int s = 16 * nt + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          const float dz = from_bf16(W.h[0][s * LD64 + u]) > 0.0f ? dh[mt][nt][r] : 0.0f;
          gb1[mt][r] += dz;
          W.dzt[u * LD32 + s] = to_bf16(dz);
        }
      }
    __syncthreads();
    // 12. dW1^T += dZ1^T . X
    wave_gemm<4, 2, 1>(gw1, W.dzt, LD32, W.xt, LD32, lane);
    __syncthreads();
  }

  // Per-wave partial gradient in the parameter layout.
  float* out = a.partial + static_cast<size_t>(gwave) * (kMlpParams + 1);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = 16 * mt + 4 * g + r;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) out[kMlpW1 + (16 * nt + j) * 64 + o] = gw1[mt][nt][r];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        out[kMlpW2 + (16 * nt + j) * 64 + o] = gw2[mt][nt][r];
        out[kMlpW3 + (16 * nt + j) * 64 + o] = gw3[mt][nt][r];
      }
      // bias sums over the 16 sample lanes of this lane group
      float v1 = gb1[mt][r], v2 = gb2[mt][r];
      for (int off = 1; off < 16; off <<= 1) {
        v1 += __shfl_xor(v1, off);
        v2 += __shfl_xor(v2, off);
      }
      if (j == 0) {
        out[kMlpB1 + o] = v1;
        out[kMlpB2 + o] = v2;
      }
    }
  out[kMlpB3 + lane] = gb3;
  out[kMlpW4 + lane] = gw4;
  float lsum = loss;
  for (int off = 1; off < 64; off <<= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0) {
    out[kMlpB4] = gb4;
    out[kMlpParams] = lsum;
  }
}

// Fixed-order reduction of the per-wave partials, then Adam on the FP32
// master weights (step t >= 1).  With lr = 0 only the gradient is produced.
__global__ void k_mlp_adam(const float* partial, uint32_t n_waves, float* params, float* m, float* v, float* grad_out,
                           float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > static_cast<uint32_t>(kMlpParams)) return;
  float gsum = 0.0f;
  for (uint32_t w = 0; w < n_waves; ++w) gsum += partial[static_cast<size_t>(w) * (kMlpParams + 1) + i];
  if (grad_out) grad_out[i] = gsum;
  if (i == static_cast<uint32_t>(kMlpParams) || lr == 0.0f) return;
  const float mi = beta1 * m[i] + (1.0f - beta1) * gsum;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gsum * gsum;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Minibatch indices: a seeded random sample (with replacement) of 0..n-1.
__global__ void k_mlp_batch_indices(uint32_t n, uint32_t batch, uint32_t seed, uint32_t* idx) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < batch; i += gridDim.x * blockDim.x)
    idx[i] = hash32(i * 0x9e3779b9u + seed * 0x85ebca6bu + 7u) % n;
}

__global__ void k_iota(uint32_t n, uint32_t* idx) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) idx[i] = i;
}

// ---- spot check ---------------------------------------------------------------

// Every k-th live cache -> two compact copies (position kept, SH zeroed): set
// 0 is lit with cone-traced visibility, set 1 with the network's.  Counters
// [0] and [1] describe the two sets; counter [2] (the fallback gate) is
// written by k_spot_error.
__global__ void k_spot_select(const Counter* counter, uint32_t k, const float* caches, int stride, float* spot,
                              uint32_t spot_cap, Counter* spot_ctr) {
  const uint32_t live = counter->live;
  const uint32_t n_spot = min(spot_cap, (live + k - 1u) / k);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    // [0], [1]: set a (traced) and set b (network), n_spot each; [3]: both
    // sets as one list of 2 n_spot (set b follows set a); [2] is the gate.
    for (int c = 0; c < 4; ++c) {
      if (c == 2) continue;
      const uint32_t n = c == 3 ? 2u * n_spot : n_spot;
      spot_ctr[c].groups_x = (n + 63u) / 64u;
      spot_ctr[c].groups_y = spot_ctr[c].groups_z = 1u;
      spot_ctr[c].total = n;
      spot_ctr[c].live = n;
      spot_ctr[c].tiles = (n + kGatherTile - 1) / kGatherTile;
      spot_ctr[c].splits = 1u;
      spot_ctr[c].phantoms = 0u;
    }
  }
  if (i >= n_spot) return;
  const float* src = caches + static_cast<size_t>(i) * k * stride;
  float* a = spot + static_cast<size_t>(i) * stride;
  float* b = spot + (static_cast<size_t>(n_spot) + i) * stride;
  for (int f = 0; f < stride; ++f) {
    const float v = f < 3 ? src[f] : 0.0f;
    a[f] = v;
    b[f] = v;
  }
}

// One workgroup: relative L2 error of the network-lit SH against the
// cone-traced SH over the spot set (fixed-order reduction), the fallback
// decision, and the gate counter that enables the full cone trace.
__global__ __launch_bounds__(256) void k_spot_error(const float* spot, uint32_t spot_cap, int stride, int last_field,
                                                    const Counter* spot_ctr, float threshold, const Counter* counter,
                                                    uint32_t* stats, Counter* gate) {
  __shared__ float num_s[256], den_s[256];
  const uint32_t n_spot = spot_ctr[0].live;
  const uint32_t fields = static_cast<uint32_t>(last_field - 3);  // SH fields 4..last_field
  const uint32_t total = n_spot * fields;
  float num = 0.0f, den = 0.0f;
  for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
    const uint32_t c = t / fields, f = 4u + t % fields;
    const float ref = spot[static_cast<size_t>(c) * stride + f];
    const float net = spot[(static_cast<size_t>(n_spot) + c) * stride + f];
    const float d = net - ref;
    num = __builtin_fmaf(d, d, num);
    den = __builtin_fmaf(ref, ref, den);
  }
  num_s[threadIdx.x] = num;
  den_s[threadIdx.x] = den;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) {
      num_s[threadIdx.x] += num_s[threadIdx.x + off];
      den_s[threadIdx.x] += den_s[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float n = num_s[0], d = den_s[0];
    const float err = d > 0.0f ? sqrtf(n / d) : (n > 0.0f ? 1.0f : 0.0f);
    const bool fall = !(err <= threshold);
    stats[0] = __float_as_uint(err);
    stats[1] = fall ? 1u : 0u;
    stats[2] = n_spot;
    *gate = *counter;
    gate->live = fall ? counter->live : 0u;
  }
}

}  // namespace

size_t mlp_train_lds_bytes() { return sizeof(TrainLds); }

void launch_spot_select(const Counter* counter, uint32_t k, const float* caches, int stride, float* spot,
                        uint32_t spot_cap, Counter* spot_ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_spot_select, dim3((spot_cap + 255) / 256), dim3(256), 0, s, counter, k, caches, stride, spot,
                     spot_cap, spot_ctr);
}

void launch_spot_error(const float* spot, uint32_t spot_cap, int stride, int sh_bands, const Counter* spot_ctr,
                       float threshold, const Counter* counter, uint32_t* stats, Counter* gate, hipStream_t s) {
  const int last_field = sh_bands >= 2 ? 30 : 15;
  hipLaunchKernelGGL(k_spot_error, dim3(1), dim3(256), 0, s, spot, spot_cap, stride, last_field, spot_ctr, threshold,
                     counter, stats, gate);
}

FeatureArgsHost make_feature_args(const drv_volume& vol, const float* cam, const VoxelView& vox) {
  FeatureArgsHost f;
  for (int i = 0; i < 3; ++i) {
    f.vmin[i] = vol.volume_world_min[i];
    f.cam[i] = cam[i];
  }
  f.voxel_size = vol.voxel_size_in_world;
  f.vox = vox;
  return f;
}

static FeatureCtx to_ctx(const FeatureArgsHost& h) {
  FeatureCtx c;
  for (int i = 0; i < 3; ++i) {
    c.vmin[i] = h.vmin[i];
    c.cam[i] = h.cam[i];
  }
  c.voxel_size = h.voxel_size;
  c.vox = h.vox;
  return c;
}

void launch_mlp_pack(const float* params, void* pack, void* train_w, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, s, params, static_cast<bf16x8*>(pack),
                     static_cast<__bf16*>(train_w));
}

void launch_visibility_mlp(const FeatureArgsHost& fh, const void* pack, const float* params, const float* caches,
                           int stride, const Counter* counter, const Counter* first, const ShadowBlock* blocks,
                           uint32_t n_blocks, float* vis, uint32_t pitch, int grid, hipStream_t s) {
  InferArgs a;
  a.fc = to_ctx(fh);
  a.pack = static_cast<const bf16x8*>(pack);
  a.params = params;
  a.caches = caches;
  a.stride = stride;
  a.counter = counter;
  a.first = first;
  a.blocks = blocks;
  a.n_blocks = n_blocks;
  a.vis = vis;
  a.pitch = pitch;
  hipLaunchKernelGGL(k_visibility_mlp, dim3(grid), dim3(64 * kInferWaves), 0, s, a);
}

void launch_mlp_sample_pairs(const Counter* counter, uint32_t n_blocks, uint32_t n, uint32_t seed, uint32_t* pairs,
                             uint32_t* n_out, hipStream_t s) {
  const unsigned blocks = std::max(1u, std::min(4096u, (n + 255u) / 256u));
  hipLaunchKernelGGL(k_mlp_sample_pairs, dim3(blocks), dim3(256), 0, s, counter, n_blocks, n, seed, pairs, n_out);
}

void launch_mlp_features(const FeatureArgsHost& fh, const float* caches, int stride, uint32_t n_caches,
                         const ShadowBlock* blocks, uint32_t n_blocks, const uint32_t* pairs, uint32_t n, float* x,
                         const float* vis, uint32_t pitch, float* y, hipStream_t s) {
  if (n == 0) return;
  FeatArgs a;
  a.fc = to_ctx(fh);
  a.caches = caches;
  a.stride = stride;
  a.blocks = blocks;
  a.pairs = pairs;
  a.n = n;
  a.n_caches = n_caches;
  a.n_blocks = n_blocks;
  a.x = x;
  a.vis = vis;
  a.pitch = pitch;
  a.y = y;
  hipLaunchKernelGGL(k_mlp_features, dim3((n + 255) / 256), dim3(256), 0, s, a);
}

void launch_mlp_forward(const void* pack, const float* params, const float* x, uint32_t n, float* p, int grid,
                        hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64 * kInferWaves), 0, s, static_cast<const bf16x8*>(pack), params, x, n, p);
}

void launch_mlp_batch_indices(uint32_t n, uint32_t batch, uint32_t seed, uint32_t* idx, bool all, hipStream_t s) {
  const unsigned blocks = std::max(1u, std::min(4096u, (batch + 255u) / 256u));
  if (all)
    hipLaunchKernelGGL(k_iota, dim3(blocks), dim3(256), 0, s, batch, idx);
  else
    hipLaunchKernelGGL(k_mlp_batch_indices, dim3(blocks), dim3(256), 0, s, n, batch, seed, idx);
}

uint32_t mlp_train_waves(uint32_t batch, int cu_count) {
  const uint32_t chains = (batch + kTrainChain - 1) / kTrainChain;
  const uint32_t max_waves = static_cast<uint32_t>(cu_count) * 2u * kTrainWaves;
  const uint32_t waves = std::min(chains, max_waves);
  return ((waves + kTrainWaves - 1) / kTrainWaves) * kTrainWaves;
}

void launch_mlp_train(const float* x, const float* y, const uint32_t* idx, uint32_t batch, const float* params,
                      const void* train_w, float* partial, uint32_t n_waves, hipStream_t s) {
  TrainArgs a;
  a.x = x;
  a.y = y;
  a.idx = idx;
  a.n_batch = batch;
  a.params = params;
  a.train_w = static_cast<const __bf16*>(train_w);
  a.partial = partial;
  a.inv_batch = 1.0f / static_cast<float>(batch);
  const uint32_t chains = (batch + kTrainChain - 1) / kTrainChain;
  a.chains_per_wave = (chains + n_waves - 1) / n_waves;
  static bool lds_raised = false;  // the training kernel needs more than 64 KiB of LDS
  if (!lds_raised) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_train), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(sizeof(TrainLds)));
    lds_raised = true;
  }
  hipLaunchKernelGGL(k_mlp_train, dim3(n_waves / kTrainWaves), dim3(64 * kTrainWaves), sizeof(TrainLds), s, a);
}

void launch_mlp_adam(const float* partial, uint32_t n_waves, float* params, float* m, float* v, float* grad_out,
                     float lr, float beta1, float beta2, float eps, float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, partial, n_waves, params, m, v,
                     grad_out, lr, beta1, beta2, eps, bc1, bc2);
}

}  // namespace drvk
