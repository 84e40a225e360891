// Issue rate of the FP32 vector instruction kinds the cone trace and the SH
// gather use, for calibrating the VALU figures of the kernel profiles: each
// lane runs 8 independent chains of one instruction kind, 32 workgroups of
// 256 threads per CU, so the SIMDs always have waves ready to issue.  A
// calibration aid only (libdrv_calib.so, not part of the drv.h interface).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 8;

template <int kOp>
__global__ __launch_bounds__(256) void k_rate(float* out, float seed, int iters) {
  float a[kChains];
  uint32_t u[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    a[i] = seed + static_cast<float>(threadIdx.x) * 1e-3f + static_cast<float>(i);
    u[i] = threadIdx.x + static_cast<uint32_t>(i);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
      if (kOp == 0) a[i] = __builtin_fmaf(a[i], 0.999f, 0.001f);                       // v_fma
      if (kOp == 1) a[i] = a[i] + 0.5f;                                                    // v_add
      if (kOp == 2) a[i] = __builtin_amdgcn_rsqf(a[i]);                                   // v_rsq
      if (kOp == 3) a[i] = __builtin_amdgcn_rcpf(a[i]);                                   // v_rcp
      if (kOp == 4) a[i] = __builtin_amdgcn_fmed3f(a[i], 0.5f, 3.0f) + 0.25f;             // v_med3 + v_add
      if (kOp == 5) a[i] = __builtin_amdgcn_fractf(a[i]) + 1.25f;                          // v_fract + v_add
      if (kOp == 6) u[i] = static_cast<uint32_t>(static_cast<float>(u[i]) * 1.0f) + 1u;  // 2 conversions + v_add
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += a[i] + static_cast<float>(u[i]);
  out[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

template <int kOp>
void launch(float* out, int blocks, int iters) {
  hipLaunchKernelGGL(k_rate<kOp>, dim3(blocks), dim3(256), 0, 0, out, 1.5f, iters);
}

}  // namespace

// Times instruction kind `op` (0 fma, 1 add, 2 rsq, 3 rcp, 4 med3 + add,
// 5 fract + add, 6 two conversions + add) on device 0: `iters` loop
// iterations of 8 chains in every lane of 32 x `cus` workgroups.  Writes the
// kernel time in ms and the number of wave64 instructions of that kind per
// SIMD; returns 0, or -1 on a HIP error or an unknown op.
extern "C" int drv_calib_valu(int op, int iters, int cus, float* ms, double* wave_insts_per_simd) {
  if (op < 0 || op > 6 || iters <= 0 || cus <= 0 || !ms || !wave_insts_per_simd) return -1;
  const int blocks = 32 * cus;
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * static_cast<size_t>(blocks) * 256) != hipSuccess) return -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
      (void)hipEventRecord(e0, nullptr);
      switch (op) {
        case 0: launch<0>(out, blocks, iters); break;
        case 1: launch<1>(out, blocks, iters); break;
        case 2: launch<2>(out, blocks, iters); break;
        case 3: launch<3>(out, blocks, iters); break;
        case 4: launch<4>(out, blocks, iters); break;
        case 5: launch<5>(out, blocks, iters); break;
        default: launch<6>(out, blocks, iters); break;
      }
      (void)hipEventRecord(e1, nullptr);
    }
    if (hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess &&
        hipEventElapsedTime(ms, e0, e1) == hipSuccess) {
      *wave_insts_per_simd = static_cast<double>(blocks) * 4.0 * iters * kChains / (4.0 * cus);
      rc = 0;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}
