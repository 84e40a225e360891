// Allocation-free light pass (drv_light_caches_noalloc): one light of one
// view lit by the persistent work-queue kernel of the batched path, with
// every buffer beyond the caller's inputs and the context's cache entries -
// the VPL and shadow-block records, the queue descriptor and the split
// partials - in one file-scope device array.
//
// The array is a property of the loaded code object: each GPU holds its own
// copy, at its own address, from the first use of this library on it until
// the process ends.  Its device address comes from hipGetSymbolAddress, taken
// once per device when the first context is created on it; host code never
// passes the array's name itself (a host-side stand-in) to a kernel.
//
// The fused kernel traces each (cache, shadow block) pair inside the VPL sum
// (cacheLightingRSM.comp:167-279), so no per-pair visibility buffer is needed
// and the scratch size does not depend on the cache count; the plan kernel
// lengthens the VPL tile until the split partials fit the array.
#include <mutex>

#include "drv_internal.hpp"

namespace drvk {

// The scratch array (device memory of every GPU the library runs on).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

// Layout of the scratch array.
constexpr size_t kDescOffset = 0;       // one BatchView
constexpr size_t kQueueOffset = 1024;   // one BatchQueue
constexpr size_t kVplOffset = 4096;     // kStaticMaxVpls Vpl records
constexpr size_t kBlockOffset = kVplOffset + kStaticMaxVpls * sizeof(Vpl);
constexpr size_t kPartialOffset = kBlockOffset + kStaticMaxVpls * sizeof(ShadowBlock);
static_assert(sizeof(BatchView) <= kQueueOffset - kDescOffset, "descriptor slot");
static_assert(sizeof(BatchQueue) <= kVplOffset - kQueueOffset, "queue slot");
static_assert(kPartialOffset % 256 == 0 && kPartialOffset < kStaticScratchBytes, "partial slot");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_ptr[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

// The view descriptor arrives as a kernel argument, so no host-to-device copy
// is needed.
__global__ void k_static_setup(BatchView view, BatchView* desc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *desc = view;
}

}  // namespace

hipError_t static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_done[device]) return g_err[device];
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_done[device] = true;
  g_err[device] = e;
  g_ptr[device] = e == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
  return e;
}

unsigned char* static_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_ptr[device];
}

void launch_light_static(unsigned char* base, const drv_rsm& rsm, const drv_spot_light& light, BatchView view,
                         int stride, int sh_bands, int shadow, const drv_sh_constants& shc, int grid,
                         hipStream_t s) {
  Vpl* vpls = reinterpret_cast<Vpl*>(base + kVplOffset);
  ShadowBlock* blocks = reinterpret_cast<ShadowBlock*>(base + kBlockOffset);
  BatchView* desc = reinterpret_cast<BatchView*>(base + kDescOffset);
  BatchQueue* queue = reinterpret_cast<BatchQueue*>(base + kQueueOffset);
  float* partial = reinterpret_cast<float*>(base + kPartialOffset);
  const size_t partial_floats = (kStaticScratchBytes - kPartialOffset) / sizeof(float);
  launch_prepare_vpls(rsm, light, vpls, shadow ? blocks : nullptr, 0, nullptr, nullptr, s);
  view.vpls = vpls;
  view.blocks = shadow ? blocks : nullptr;
  hipLaunchKernelGGL(k_static_setup, dim3(1), dim3(64), 0, s, view, desc);
  // 32 caches per wave (two lane groups per cache); the VPL tile starts at
  // 256 VPLs (or one shadow block) and the plan kernel doubles it until the
  // partials fit.
  const uint32_t tile = view.interval > 256u ? view.interval : 256u;
  launch_batch(desc, 1u, queue, stride, sh_bands, shadow, 32u, tile, partial, partial_floats, shc, grid, s);
}

}  // namespace drvk
