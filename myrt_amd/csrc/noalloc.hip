// Closest-hit trace that allocates nothing, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the kernels of mrt_trace_closest_device with their scratch (the counters, the
// traversal spill stack and, with SDF drawables, the per-ray SDF hit points and hit list) carved
// from one file-scope device array of kScratchBytes per GPU, instead of the scene's own buffers.
// The array's device address is looked up once per device with hipGetSymbolAddress; the host
// symbol itself is only a stand-in and is never passed to a kernel. Counters are cleared by a
// fill kernel, so the launches can be captured. Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU (two streams, two threads) race and must be serialised by the
// caller.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/myrt.h"
#include "abi_guard.hpp"
#include "scene_internal.hpp"

namespace mrt {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // per GPU, held until the process ends
constexpr size_t kCtrlBytes = 256;                  // ray counter, flags, SDF hit count, stats
constexpr size_t kMaxSdfRays = size_t(1) << 20;     // SDF side data: 20 B per ray
constexpr int kMaxDevices = 64;

}  // namespace
}  // namespace mrt

static __device__ __align__(256) uint8_t scratch[mrt::kScratchBytes];

namespace mrt {
namespace {

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call there).
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(MRT_ERR_LIMIT, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress of the trace scratch");
    if (!p) return fail(MRT_ERR_HIP, "hipGetSymbolAddress returned no address for the trace scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MRT_OK;
}

}  // namespace
}  // namespace mrt

using namespace mrt;

extern "C" {

int mrt_trace_closest_device_noalloc(mrt_scene sh, const mrt_ray* d_rays, size_t n, mrt_hit* d_hits, void* stream) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && (!d_rays || !d_hits))) return fail(MRT_ERR_INVALID, "null argument");
  if (!s->prepared) return fail(MRT_ERR_NOT_PREPARED, "trace before mrt_scene_prepare");
  if (n == 0) return MRT_OK;
  if (n >= 0x7fffffffu) return fail(MRT_ERR_LIMIT, "at most 2^31-1 rays per call");
  const bool sdf = !s->sdf_drawables.empty();
  if (sdf && n > kMaxSdfRays)
    return fail(MRT_ERR_LIMIT, "at most 1048576 rays per call when the scene has SDF drawables");
  void* base = nullptr;
  const int rc = scratch_address(&base);
  if (rc != MRT_OK) return rc;
  uint8_t* p = static_cast<uint8_t*>(base);
  Scene::TraceArena a{};
  a.ctrl = reinterpret_cast<uint32_t*>(p);
  size_t off = kCtrlBytes;
  if (sdf) {
    a.sdf_side = reinterpret_cast<float4*>(p + off);
    off += n * sizeof(float4);
    a.sdf_hits = reinterpret_cast<uint32_t*>(p + off);
    off += n * sizeof(uint32_t);
    off = (off + 255) & ~size_t(255);
    a.sdf_cap = n;
  }
  // The rest holds the spill stack; the persistent grid shrinks to what fits.
  a.spill = reinterpret_cast<uint32_t*>(p + off);
  a.spill_entries = (kScratchBytes - off) / sizeof(uint32_t);
  return s->trace_in(a, d_rays, n, d_hits, false, false, static_cast<hipStream_t>(stream), nullptr);
  MRT_GUARD_END
}

}  // extern "C"
