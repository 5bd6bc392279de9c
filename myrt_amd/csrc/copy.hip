// Plain streaming copy, used to calibrate hardware byte counters against a known traffic, and a
// small fill kernel that clears counters where a hipMemsetAsync would not do (a caller may
// capture the launches into a HIP graph).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mrt {
namespace {

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(64) void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = threadIdx.x; i < n; i += 64) p[i] = v;
}

}  // namespace

int launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream) {
  const size_t n16 = bytes / 16;
  const size_t blocks = (n16 + 255) / 256;
  const unsigned grid = unsigned(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, stream, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n16);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
