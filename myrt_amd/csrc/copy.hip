// Plain streaming copy, used to calibrate hardware byte counters against a known traffic.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mrt {
namespace {

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

int launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream) {
  const size_t n16 = bytes / 16;
  const size_t blocks = (n16 + 255) / 256;
  const unsigned grid = unsigned(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, stream, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n16);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
