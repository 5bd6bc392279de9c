// Ray binning before traversal: a one-pass counting sort of ray indices by a 15-bit key
// (direction octant, then a 4-bit-per-axis Morton code of the origin inside the scene box), so
// that the rays of one wave start close together and head the same way. Order inside a bin is
// whatever the atomics give; traversal results do not depend on it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "raysort.hpp"

namespace mrt {
namespace rsk {

constexpr int kBlock = 256;

__device__ inline uint32_t spread4(uint32_t v) {  // 4 bits -> every third bit
  v &= 0xfu;
  return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}

__device__ inline uint32_t ray_key(const RaySortArgs& a, uint32_t i) {
  const float4* rp = reinterpret_cast<const float4*>(a.rays) + 2 * size_t(i);
  const float4 o = rp[0], d = rp[1];
  const uint32_t oct = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
  uint32_t c[3];
  const float ov[3] = {o.x, o.y, o.z};
  for (int k = 0; k < 3; ++k) {
    const float f = (ov[k] - a.lo[k]) * a.inv_ext[k] * 16.0f;
    c[k] = f > 0.f ? (f < 15.0f ? uint32_t(f) : 15u) : 0u;  // NaN -> 0
  }
  return (oct << 12) | (spread4(c[0]) << 2) | (spread4(c[1]) << 1) | spread4(c[2]);
}

__global__ __launch_bounds__(kBlock) void histogram_kernel(RaySortArgs a) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
  if (i >= n) return;
  const uint32_t key = ray_key(a, i);
  a.keys[i] = key;
  atomicAdd(a.counts + key, 1u);
}

// Exclusive scan of the kRaySortBins counters in one block; the cursors start at the offsets.
__global__ __launch_bounds__(1024) void scan_kernel(RaySortArgs a) {
  __shared__ uint32_t part[1024];
  constexpr int per = kRaySortBins / 1024;
  const int t = threadIdx.x;
  uint32_t local[per];
  uint32_t s = 0;
  for (int k = 0; k < per; ++k) {
    local[k] = a.counts[t * per + k];
    s += local[k];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;  // exclusive prefix of this thread's chunk
  for (int k = 0; k < per; ++k) {
    a.counts[t * per + k] = run;
    run += local[k];
  }
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(RaySortArgs a) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
  if (i >= n) return;
  const uint32_t slot = atomicAdd(a.counts + a.keys[i], 1u);
  a.perm[slot] = i;
}

}  // namespace rsk

int launch_ray_sort(const RaySortArgs& a, hipStream_t s) {
  if (a.n == 0) return 0;
  if (hipMemsetAsync(a.counts, 0, kRaySortBins * sizeof(uint32_t), s) != hipSuccess) return -1;
  const unsigned grid = unsigned((a.n + rsk::kBlock - 1) / rsk::kBlock);
  hipLaunchKernelGGL(rsk::histogram_kernel, dim3(grid), dim3(rsk::kBlock), 0, s, a);
  hipLaunchKernelGGL(rsk::scan_kernel, dim3(1), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(rsk::scatter_kernel, dim3(grid), dim3(rsk::kBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
