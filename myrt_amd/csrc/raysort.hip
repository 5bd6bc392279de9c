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

// ---- XCD groups: 8 slabs, block-local histograms (one global atomic per block and group) ----
constexpr int kGroupPer = 16;  // rays per thread

__device__ inline uint32_t group_key(const RayGroupArgs& a, uint32_t i) {
  const float4 o = reinterpret_cast<const float4*>(a.rays)[2 * size_t(i)];
  const float v = a.axis == 0 ? o.x : (a.axis == 1 ? o.y : o.z);
  const float f = (v - a.lo) * a.inv_ext * float(kXcdGroups);
  return f > 0.f ? (f < float(kXcdGroups - 1) ? uint32_t(f) : kXcdGroups - 1) : 0u;  // NaN -> 0
}

__global__ __launch_bounds__(kBlock) void group_histogram_kernel(RayGroupArgs a) {
  __shared__ uint32_t hist[kXcdGroups];
  if (threadIdx.x < kXcdGroups) hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
  const uint32_t base = blockIdx.x * kBlock * kGroupPer;
  for (int j = 0; j < kGroupPer; ++j) {
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    if (i < n) {
      const uint32_t k = group_key(a, i);
      a.keys[i] = uint8_t(k);
      atomicAdd(&hist[k], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < kXcdGroups && hist[threadIdx.x]) atomicAdd(a.counts + threadIdx.x, hist[threadIdx.x]);
}

__global__ void group_scan_kernel(RayGroupArgs a) {
  if (threadIdx.x != 0) return;
  uint32_t run = 0;
  for (uint32_t g = 0; g < kXcdGroups; ++g) {
    const uint32_t c = a.counts[g];
    a.counts[g] = run;
    a.group_start[g] = run;
    run += c;
  }
  a.group_start[kXcdGroups] = run;
}

__global__ __launch_bounds__(kBlock) void group_scatter_kernel(RayGroupArgs a) {
  __shared__ uint32_t hist[kXcdGroups];
  if (threadIdx.x < kXcdGroups) hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
  const uint32_t base = blockIdx.x * kBlock * kGroupPer;
  uint32_t key[kGroupPer];
  for (int j = 0; j < kGroupPer; ++j) {
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    key[j] = i < n ? uint32_t(a.keys[i]) : kXcdGroups;
    if (key[j] < kXcdGroups) atomicAdd(&hist[key[j]], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kXcdGroups && hist[threadIdx.x]) hist[threadIdx.x] = atomicAdd(a.counts + threadIdx.x, hist[threadIdx.x]);
  __syncthreads();
  for (int j = 0; j < kGroupPer; ++j) {
    if (key[j] >= kXcdGroups) continue;
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    a.perm[atomicAdd(&hist[key[j]], 1u)] = i;
  }
}

}  // namespace rsk

int launch_ray_groups(const RayGroupArgs& a, hipStream_t s) {
  if (a.n == 0) return 0;
  if (hipMemsetAsync(a.counts, 0, (kXcdGroups + 1) * sizeof(uint32_t), s) != hipSuccess) return -1;
  const unsigned grid = unsigned((a.n + rsk::kBlock * rsk::kGroupPer - 1) / (rsk::kBlock * rsk::kGroupPer));
  hipLaunchKernelGGL(rsk::group_histogram_kernel, dim3(grid), dim3(rsk::kBlock), 0, s, a);
  hipLaunchKernelGGL(rsk::group_scan_kernel, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(rsk::group_scatter_kernel, dim3(grid), dim3(rsk::kBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

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
