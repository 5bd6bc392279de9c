// Shading order by material: the path queue's closest hits are binned by material (misses in
// bin 0, SDF hits by the material in the hit record, triangle hits by their instance's
// material), and the shade kernel reads its inputs through the resulting permutation. Each
// block histograms kMatSortItems entries in LDS and adds only its non-zero bins to the global
// counters, so a scene with a handful of materials costs a few thousand global atomics per
// bounce rather than one per path. Order inside a bin is whatever the atomics give; every
// path's result is the same in any order (the sampler is indexed by pixel and sample).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "matsort.hpp"

namespace mrt {
namespace msk {

constexpr int kBlock = 256;
constexpr int kPer = int(kMatSortItems) / kBlock;

__device__ inline uint32_t live_count(const MatSortArgs& a) { return min(*a.n_dev, a.n); }

__device__ inline uint32_t material_key(const MatSortArgs& a, uint32_t i) {
  const uint4 h = reinterpret_cast<const uint4*>(a.hits)[i];
  if (h.z == 0xFFFFFFFFu) return 0u;
  const uint32_t m = (h.z & 0x80000000u) ? h.y : a.inst_material[h.z];
  return 1u + min(m, kMatSortBins - 2u);
}

__global__ __launch_bounds__(kBlock) void histogram_kernel(MatSortArgs a) {
  __shared__ uint32_t hist[kMatSortBins];
  for (uint32_t k = threadIdx.x; k < kMatSortBins; k += kBlock) hist[k] = 0;
  __syncthreads();
  const uint32_t n = live_count(a);
  const uint32_t base = blockIdx.x * kMatSortItems;
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    if (i < n) {
      const uint32_t key = material_key(a, i);
      a.keys[i] = key;
      atomicAdd(&hist[key], 1u);
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kMatSortBins; k += kBlock)
    if (hist[k]) atomicAdd(a.counts + k, hist[k]);
}

// Exclusive scan of the kMatSortBins counters in one block of kMatSortBins threads.
__global__ __launch_bounds__(kMatSortBins) void scan_kernel(MatSortArgs a) {
  __shared__ uint32_t part[kMatSortBins];
  const uint32_t t = threadIdx.x;
  const uint32_t c = a.counts[t];
  part[t] = c;
  __syncthreads();
  for (uint32_t off = 1; off < kMatSortBins; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  a.counts[t] = part[t] - c;
}

// Each block reserves one range per non-zero bin, then places its entries inside those ranges.
__global__ __launch_bounds__(kBlock) void scatter_kernel(MatSortArgs a) {
  __shared__ uint32_t hist[kMatSortBins];
  for (uint32_t k = threadIdx.x; k < kMatSortBins; k += kBlock) hist[k] = 0;
  __syncthreads();
  const uint32_t n = live_count(a);
  const uint32_t base = blockIdx.x * kMatSortItems;
  uint32_t key[kPer];
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    key[j] = i < n ? a.keys[i] : kMatSortBins;
    if (key[j] < kMatSortBins) atomicAdd(&hist[key[j]], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kMatSortBins; k += kBlock)
    if (hist[k]) hist[k] = atomicAdd(a.counts + k, hist[k]);
  __syncthreads();
  for (int j = 0; j < kPer; ++j) {
    if (key[j] >= kMatSortBins) continue;
    const uint32_t i = base + uint32_t(j * kBlock) + threadIdx.x;
    a.perm[atomicAdd(&hist[key[j]], 1u)] = i;
  }
}

}  // namespace msk

int launch_material_sort(const MatSortArgs& a, hipStream_t s) {
  if (a.n == 0) return 0;
  if (hipMemsetAsync(a.counts, 0, kMatSortBins * sizeof(uint32_t), s) != hipSuccess) return -1;
  const unsigned grid = unsigned((a.n + kMatSortItems - 1) / kMatSortItems);
  hipLaunchKernelGGL(msk::histogram_kernel, dim3(grid), dim3(msk::kBlock), 0, s, a);
  hipLaunchKernelGGL(msk::scan_kernel, dim3(1), dim3(kMatSortBins), 0, s, a);
  hipLaunchKernelGGL(msk::scatter_kernel, dim3(grid), dim3(msk::kBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
