// Ray binning before traversal (raysort.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

constexpr uint32_t kRaySortBins = 1u << 15;  // 3 octant bits + 12 Morton bits

struct RaySortArgs {
  const void* rays;        // mrt_ray[n]
  uint32_t n;              // buffer size (rays beyond *n_dev are ignored)
  const uint32_t* n_dev;   // device-side ray count, if set
  float lo[3];             // scene box
  float inv_ext[3];        // 1 / extent per axis
  uint32_t* keys;          // [n]
  uint32_t* counts;        // [kRaySortBins]
  uint32_t* perm;          // [n]: sorted position -> ray index
};

int launch_ray_sort(const RaySortArgs& a, hipStream_t s);

}  // namespace mrt
