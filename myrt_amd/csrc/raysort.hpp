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

// Ray grouping for XCD-local traversal: rays are split into kXcdGroups slabs of their origin
// along the scene box's widest axis, and perm lists each group's rays contiguously
// (group_start[g] .. group_start[g + 1]).
constexpr uint32_t kXcdGroups = 8;
struct RayGroupArgs {
  const void* rays;        // mrt_ray[n]
  uint32_t n;              // buffer size
  const uint32_t* n_dev;   // device-side ray count, if set
  float lo, inv_ext;       // the widest axis of the scene box
  int axis;
  uint8_t* keys;           // [n]
  uint32_t* counts;        // [kXcdGroups + 1]: histogram, then cursors; group_start follows
  uint32_t* group_start;   // [kXcdGroups + 1]
  uint32_t* perm;          // [n]
};
int launch_ray_groups(const RayGroupArgs& a, hipStream_t s);

}  // namespace mrt
