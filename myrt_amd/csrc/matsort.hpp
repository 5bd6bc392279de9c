// Shading order by material (matsort.hip): a counting sort of the live path queue by the
// material of each closest hit, so that the lanes of one shading wave take the same BSDF branch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

constexpr uint32_t kMatSortBins = 1024;  // bin 0: misses; bin 1 + m for material m (m >= 1022 share)
constexpr uint32_t kMatSortItems = 4096;  // queue entries per block (16 per thread)

struct MatSortArgs {
  const void* hits;                // mrt_hit[n]
  const uint32_t* n_dev;           // live queue length (device)
  uint32_t n;                      // queue capacity
  const uint32_t* inst_material;   // instance -> material
  uint32_t* keys;                  // [n]
  uint32_t* counts;                // [kMatSortBins]: histogram, then the bins' write cursors
  uint32_t* perm;                  // [n]: shading position -> queue entry
};

int launch_material_sort(const MatSortArgs& a, hipStream_t s);

}  // namespace mrt
