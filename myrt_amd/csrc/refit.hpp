// Instance-BVH refit (refit.hip) for frames that only move instances.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.hpp"

namespace mrt {

struct RefitParams {
  const float* matrices;      // n_inst column-major 4x4 object -> world matrices
  uint32_t n_inst;
  InstanceRecord* inst;       // rows of the inverse are rewritten; geom, blas_root, inst kept
  const float* geo_box;       // per geometry: min xyz, max xyz
  float* inst_box;            // per instance world box (written)
  WideNode* tlas;             // quantised child boxes rewritten in place
  const uint32_t* items;      // TLAS leaf slots -> instance index
  const uint32_t* level_nodes;  // node ids, deepest level first
  float* node_box;            // per wide node: union of its children's boxes (written)
  uint32_t* status;           // bit 0: a matrix was singular
  double pad_rel;             // the collapse's relative padding of child boxes
};

// Runs the instance pass and one launch per level on stream s (no host synchronisation).
// level_start has n_levels + 1 entries into level_nodes. Returns nonzero on a launch error.
int launch_refit(const RefitParams& p, const uint32_t* level_start, uint32_t n_levels, hipStream_t s);
// The same passes in one workgroup and one launch, for TLASes of up to kRefitOneBlockNodes nodes;
// d_level_start is the device copy of level_start, and the status word is cleared by the kernel.
constexpr uint32_t kRefitOneBlockNodes = 4096;
int launch_refit_one_block(const RefitParams& p, const uint32_t* d_level_start, uint32_t n_levels, hipStream_t s);

}  // namespace mrt
