// Launcher of the SDF sphere-tracing pass (sdf.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdf.hpp"

namespace mrt {

struct SdfParams {
  const SdfDrawable* drawables;
  const SdfInsn* prog;     // compiled tapes; drawables address them
  const SdfGrid* grids;    // per drawable
  const uint64_t* masks;   // culling masks of all grids
  uint32_t num_drawables;
  const void* rays;        // mrt_ray[n]
  void* out;               // mrt_hit[n] (closest) or uint8_t[n] (occluded)
  uint32_t n;
  const uint32_t* n_dev;   // device-side ray count, if set
  int steps;               // sdf_marching_steps (sequential_pathtracer.hpp:135, default 400)
  float eps;               // sdf_marching_epsilon (default 1e-5)
  float4* side;            // [n]: object-space hit point and ray-length scale of SDF hits
  uint32_t prog_len;       // instructions in prog (all drawables)
  uint32_t chain_only;     // 1: every drawable's program has a culling grid (a leaner kernel serves)
  uint32_t* hit_list;      // closest hits: rays that hit an SDF, appended by the march (n entries), or null
  uint32_t* hit_count;     // entries in hit_list, zero before the launch
};

int launch_sdf(const SdfParams& p, bool any, int num_cus, hipStream_t stream);

}  // namespace mrt
