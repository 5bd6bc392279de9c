// Culling grids of SDF union chains (sdf_grid.hip): host build and GPU rebuild.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "bvh_build.hpp"
#include "sdf.hpp"

namespace mrt {

// Cells per axis of a chain's grid: finer for longer chains (more primitives to cull), 2 MiB of
// masks at 64 and 16 MiB at 128.
int sdf_grid_dim(uint32_t n);
// Whether a compiled program gets a grid: one union chain of exact primitives with a non-empty box.
bool sdf_grid_applies(const SdfInsn* prog, uint32_t n, const Box& box);
// The grid's dimension, origin and scale over `box` (mask_offset untouched).
void sdf_grid_frame(const Box& box, uint32_t n, SdfGrid& g);
// Host build: fills g and appends dim^3 masks (g.dim = 0 and nothing appended if no grid applies).
void build_sdf_grid(const SdfInsn* prog, uint32_t n, const Box& box, SdfGrid& g, std::vector<uint64_t>& masks);
// The same masks built on the GPU from the device copy of the program into d_masks (dim^3
// entries); d_coarse holds kSdfGridCoarse^3 entries of scratch. Returns nonzero on a launch error.
int launch_sdf_grid(const SdfInsn* d_prog, uint32_t n, const Box& box, uint64_t* d_coarse, uint64_t* d_masks,
                    hipStream_t stream);

}  // namespace mrt
