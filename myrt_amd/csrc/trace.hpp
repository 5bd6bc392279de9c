// Host-visible interface of the traversal kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.hpp"

namespace mrt {

struct TraceParams {
  const WideNode* tlas_nodes;
  const uint32_t* tlas_items;      // TLAS leaf slot -> instance index
  const InstanceRecord* instances;
  const WideNode* blas_nodes;
  const TriRecord* tris;
  const void* rays;                // mrt_ray[n]
  void* out;                       // mrt_hit[n] or uint8_t[n]
  uint32_t n;
  const uint32_t* n_dev;           // if set, the ray count is read from device memory (wavefront queues)
  uint32_t has_tlas;
  uint32_t* counter;               // ray counter, zero before the launch
  uint32_t* flags;                 // bit 0: stack overflow
  uint32_t* spill;                 // [spill_depth][total_threads]
  uint32_t spill_depth;
  uint32_t total_threads;
  uint32_t refill_below;           // refill a wave when fewer lanes than this are busy
  uint32_t grab;                   // rays a wave reserves per atomic on the ray counter (>= 1)
  uint32_t lds_depth;              // LDS stack entries per lane (a power of two)
  uint32_t waves;                  // waves per SIMD the kernel is compiled for: 6 or 7
  uint32_t single_instance;        // the frame's only instance record, or 0xFFFFFFFF
  uint32_t leaf_wait;              // a lane at a leaf waits until leaf_wait/64 of the busy lanes are (0: never)
  unsigned long long* stats;       // rays, node visits, triangle tests, instance visits
};

int launch_trace(const TraceParams& p, bool any, bool stats, int grid, hipStream_t stream);
int trace_blocks_per_cu(bool any, bool stats, uint32_t waves, uint32_t lds_depth, bool single);
constexpr int kTraceBlock = 256;
constexpr uint32_t kTraceStackLds = 16;  // default LDS stack entries per lane

}  // namespace mrt
