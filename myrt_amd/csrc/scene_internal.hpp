// Scene state behind the mrt_scene handle.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "bvh_build.hpp"
#include "device_types.hpp"
#include "material_layout.hpp"
#include "pathtrace.hpp"
#include "sdf.hpp"
#include "wide_bvh.hpp"

namespace mrt {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
bool invert_affine_rows(const float m[16], float out[12]);
Box transform_box(const Box& b, const float m[16]);

// geometry_t plus the arrays push_geometry appends (src/pathtracer/scene.cpp:7-12, 43-71).
struct Geometry {
  std::vector<uint32_t> idx;
  std::vector<float> pos, nrm, uv;
  BinaryBvh bvh;  // object BVH, split_aabbs
  WideBvh wide;
  Box box;        // bvh root box = geometry_t::aabb
};

// drawable_geometry_t (src/pathtracer/scene.hpp:97-104).
struct Instance {
  uint32_t geom = 0, material = 0;
  float m[16];     // column-major
  float inv[12];   // rows 0-2 of the inverse
  Box world_box;
};

// An enqueued SDF (scene::enqueue(sdf, mat4), src/pathtracer/scene.cpp:411-417).
struct SdfInstance {
  uint32_t sdf = 0;
  float m[16];
  float inv[12];
};

struct Scene {
  std::vector<Geometry> geoms;
  // SDF programs (tapes concatenated), their object-space bounds and this frame's drawables.
  std::vector<SdfNode> sdf_tape;
  std::vector<uint32_t> sdf_offset, sdf_len;
  std::vector<SdfInsn> sdf_prog;  // compiled tapes (device form)
  std::vector<uint32_t> sdf_prog_offset, sdf_prog_len;
  std::vector<SdfGrid> sdf_grid;    // per program
  std::vector<uint64_t> sdf_masks;  // culling masks of all grids
  std::vector<Box> sdf_bounds;
  std::vector<SdfInstance> sdf_pending, sdf_current;
  std::vector<SdfDrawable> sdf_drawables;
  bool sdf_dirty = false;
  bool sdf_masks_dirty = false;       // masks appended on the host since the last upload
  bool sdf_masks_host_stale = false;  // the device holds GPU-rebuilt masks the host copy lacks
  std::vector<uint32_t> sdf_grid_rebuild;  // programs whose grids the next prepare rebuilds on the GPU
  uint64_t* d_sdf_coarse = nullptr;   // coarse-pass scratch of the GPU grid build
  int sdf_steps = 400;      // sequential_pathtracer.hpp:135
  float sdf_eps = 1e-5f;    // sequential_pathtracer.hpp:136
  SdfDrawable* d_sdf_drawables = nullptr;
  SdfInsn* d_sdf_prog = nullptr;
  SdfGrid* d_sdf_grids = nullptr;
  uint64_t* d_sdf_masks = nullptr;
  size_t cap_sdf_grids = 0, cap_sdf_masks = 0;
  size_t cap_sdf_dr = 0, cap_sdf_tape = 0;
  std::vector<Instance> pending;  // enqueued since the last prepare
  std::vector<Instance> current;  // the prepared frame
  bool geoms_dirty = false, inst_dirty = false, prepared = false;

  // Shading data: material layouts and parameters (material_buffer, src/material.hpp:72-118),
  // albedo textures, quad lights, environment.
  MaterialStore mats;
  std::vector<TexInfo> textures;   // albedo textures (mrt_scene_push_texture), ids from 1
  std::vector<float4> texels;
  std::vector<QuadLight> lights;
  float env_scale = 1.0f;
  bool shading_dirty = true;

  // Host copies of the device arrays.
  std::vector<WideNode> blas_nodes, tlas_nodes;
  std::vector<TriRecord> tris;
  std::vector<uint32_t> tlas_items, blas_root;
  std::vector<InstanceRecord> inst_records;
  BinaryBvh tlas;
  uint32_t blas_depth = 0, tlas_depth = 0, max_stack = 0;
  std::vector<GeoInfo> geo_info;
  std::vector<uint32_t> all_idx, inst_material;
  std::vector<float4> all_pos, all_nrm;
  std::vector<float2> all_uv;

  // Device arrays.
  WideNode *d_blas = nullptr, *d_tlas = nullptr;
  TriRecord* d_tris = nullptr;
  uint32_t* d_tlas_items = nullptr;
  InstanceRecord* d_inst = nullptr;
  size_t cap_blas = 0, cap_tris = 0, cap_tlas = 0, cap_items = 0, cap_inst = 0;
  GeoInfo* d_geo_info = nullptr;
  uint32_t *d_idx = nullptr, *d_inst_material = nullptr;
  float4 *d_pos = nullptr, *d_nrm = nullptr;
  Material* d_materials = nullptr;
  QuadLight* d_lights = nullptr;
  float2* d_uv = nullptr;
  TexInfo* d_textures = nullptr;
  float4* d_texels = nullptr;
  size_t cap_uv = 0, cap_textures = 0, cap_texels = 0;
  size_t cap_geo_info = 0, cap_idx = 0, cap_inst_mat = 0, cap_pos = 0, cap_nrm = 0, cap_mat = 0, cap_lights = 0;
  // Per-trace scratch. Context 0 serves every trace by default; the others let traces run at the
  // same time on other streams (the renderer's shadow rays and its second sample pipeline).
  struct TraceScratch {
    uint32_t* ctrl = nullptr;  // 256 B: [0] ray counter, [1] flags, [2] SDF ray counter, [8..15] four u64 counters
    uint32_t* spill = nullptr;  // traversal stack entries beyond the LDS part
    size_t spill_cap = 0;
    float4* sdf_side = nullptr;  // per-ray SDF hit points between the march and the normal pass
    uint32_t* sdf_hits = nullptr;  // rays that hit an SDF, for the normal pass
    size_t sdf_side_cap = 0;
  };
  static constexpr int kTraceContexts = 4;
  TraceScratch ctx[kTraceContexts];
  // Pinned host copy of every context's 64-B trace counter block, filled by read_ctrl.
  uint32_t* ctrl_host = nullptr;
  // Device staging of the host-buffer queries (mrt_trace_closest/any, mrt_scene_pick); they
  // grow to the largest call and are reused.
  void *d_host_rays = nullptr, *d_host_out = nullptr;
  size_t cap_host_rays = 0, cap_host_out = 0;
  bool sdf_chain_only = false;  // every enqueued SDF program has a culling grid
  int num_cus = 0;
  // MRT_TRACE_REFILL: refill a traversal wave below this many busy lanes; MRT_TRACE_LEAF_WAIT:
  // lanes at a leaf wait for leaf_wait/64 of the busy lanes (0: never). Unset (kKnobUnset), each
  // takes the measured default for the frame: 52 and 20 with a single instance, 56 and 16 with
  // several.
  static constexpr uint32_t kKnobUnset = 0xFFFFFFFFu;
  uint32_t refill_below = kKnobUnset;
  uint32_t leaf_wait = kKnobUnset;
  uint32_t trace_grab = 128;  // MRT_TRACE_GRAB: rays a traversal wave reserves per counter atomic
  bool sah_builder = true;  // traversal trees from build_sah_bvh (MRT_BVH_BUILDER=0: the reference builder's)
  uint32_t trace_waves = 0;  // 0: by scene (7 with several instances, else 6); MRT_TRACE_WAVES=6|7
  // Largest wide leaf: binary leaves are split at centroid medians down to this many triangles
  // (instances for the TLAS). Measured best: 2 triangles, 1 instance. MRT_BVH_LEAF and
  // MRT_TLAS_LEAF override.
  uint32_t blas_leaf = 2;
  uint32_t tlas_leaf = 1;
  uint32_t lds_depth = 16;  // LDS stack entries per lane; MRT_TRACE_LDS_STACK overrides (a power of two, 4-64)

  ~Scene();
  int build_host();
  void tlas_levels();
  int upload_all();
  int trace(const void* d_rays, size_t n, void* d_out, bool any, bool stats, hipStream_t stream,
            const uint32_t* n_dev = nullptr, int context = 0);
  // The device scratch of one trace call: 256 B of counters (zeroed by the call), the traversal
  // spill stack (spill_entries 4-byte entries; the persistent grid is cut down to what fits) and,
  // for closest hits with SDF drawables, per-ray SDF hit points and the SDF hit list (sdf_cap rays).
  struct TraceArena {
    uint32_t* ctrl;
    uint32_t* spill;
    size_t spill_entries;
    float4* sdf_side;
    uint32_t* sdf_hits;
    size_t sdf_cap;
  };
  // Entries of spill stack per traversal thread, and the full persistent grid of a launch.
  uint32_t spill_depth() const;
  int full_grid(bool any, bool stats, size_t n) const;
  // The launches of a trace into caller-provided scratch; allocates nothing. Counters are cleared
  // with a fill kernel, so the launches may be captured into a HIP graph.
  int trace_in(const TraceArena& a, const void* d_rays, size_t n, void* d_out, bool any, bool stats,
               hipStream_t stream, const uint32_t* n_dev);
  PtScene shading_view() const;
  int read_ctrl(uint32_t* flags, unsigned long long* stats, hipStream_t stream);

  // Instance-BVH refit (refit.hip). A frame that changes only the matrices of its drawables
  // refits the prepared TLAS on the GPU when refit_limit allows (0: rebuild on every change, as
  // the reference does); mrt_scene_update_transforms_device refits from device matrices.
  uint32_t refit_limit = 0;
  uint32_t refits_done = 0;           // refits since the last rebuild
  bool device_newer = false;          // device transforms the host copies lack
  bool tlas_host_stale = false;       // tlas_nodes / inst_records lag the device
  std::vector<uint32_t> tlas_level_nodes, tlas_level_start;  // wide TLAS nodes, deepest level first
  std::vector<float> geo_boxes;       // per geometry: min xyz, max xyz
  uint32_t *d_level_nodes = nullptr, *d_level_start = nullptr;
  float *d_geo_box = nullptr, *d_inst_box = nullptr, *d_node_box = nullptr, *d_matrices = nullptr;
  uint32_t* d_refit_status = nullptr;
  size_t cap_level_nodes = 0, cap_level_start = 0, cap_geo_box = 0, cap_inst_box = 0, cap_node_box = 0, cap_matrices = 0;
  int refit(const float* d_matrices_colmajor, hipStream_t stream);
  int sync_host_masks();
  int sync_host_tlas();
};

}  // namespace mrt
