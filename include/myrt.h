/*
 * myrt.h - C ABI of the MI355X path-tracing core (libmyrt_amd.so).
 *
 * Every entry point is extern "C", returns an int status (MRT_OK = 0, negative = error) and
 * reports the reason through mrt_last_error(). Handles are opaque. Buffers are plain pointers
 * plus element counts; nothing here depends on torch or on C++ types.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * johannes-braun/myrt). Threading: a handle may be used by one host thread at a time.
 */
#ifndef MYRT_H
#define MYRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
enum {
  MRT_OK = 0,
  MRT_ERR_INVALID = -1,      /* bad argument or handle */
  MRT_ERR_HIP = -2,          /* a HIP runtime call failed */
  MRT_ERR_NOT_PREPARED = -3, /* trace before mrt_scene_prepare */
  MRT_ERR_LIMIT = -4,        /* a documented size limit was exceeded */
  MRT_ERR_INTERNAL = -5
};

/* BVH build modes: the reference's primitive_split_func values (src/pathtracer/bvh.hpp:159-160). */
enum { MRT_SPLIT_NONE = 0 /* dont_split */, MRT_SPLIT_AABBS = 1 /* split_aabbs */ };

/* ---- plain records ---------------------------------------------------------------------- */

/* Ray record, 32 B: origin, t_min, direction (not necessarily normalised), t_max.
 * Replaces ray_t {origin, direction, length} (src/pathtracer/bvh.hpp:88-96); a hit is accepted
 * for t_min < t < t_max. */
typedef struct mrt_ray {
  float o[3];
  float tmin;
  float d[3];
  float tmax;
} mrt_ray;

/* Closest-hit record, 16 B: distance, triangle index within its geometry, instance index and
 * barycentrics (b1 weights vertex 1, b2 vertex 2; each stored as unorm16, b1 in the low half).
 * A miss has t = +inf and prim = inst = 0xFFFFFFFF. Replaces scene::hit
 * (src/pathtracer/scene.hpp:160-165) and hit_t (src/pt/pt_trace.comp:56-66). */
typedef struct mrt_hit {
  float t;
  uint32_t prim;
  uint32_t inst;
  uint32_t bary;
} mrt_hit;

/* Binary BVH node, 48 B, identical layout to aligned_node_t / bvh_node_t
 * (src/pathtracer/bvh.hpp:98-139): bit 31 of type_and_parent marks a leaf, the low 31 bits hold
 * the parent index. Leaves reference [first, second] (inclusive) of the reordered index array;
 * inner nodes hold their two child node indices. */
typedef struct mrt_bvh_node {
  float min[3];
  uint32_t type_and_parent;
  float max[3];
  uint32_t first;
  uint32_t second;
  uint32_t pad[3];
} mrt_bvh_node;

/* Axis-aligned box with optional weighted centroid: aabb_t (src/pathtracer/bvh.hpp:25-86). */
typedef struct mrt_aabb {
  float min[3];
  float max[3];
  float centroid[3];
  uint32_t has_centroid;
} mrt_aabb;

/* Material record, 32 B: the reference's wavefront material {albedo: vec4unorm8, ior, roughness,
 * metallic, transmission, albedo_texture} (src/main.cpp:454-464, layout rules src/types.hpp:216-280);
 * albedo is RGBA8 with red in the low byte (packUnorm4x8), texture 0 = none. */
typedef struct mrt_material {
  uint32_t albedo_rgba8;
  float ior;
  float roughness;
  float metallic;
  float transmission;
  uint32_t texture;
  uint32_t pad[2];
} mrt_material;

/* One-sided quad area light: corner + a*u + b*v (a, b in [0,1]) emitting `emission` (linear RGB
 * radiance) towards normalize(cross(u, v)). Replaces the two hard-coded sphere lights of
 * src/pt/pt_trace.comp:141-146. */
typedef struct mrt_quad_light {
  float corner[3];
  float u[3];
  float v[3];
  float emission[3];
} mrt_quad_light;

/* One instruction of an SDF program, 32 B. A program is the postfix form of the reference's
 * sdf:: tree (src/pathtracer/sdf.hpp:24-67, 399-633): primitives push a distance, operators
 * combine the top two (lhs pushed first), PUSH_TRANSLATE / PUSH_MIRROR transform the position
 * for the single operand that follows, up to the matching POP. */
typedef struct mrt_sdf_node {
  uint32_t op;
  uint32_t material;  /* primitives: material index of the surface */
  float p[6];
} mrt_sdf_node;
enum {
  MRT_SDF_SPHERE = 1,        /* p0 radius */
  MRT_SDF_BOX = 2,           /* p0-p2 half extents */
  MRT_SDF_TORUS = 3,         /* p0 large radius, p1 small radius */
  MRT_SDF_CAPSULE = 4,       /* vertical capsule: p0 height, p1 radius */
  MRT_SDF_MENGER = 5,        /* unit Menger sponge */
  MRT_SDF_UNION = 16,
  MRT_SDF_SUBTRACT = 17,     /* max(-lhs, rhs) */
  MRT_SDF_INTERSECT = 18,
  MRT_SDF_SMOOTH_UNION = 19, /* p0 k */
  MRT_SDF_SMOOTH_SUBTRACT = 20,
  MRT_SDF_SMOOTH_INTERSECT = 21,
  MRT_SDF_PUSH_TRANSLATE = 32, /* p0-p2 offset */
  MRT_SDF_PUSH_MIRROR = 33,    /* p0 axis (0, 1, 2) */
  MRT_SDF_POP = 34
};

typedef struct mrt_render_stats {
  uint64_t samples;      /* pixel samples rendered since creation */
  double milliseconds;   /* wall time inside mrt_render */
  uint64_t rays;         /* closest-hit plus shadow rays traced since creation */
  uint64_t nonfinite;    /* non-finite values rejected since creation (pt_color.comp:84-98, 156-166;
                            pt_trace.comp:186-189): a non-finite geometric normal makes the hit a
                            miss, a non-finite shading normal falls back to the geometric one,
                            non-finite contributions are dropped and a non-finite throughput ends
                            the path */
} mrt_render_stats;

/* Radiance-cache query, 64 B: shading point, outgoing direction (towards the viewer), shading
 * normal, GGX roughness and albedo. */
typedef struct mrt_cache_query {
  float pos[3];
  float roughness;
  float dir[3];
  float pad0;
  float nrm[3];
  float pad1;
  float albedo[3];
  float pad2;
} mrt_cache_query;

typedef struct mrt_scene_s* mrt_scene;
typedef struct mrt_bvh_s* mrt_bvh;
typedef struct mrt_renderer_s* mrt_renderer;
typedef struct mrt_cache_s* mrt_cache;
typedef struct mrt_denoiser_s* mrt_denoiser;
typedef struct mrt_obj_s* mrt_obj;

/* One triangulated OBJ group with the MTL fields of its material (obj.hpp material_t). */
typedef struct mrt_obj_group_info {
  uint64_t n_idx;
  uint64_t n_vtx;
  float diffuse[3];  /* Kd */
  float specularity; /* Ns */
  float ior;         /* Ni */
  float dissolve;    /* d */
  char name[128];    /* "object/group", truncated */
  char material_name[64];
  char albedo_texture[256];
} mrt_obj_group_info;

/* Counters filled by mrt_scene_prepare (scene::prepare_result_t, src/pathtracer/scene.hpp:84-91,
 * extended with the device-side sizes of the wide BVH). */
typedef struct mrt_prepare_result {
  int32_t geometries_changed;
  int32_t drawables_changed;
  uint32_t num_geometries;
  uint32_t num_instances;
  uint32_t blas_wide_nodes;
  uint32_t blas_tri_records;
  uint32_t tlas_wide_nodes;
  uint32_t max_stack_entries;
  int32_t tlas_refitted; /* the frame moved instances only and the TLAS was refitted on the GPU */
  uint32_t reserved;
  uint64_t device_bytes;
} mrt_prepare_result;

/* Per-launch traversal counters (sum over the rays of one trace call). */
typedef struct mrt_trace_stats {
  uint64_t rays;
  uint64_t node_visits; /* wide 64-B node fetches */
  uint64_t tri_tests;   /* 48-B triangle record fetches */
  uint64_t instance_visits;
} mrt_trace_stats;

/* ---- errors, device plumbing ------------------------------------------------------------- */
const char* mrt_last_error(void);
int mrt_device_info(char* name_out, size_t name_cap, int* num_cus);
int mrt_device_alloc(size_t bytes, void** out);
int mrt_device_free(void* ptr);
int mrt_memcpy_h2d(void* dst, const void* src, size_t bytes);
int mrt_memcpy_d2h(void* dst, const void* src, size_t bytes);
int mrt_device_synchronize(void);
/* A plain streaming copy kernel (bytes a multiple of 16) on `stream`; reads and writes exactly
 * `bytes` once, so it calibrates hardware byte counters. No reference counterpart. */
int mrt_device_copy(void* dst, const void* src, size_t bytes, void* stream);
/* Stream-ordered GPU timer: records start/stop events on `stream` (NULL = default stream). */
typedef struct mrt_timer_s* mrt_timer;
int mrt_timer_create(mrt_timer* out);
int mrt_timer_destroy(mrt_timer t);
int mrt_timer_start(mrt_timer t, void* stream);
int mrt_timer_stop(mrt_timer t, void* stream);
int mrt_timer_elapsed_ms(mrt_timer t, float* ms); /* synchronises on the stop event */

/* ---- BVH builder (host) ------------------------------------------------------------------ */
/* generate_triangle_bounds (src/pathtracer/bvh.cpp:136-152): one padded box per triangle with
 * the weighted centroid (p0+p1+p2)/3. */
int mrt_generate_triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3, size_t n_vtx,
                                 mrt_aabb* out_boxes);
/* bvh::bvh(aabbs, split) (src/pathtracer/bvh.cpp:154-161, 356-512): binned-SAH build, bit-exact
 * with the reference's node order and reordered index array. */
int mrt_bvh_build(const mrt_aabb* boxes, size_t n, int split_mode, mrt_bvh* out);
int mrt_bvh_build_triangles(const uint32_t* idx, size_t n_idx, const float* pos3, size_t n_vtx, int split_mode,
                            mrt_bvh* out);
int mrt_bvh_sizes(mrt_bvh b, size_t* n_nodes, size_t* n_indices);
/* bvh::nodes() and bvh::reordered_indices() (src/pathtracer/bvh.hpp:174-175). */
int mrt_bvh_copy(mrt_bvh b, mrt_bvh_node* nodes, uint32_t* indices);
int mrt_bvh_destroy(mrt_bvh b);

/* ---- scene ------------------------------------------------------------------------------- */
/* scene::scene / ~scene (src/pathtracer/scene.hpp:116-117). */
int mrt_scene_create(mrt_scene* out);
int mrt_scene_destroy(mrt_scene s);
/* scene::push_geometry (src/pathtracer/scene.hpp:128-129, scene.cpp:43-71). Copies its inputs.
 * nrm3 and uv2 may be NULL. Builds the object BVH with split_aabbs, as the reference does. */
int mrt_scene_push_geometry(mrt_scene s, const uint32_t* idx, size_t n_idx, const float* pos3, const float* nrm3,
                            const float* uv2, size_t n_vtx, uint32_t* out_geom);
/* scene::enqueue(geometry, material, mat4) (src/pathtracer/scene.hpp:144-145, scene.cpp:255-271).
 * m4x4 is column-major like rnu::mat4. out_inst receives the instance index used in mrt_hit.inst
 * after the next prepare (the enqueue order of this frame). */
int mrt_scene_enqueue_geometry(mrt_scene s, uint32_t geom, uint32_t material, const float m4x4_colmajor[16],
                               uint32_t* out_inst);
/* scene::prepare (src/pathtracer/scene.hpp:182, scene.cpp:330-409): builds the instance BVH with
 * dont_split over this frame's drawables, collapses every BVH to the 64-B wide format and uploads
 * it. Drawables are consumed, as in the reference; an empty frame keeps the previous one. */
int mrt_scene_prepare(mrt_scene s, mrt_prepare_result* out);
/* Refit instead of rebuild (no reference counterpart; the reference rebuilds its instance BVH on
 * every drawable-hash change, src/pathtracer/scene.cpp:255-271, 336-344). With max_refits > 0 a
 * prepare whose drawables differ from the prepared frame only in their matrices (same count, same
 * geometry and material at each position) keeps the TLAS topology and refits its boxes on the GPU;
 * after max_refits refits in a row the next change rebuilds. 0 (the default) rebuilds every time.
 * Closest and any hits do not depend on the choice; traversal cost may grow as instances move. */
int mrt_scene_set_tlas_refit(mrt_scene s, uint32_t max_refits);
/* Moves the prepared frame's instances to n device-resident column-major 4x4 matrices (n = the
 * prepared instance count, in enqueue order) and refits the TLAS, all on `stream` (a hipStream_t;
 * NULL = the null stream) with no host synchronisation: for transforms produced on the GPU. The
 * inverses and world boxes are computed on the device exactly as mrt_scene_enqueue_geometry and
 * mrt_scene_prepare compute them on the host. A singular matrix makes its instance unhittable and
 * sets bit 0 of mrt_scene_update_status. The next prepare re-evaluates the enqueued drawables. */
int mrt_scene_update_transforms_device(mrt_scene s, const float* d_m4x4_colmajor, size_t n, void* stream);
/* Synchronises and reports the last refit's status bits (bit 0: a singular matrix). */
int mrt_scene_update_status(mrt_scene s, uint32_t* status);
/* The instance inverse matrices of the prepared frame (row-major 3x4 floats, 12 per instance). */
int mrt_scene_instance_inverses(mrt_scene s, float* out, size_t n_inst);
/* The host half of mrt_scene_prepare (BVH builds and the wide collapse) without touching the GPU,
 * and a copy of the resulting device-format arrays: 64-B wide nodes, 48-B triangle records,
 * TLAS leaf slots and 64-B instance records. Sizes come from mrt_prepare_result. These replace
 * scene::get_scene_buffers (src/pathtracer/scene.hpp:184-186), which exposes the GL buffers. */
int mrt_scene_build_host(mrt_scene s, mrt_prepare_result* out);
int mrt_scene_copy_wide(mrt_scene s, void* blas_nodes, void* tris, void* tlas_nodes, uint32_t* tlas_items,
                        void* instances);

/* Albedo texture (the bindless albedo_texture of the material layout, main.cpp:454-464): w x h
 * rgba floats, row 0 at v = 0. Returns its id (>= 1) for mrt_material.texture; a textured hit
 * takes albedo = texture(u, 1 - v)^2.2 with bilinear filtering and wrap (material_new.glsl:29-31). */
int mrt_scene_push_texture(mrt_scene s, const float* rgba, int w, int h, uint32_t* out_id);
/* The reference application's material (src/main.cpp:447-464, pushed through
 * scene::push_material, src/pathtracer/scene.hpp:134-140) as one record: a material of the
 * built-in 7-block layout {albedo: color_t{value: vec4unorm8}, ior, roughness, metallic,
 * transmission, albedo_texture: uvec2} with texture = the uvec2's low word. The material index is
 * what mrt_scene_enqueue_geometry and SDF programs take; materials of every type share it. */
int mrt_scene_push_material(mrt_scene s, const mrt_material* mat, uint32_t* out_id);
/* Writes a record through the material's type bindings (below); an input its type does not
 * bind must keep its default (albedo 0xFFFFFFFF, ior 1.5, roughness 1, others 0). */
int mrt_scene_set_material(mrt_scene s, uint32_t id, const mrt_material* mat);
/* The record the shaders read for a material of any type (scene::get_material_buffer,
 * src/pathtracer/scene.hpp:178, resolved through the type's bindings). */
int mrt_scene_get_material(mrt_scene s, uint32_t id, mrt_material* out);
/* set_parameter(material, name, value) with the value as floats: a vec4unorm8 parameter (or a
 * layout wrapping one, like color_t) takes 3 or 4 values in [0, 1] packed to RGBA8 (alpha 1 if
 * absent); integer types take integral values; a parameter feeding the texture input also takes
 * the texture id alone and must name a pushed texture; float types take exactly their component
 * count (parameter.cpp:40-41). An unknown name is an error here. */
int mrt_scene_set_material_param(mrt_scene s, uint32_t id, const char* name, const float* value, size_t n);

/* ---- named-parameter material layouts ---------------------------------------------------
 * object_type / types_registry (src/types.hpp:216-310): a layout is a named list of parameters,
 * each a constant type ("float", "int", "uint", "vec2".."vec4", "vec4unorm8", "ivec2".."ivec4",
 * "uvec2".."uvec4", "mat2".."mat4"; 1, 1, 1, 2-4, 1, 2-4, 2-4, 4/9/16 blocks) or an earlier layout,
 * packed as consecutive 4-byte blocks in declaration order (object_type::add_parameter,
 * src/types.hpp:250-258). A layout can no longer grow once a material type or another layout
 * uses it. Layout and material type ids are per scene. */
/* object_type(name) registered with types_registry::push_back (src/types.hpp:216-224, 284-290). */
int mrt_scene_register_layout(mrt_scene s, const char* name, uint32_t* out_layout);
/* object_type::add_parameter(name, type) (src/types.hpp:250-258) with the type found by name as
 * types_registry::find does (src/types.hpp:292-299): the first layout of that name, else the
 * constant type of that identifier. */
int mrt_scene_layout_add_parameter(mrt_scene s, uint32_t layout, const char* name, const char* type_name);
/* types_registry::find: *constant = the constant type's parameter_type_id (src/types.hpp:27-44),
 * or -1 and *layout = the layout id. Unknown names are an error. */
int mrt_scene_find_type(mrt_scene s, const char* name, int32_t* constant, uint32_t* layout);
/* object_type::blocks() and parameters() (src/types.hpp:240-262). */
int mrt_scene_layout_info(mrt_scene s, uint32_t layout, uint32_t* blocks, uint32_t* n_params);
int mrt_scene_layout_parameter(mrt_scene s, uint32_t layout, uint32_t index, char* name, size_t name_cap,
                               uint32_t* offset, uint32_t* blocks, int32_t* constant, uint32_t* inner_layout);
/* scene::push_material_type(type, glsl) (src/pathtracer/scene.hpp:131-133, material.hpp:18-37).
 * Every material is shaded by the GGX model (the reference's material_new.glsl); in place of the
 * GLSL `open()` that loads its inputs, input_paths names, for each input in the order albedo, ior,
 * roughness, metallic, transmission, albedo_texture, the parameter path that feeds it ("albedo"
 * or "albedo.value"; a layout of one parameter counts as that parameter). A NULL array, or a NULL
 * entry, binds the input to the parameter of the input's name if the layout has one; "" leaves
 * it unbound (its default). Albedo takes vec4unorm8, albedo_texture uint or uvec2 (low word = a
 * texture id, a nonzero high word = no texture), the others float; a mismatch is an error. */
int mrt_scene_push_material_type(mrt_scene s, uint32_t layout, const char* const* input_paths, uint32_t* out_type);
/* scene::push_material(material_type) (src/pathtracer/scene.hpp:134-140, material.hpp:81-91):
 * the material's blocks start at zero. */
int mrt_scene_push_material_of_type(mrt_scene s, uint32_t type, uint32_t* out_id);
/* scene::set_parameter<T>(material, name, value) (src/pathtracer/scene.hpp:156, 249-252;
 * material.hpp:93-102): writes `bytes` bytes, the value's object representation (a vec4unorm8 is
 * its 4 bytes r, g, b, a, parameter.cpp:130-133; a uvec2 its two words), at the parameter's
 * offset. The size must be the parameter's blocks x 4, else MRT_ERR_INVALID ("Data size not
 * suitable for this parameter"). A name the layout lacks is ignored, as in the reference, and
 * *found (if not NULL) says which happened. */
int mrt_scene_set_parameter(mrt_scene s, uint32_t material, const char* name, const void* data, size_t bytes,
                            int* found);
/* material_t {index, offset, type_info} (src/material.hpp:74-79). */
int mrt_scene_material_slot(mrt_scene s, uint32_t material, uint32_t* type, uint32_t* offset);
/* material_buffer::buffer() (src/material.hpp:104-106) as raw 4-byte blocks: copies
 * min(n, total) blocks and reports the total. */
int mrt_scene_material_blocks(mrt_scene s, uint32_t* out, size_t n, size_t* total);

/* scene::push_sdf(sdf_info_t{root}) (src/pathtracer/scene.cpp:189-204) with the tree as a postfix
 * program; the program is validated (stack discipline, depth <= 8, modifiers <= 4 deep). */
int mrt_scene_push_sdf(mrt_scene s, const mrt_sdf_node* nodes, size_t n, uint32_t* out_sdf);
/* scene::set_parameter(sdf, parameter, value) (src/pathtracer/scene.hpp:150-153; sdf.cpp:346-407,
 * parameter.cpp:57-109): writes count floats into p[slot .. slot + count) of node `node` of a
 * pushed program (its tape index, as pushed) without re-pushing it. The program's bounds and
 * compiled form are updated at once; a program's culling grid is rebuilt on the GPU at the next
 * mrt_scene_prepare, which is also when the change reaches the device (as in the reference, whose
 * prepare uploads the parameter buffer). Values must be finite; an update whose bounds come out
 * invalid is rejected and changes nothing. */
int mrt_scene_set_sdf_param(mrt_scene s, uint32_t sdf, uint32_t node, uint32_t slot, const float* values,
                            size_t count);
/* A program's culling grid: dim cells per axis over its object-space bounds (0 = none, every
 * instruction is evaluated), origin lo and scale inv (cells per unit), and, if masks is not NULL,
 * its dim^3 cell masks (bit j: chain instruction j may change the distance in the cell).
 * Synchronises with the device when the masks were rebuilt there. */
int mrt_scene_copy_sdf_grid(mrt_scene s, uint32_t sdf, float lo[3], float inv[3], uint32_t* dim, uint64_t* masks,
                            size_t n);
/* scene::enqueue(sdf, mat4) (src/pathtracer/scene.hpp:147-148, scene.cpp:411-417). SDF hits come
 * back with inst = 0x80000000 | drawable index, prim = material and an octahedral normal in bary. */
int mrt_scene_enqueue_sdf(mrt_scene s, uint32_t sdf, const float m4x4_colmajor[16], uint32_t* out_drawable);
/* The prepared SDF drawables, 80 B each: rows 0-2 of M^-1, world box (lo, tape offset, hi, tape
 * length). Like mrt_scene_copy_wide, it exposes the device records (scene::get_scene_buffers). */
int mrt_scene_sdf_drawables(mrt_scene s, void* out, size_t n);
/* set_sdf_marching_steps / set_sdf_marching_epsilon (src/pathtracer/sequential_pathtracer.hpp:44-45). */
int mrt_scene_set_sdf_marching(mrt_scene s, int steps, float eps);
/* Lights and the sky-gradient environment scale (src/pt/pt_color.comp:189-201). */
int mrt_scene_set_lights(mrt_scene s, const mrt_quad_light* lights, size_t n);
int mrt_scene_set_environment(mrt_scene s, float scale);

/* ---- Wavefront OBJ/MTL ingest (src/obj.hpp, src/obj.cpp) ----------------------------------- */
/* myrt::obj::load_obj followed by triangulate for every object (src/obj.cpp:96-361): one group
 * per `g` / `usemtl` run, named "object/group", fan-triangulated with vertices shared per
 * distinct (position, normal, uv). Host only. An index outside its list is an error. */
int mrt_obj_load(const char* path, mrt_obj* out);
int mrt_obj_destroy(mrt_obj o);
int mrt_obj_num_groups(mrt_obj o, size_t* n);
/* A group's sizes, name and MTL entry (material_t, src/obj.hpp); albedo_texture is map_Kd
 * joined to the OBJ's folder, "" when there is none. */
int mrt_obj_group(mrt_obj o, size_t i, mrt_obj_group_info* info);
/* Copies a group's triangle indices (n_idx) and n_vtx positions, normals and uvs; any pointer
 * may be NULL. */
int mrt_obj_group_copy(mrt_obj o, size_t i, uint32_t* idx, float* pos3, float* nrm3, float* uv2);
/* load_object_file (src/main.cpp:902-949): pushes every non-empty group as a geometry and a
 * material {albedo = unorm8(Kd, 1), ior = Ni, roughness = max(0.054, (1/Ns)^(1/pi)),
 * transmission = 1 - d}. n = the number of groups; an empty group's ids are 0xFFFFFFFF. The
 * caller loads map_Kd images and sets the materials' albedo_texture, and enqueues the pairs
 * with its import scale, as the reference's caller does. */
int mrt_scene_push_obj(mrt_scene s, mrt_obj o, uint32_t* geometries, uint32_t* materials, size_t n);

/* ---- ray queries (GPU) ------------------------------------------------------------------- */
/* nearest_hit / any_hit (src/pt/pt_trace.comp:311-386) over the two-level BVH
 * (src/pt/bvh.impl.glsl:8-72), with a watertight triangle test and lowest-(inst, prim) ties. */
int mrt_trace_closest(mrt_scene s, const mrt_ray* rays, size_t n, mrt_hit* hits);      /* host buffers */
int mrt_trace_any(mrt_scene s, const mrt_ray* rays, size_t n, uint8_t* occluded);      /* host buffers */
int mrt_trace_closest_device(mrt_scene s, const mrt_ray* d_rays, size_t n, mrt_hit* d_hits, void* stream);
int mrt_trace_any_device(mrt_scene s, const mrt_ray* d_rays, size_t n, uint8_t* d_occluded, void* stream);
/* mrt_trace_closest_device without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same kernels and results; their scratch (counters, traversal
 * spill stack, SDF hit points) lives in one 32 MiB device array per GPU that the library holds
 * until the process ends, and the persistent grid shrinks to the spill stack that fits (deep
 * trees then run on fewer waves). The scene must be prepared. Limits: n < 2^31, and n <= 2^20
 * when the scene has SDF drawables (MRT_ERR_LIMIT beyond). Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (two streams or threads) race; serialise them. */
int mrt_trace_closest_device_noalloc(mrt_scene s, const mrt_ray* d_rays, size_t n, mrt_hit* d_hits, void* stream);
/* Same traversal with per-launch fetch counters (slower; used to price bytes per ray). */
int mrt_trace_stats_device(mrt_scene s, const mrt_ray* d_rays, size_t n, int any_hit, mrt_trace_stats* out,
                           void* stream);
/* scene::pick (src/pathtracer/scene.cpp:273-314): closest hit of one ray; returns 1 in *hit_found
 * when something was hit. */
int mrt_scene_pick(mrt_scene s, const mrt_ray* ray, mrt_hit* hit, int* hit_found);

/* ---- wavefront renderer (sequential_pathtracer, src/pathtracer/sequential_pathtracer.hpp:26-58) */
int mrt_renderer_create(mrt_renderer* out);
int mrt_renderer_destroy(mrt_renderer r);
/* set_view_matrix + set_projection_matrix (column-major like rnu::mat4); restarts accumulation. */
int mrt_renderer_set_camera(mrt_renderer r, const float view_colmajor[16], const float proj_colmajor[16]);
/* set_num_bounces (default 10, sequential_pathtracer.hpp:134). */
int mrt_renderer_set_num_bounces(mrt_renderer r, int n);
/* Shade each bounce's paths in material order (a counting sort of the live queue by the hit's
 * material, pt_ray_filter.comp:49-63 leaves them in atomic order). Results do not change beyond
 * fp32 summation order; the default is off (MRT_MATERIAL_SORT=1 at creation switches it on). */
int mrt_renderer_set_material_sort(mrt_renderer r, int enabled);
/* set_dof_enabled / set_lens_size / set_focus / set_bokeh_mask (sequential_pathtracer.hpp:30-33;
 * defaults off, 10 x 10, 10). A thin lens around each pinhole ray: lens offset o uniform in the
 * disk of radius 0.49, focal point at `focus` along the pinhole ray, new direction
 * normalize(o * lens / 2, focus) in a frame around the pinhole direction (pt_generate.comp:38-52).
 * The mask (rgba floats, w x h, row 0 at v = 0; w < 10 means none, bokeh.glsl:24-25) is sampled
 * bilinearly at 0.5 + o; black offsets are redrawn up to 8 times and the mask colour tints the
 * path. Each setter restarts accumulation when the value changes. */
int mrt_renderer_set_dof_enabled(mrt_renderer r, int enabled);
int mrt_renderer_set_lens_size(mrt_renderer r, float w, float h);
int mrt_renderer_set_focus(mrt_renderer r, float focus);
int mrt_renderer_set_bokeh_mask(mrt_renderer r, const float* rgba, int w, int h);
/* set_cubemap (sequential_pathtracer.hpp:34): environment for rays that leave the scene, six
 * faces (+x, -x, +y, -y, +z, -z; OpenGL face orientation) of size x size rgba floats, bilinear
 * within a face and scaled by mrt_scene_set_environment; size <= 10 restores the sky gradient
 * (pt_color.comp:191-199). */
int mrt_renderer_set_cubemap(mrt_renderer r, const float* faces, int size);
/* Attaches a radiance cache (no reference counterpart). mode 0: off; 1: train only (every pass
 * builds a training batch of up to 2^18 path-vertex samples from its training paths, one drawn
 * from every round(1 / train_fraction) consecutive paths of the pixel-major batch, which stay
 * unbiased, and takes `train_steps` Adam steps on disjoint slices of it); 2: train and
 * end the other paths in the cache at vertex `min_bounce` (0 = primary hit) if its roughness is
 * >= 0.25, or at any later vertex; never on dielectrics. The accumulated image is kept; call
 * mrt_renderer_invalidate to restart it. */
int mrt_renderer_set_cache(mrt_renderer r, mrt_cache c, int mode, float train_fraction, int train_steps,
                           int min_bounce);
/* Path length under mode 2 (no reference counterpart). train_end > 0 (default 4): every path,
 * training paths included, ends in the cache at vertex train_end, on dielectrics too (render paths
 * have usually ended at min_bounce already), and for a training path the cache's prediction there
 * is added to the radiance its earlier vertices learn from (self-training, as in Mueller et al.
 * 2021); that vertex itself is not a training sample. An `unbiased_fraction` of the training
 * paths (default 0) is exempt and runs to full depth. train_end = 0 keeps every training path
 * unbiased and lets paths on dielectrics run to full depth. */
int mrt_renderer_set_cache_training(mrt_renderer r, int train_end, float unbiased_fraction);
/* The sampler index of the first sample of the accumulation, 0 by default (no reference
 * counterpart): renderers given disjoint ranges [base, base + spp) render independent samples of
 * one image, so their means combine (say one renderer per GPU, or chunks of a reference image
 * accumulated in binary64 on the host). Restarts the accumulation. */
int mrt_renderer_set_sample_base(mrt_renderer r, uint32_t base);
/* invalidate_counter / sample_count. */
int mrt_renderer_invalidate(mrt_renderer r);
int mrt_renderer_sample_count(mrt_renderer r, uint32_t* out);
/* run(provider, scene, w, h) repeated spp times: prepares the scene, adds spp samples per pixel
 * to the running accumulation and synchronises. Image rows start at the bottom (GL convention). */
int mrt_render(mrt_renderer r, mrt_scene s, int width, int height, int spp);
/* The accumulated mean radiance, width*height RGBA floats (alpha = 1). */
int mrt_renderer_read(mrt_renderer r, float* rgba);
/* The same image into a device buffer of width*height RGBA floats, asynchronously on `stream`
 * (NULL: the renderer's stream, which mrt_render has synchronised). Replaces the GPU texture that
 * run() returns (src/pathtracer/sequential_pathtracer.hpp:26-28): no host copy per frame. */
int mrt_renderer_read_device(mrt_renderer r, float* d_rgba, void* stream);
int mrt_renderer_stats(mrt_renderer r, mrt_render_stats* out);

/* ---- neural radiance cache (no reference counterpart: the north star's learned part) ------
 * A 64-input frequency-encoded MLP with 5 hidden layers of 64 ReLU units and 3 outputs
 * (20,672 weights, no biases), bf16 MFMA with fp32 accumulation, Adam on fp32 master weights.
 * Parameters are packed layer by layer, W[out][in] row-major: 5 x 64x64 then 3x64. */
int mrt_cache_create(mrt_cache* out, uint64_t seed);
int mrt_cache_destroy(mrt_cache c);
/* Box the position encoding maps to [0,1]^3. */
int mrt_cache_set_bounds(mrt_cache c, const float lo[3], const float hi[3]);
int mrt_cache_set_adam(mrt_cache c, float lr, float beta1, float beta2, float eps);
int mrt_cache_get_params(mrt_cache c, float* out, size_t n);
int mrt_cache_set_params(mrt_cache c, const float* in, size_t n);
/* Predicted radiance, 3 floats per query. */
int mrt_cache_infer(mrt_cache c, const mrt_cache_query* q, size_t n, float* out_rgb);
int mrt_cache_infer_device(mrt_cache c, const mrt_cache_query* d_q, size_t n, float* d_out_rgb, void* stream);
/* One Adam step on the relative-L2 loss mean((y - t)^2 / (y^2 + 0.01)) over the batch. */
int mrt_cache_train(mrt_cache c, const mrt_cache_query* q, const float* targets_rgb, size_t n, float* loss);
int mrt_cache_train_device(mrt_cache c, const mrt_cache_query* d_q, const float* d_targets_rgb, size_t n,
                           void* stream);
/* The loss gradient (packed like the parameters) without updating anything. */
int mrt_cache_gradient(mrt_cache c, const mrt_cache_query* q, const float* targets_rgb, size_t n, float* grad,
                       float* loss);

/* ---- denoiser (no reference counterpart: the north star's convolutional denoiser, which the
 * reference's bilateral pp_denoise.comp:24-63 stands in for) ------------------------------------
 * A 5-level U-Net (32/48/64/96/128 channels, two 3x3 conv + ReLU per level and side, 2x2 average
 * pooling, nearest upsampling with skip concatenation, 1x1 output) over 10 raw features per
 * pixel: mean radiance rgb, albedo rgb, normal xyz, hit distance (inf for a miss). It predicts a
 * log-space residual: out = expm1(max(log1p(rgb) + y, 0)). Inference in fp16 MFMA, training in
 * fp32 MFMA, both accumulating in fp32; Adam on fp32 weights. Parameters are packed layer by
 * layer (encoder levels 0-4 conv a, conv b; decoder levels 3-0 conv a, conv b; output), each as
 * W[cout][tap][cin] (tap = 3*ky + kx; cin in [upsampled, skip] order) followed by b[cout]. */
int mrt_denoiser_create(mrt_denoiser* out, uint64_t seed);
int mrt_denoiser_destroy(mrt_denoiser d);
int mrt_denoiser_num_params(size_t* n);
int mrt_denoiser_get_params(mrt_denoiser d, float* out, size_t n);
int mrt_denoiser_set_params(mrt_denoiser d, const float* in, size_t n);
int mrt_denoiser_set_adam(mrt_denoiser d, float lr, float beta1, float beta2, float eps);
/* The residual y (B*H*W*3) of a batch of raw feature images (B*H*W*10); H, W multiples of 16.
 * half = 1 runs the fp16 inference kernels, 0 the fp32 training kernels. */
int mrt_denoiser_forward(mrt_denoiser d, const float* raw, int B, int H, int W, int half, float* y);
/* Loss mean(|log1p(rgb) + y - log1p(target)|) and its gradient (packed like the parameters). */
int mrt_denoiser_gradient(mrt_denoiser d, const float* raw, const float* target_rgb, int B, int H, int W, float* grad,
                          float* loss);
/* One Adam step on a batch (host or device buffers). */
int mrt_denoiser_train(mrt_denoiser d, const float* raw, const float* target_rgb, int B, int H, int W, float* loss);
int mrt_denoiser_train_device(mrt_denoiser d, const float* d_raw, const float* d_target_rgb, int B, int H, int W,
                              float* loss, void* stream);
/* Denoised rgba (H*W*4) of one raw feature image of any size. */
int mrt_denoiser_denoise(mrt_denoiser d, const float* raw, int H, int W, float* rgba);
int mrt_denoiser_denoise_device(mrt_denoiser d, const float* d_raw, int H, int W, float* d_rgba, void* stream);
/* The renderer's per-pixel feature means (H*W*8: albedo rgb, normal xyz, hit distance, 0) of the
 * primary hits, and its accumulated image passed through a denoiser (H*W*4). */
int mrt_renderer_read_aux(mrt_renderer r, float* out);
int mrt_renderer_denoise(mrt_renderer r, mrt_denoiser d, float* rgba);

/* ---- post-process chain (src/pathtracer/post_process.hpp:30-60 and src/pt/pp_*.comp) --------
 * rgba float images, row-major. MRT_POST_BILATERAL is pp_denoise.comp:24-63 (9x9 window, spatial
 * sigma `exponent`, range sigma `strength`, neighbours clamped to [0, 1]; the app's defaults are
 * 10 and 0.05, post_process.hpp:33-34); MRT_POST_TONEMAP_FILMIC is pp_tonemap_filmic.cl:1-5;
 * MRT_POST_FXAA is pp_fxaa.comp:31-105 (bilinear taps clamped to the edge). */
enum { MRT_POST_BILATERAL = 1, MRT_POST_TONEMAP_FILMIC = 2, MRT_POST_FXAA = 3 };
/* One pass on device images (in and out distinct), asynchronous on `stream`. */
int mrt_post_pass_device(int op, const float* d_in, float* d_out, int w, int h, float exponent, float strength,
                         void* stream);
/* A sequence of passes on host images. */
int mrt_post_process(const float* rgba_in, float* rgba_out, int w, int h, const int* ops, int n_ops, float exponent,
                     float strength);

#ifdef __cplusplus
}
#endif

#endif /* MYRT_H */
