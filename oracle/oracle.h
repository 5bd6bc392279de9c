/*
 * CPU oracle for the MI355X path-tracing core. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load liboracle.so,
 * and only as the checker or the timed CPU baseline, never as part of the product path.
 *
 * Every function restates the reference's algorithm (johannes-braun/myrt, file:line cited at
 * each definition) in plain single-precision C++ compiled with -ffp-contract=off.
 */
#ifndef MYRT_ORACLE_H
#define MYRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_aabb {
  float min[3];
  float max[3];
  float centroid[3];
  uint32_t has_centroid;
} orc_aabb;

typedef struct orc_node {
  float min[3];
  uint32_t type_and_parent;
  float max[3];
  uint32_t first;
  uint32_t second;
  uint32_t pad[3];
} orc_node;

typedef struct orc_ray {
  float o[3];
  float tmin;
  float d[3];
  float tmax;
} orc_ray;

typedef struct orc_hit {
  float t;
  uint32_t prim;
  uint32_t inst;
  float b1, b2;
} orc_hit;

/* A scene as flat arrays: geometries (concatenated index/position arrays with per-geometry
 * offsets) and instances (geometry id + rows 0-2 of the inverse transform). */
typedef struct orc_scene {
  const uint32_t* idx;        /* concatenated triangle indices (local to each geometry) */
  const float* pos;           /* concatenated xyz positions */
  const uint32_t* geo_idx_off; /* per geometry: first index in idx */
  const uint32_t* geo_tri_cnt; /* per geometry: triangle count */
  const uint32_t* geo_vtx_off; /* per geometry: first vertex in pos (in vertices) */
  uint32_t n_geoms;
  const uint32_t* inst_geom;   /* per instance: geometry */
  const float* inst_inv;       /* per instance: 12 floats, rows 0-2 of M^-1 */
  uint32_t n_inst;
} orc_scene;

/* generate_triangle_bounds (src/pathtracer/bvh.cpp:136-152). */
void orc_triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3, orc_aabb* out);
/* bvh::bvh + bvh::create + bvh::split + bvh::compute_split_axis (src/pathtracer/bvh.cpp:43-116,
 * 154-161, 356-512), literal: one global reference array with insertion and range shifting.
 * Returns 0 on success. Sizes via orc_bvh_sizes after a build into an internal handle. */
typedef struct orc_bvh_s* orc_bvh;
int orc_bvh_build(const orc_aabb* boxes, size_t n, int split_aabbs, orc_bvh* out);
void orc_bvh_sizes(orc_bvh b, size_t* n_nodes, size_t* n_indices);
void orc_bvh_copy(orc_bvh b, orc_node* nodes, uint32_t* indices, orc_aabb* final_boxes);
void orc_bvh_destroy(orc_bvh b);

/* Inverse of a column-major 4x4 (rows 0-2 of the inverse, row-major 3x4). Returns 0 if singular. */
int orc_invert(const float m[16], float out[12]);
/* The instance box of scene::enqueue (src/pathtracer/scene.cpp:256-261). */
void orc_transform_box(const float box_min[3], const float box_max[3], const float m[16], float out_min[3],
                       float out_max[3]);

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013) with the FP64 fallback, in the
 * arithmetic order of the product kernels. Returns 1 on a crossing, with t = T/det. */
int orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                   float* t, float* b1, float* b2);
/* ray_t::intersect(v1, v2, v3) (src/pathtracer/bvh.cpp:514-538; src/pt/intersect.glsl:3-28). */
int orc_moller_trumbore(const float o[3], const float d[3], const float v1[3], const float v2[3],
                        const float v3[3], float* t, float* bx, float* by);

/* Closest hit (lowest (inst, prim) on equal t) or any hit by testing every triangle of every
 * instance. any_hit != 0 writes hits[i].t = 1 for occluded rays and +inf otherwise. */
void orc_trace_brute(const orc_scene* s, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits, int threads);

/* The same query over per-geometry binary BVHs from orc_bvh_build (split_aabbs) and an instance
 * BVH (dont_split): stack traversal with a robust slab test. Same results as brute force. */
typedef struct orc_accel_s* orc_accel;
int orc_accel_build(const orc_scene* s, orc_accel* out);
void orc_accel_trace(orc_accel a, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits, int threads);
void orc_accel_destroy(orc_accel a);

/* The reference's own query, used as the CPU baseline: stackless parent-link traversal with
 * 32-bit bit-stacks (src/pathtracer/bvh.cpp:237-286; src/pt/bvh.impl.glsl:8-72), the slab test
 * of bvh.glsl:31-45, Moeller-Trumbore with eps 1e-23 / border 1e-6, first-found strict
 * t comparisons (src/pt/pt_trace.comp:259-309). hits[i].prim/inst = 0xFFFFFFFF on a miss. */
void orc_reference_trace(orc_accel a, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits);
/* Times orc_reference_trace under a clone of myrt::thread_pool (src/thread_pool.hpp:34-162):
 * `threads` workers (0 = hardware_concurrency), jobs of `chunk` rays submitted with run_async and
 * joined on their futures; one untimed warm-up pass, then the median of `reps` passes.
 * Returns rays per second; *threads_used receives the worker count. */
double orc_reference_time(orc_accel a, const orc_ray* rays, size_t n, int any_hit, int threads, size_t chunk,
                          int reps, int* threads_used);

/* ---- FP64 path-tracing oracle -------------------------------------------------------------- */
typedef struct orc_material {
  uint32_t albedo_rgba8;
  float ior, roughness, metallic, transmission;
  uint32_t texture;
  uint32_t pad[2];
} orc_material;

typedef struct orc_pt_scene {
  const orc_scene* geo;
  const float* nrm;              /* per-vertex normals, same layout as geo->pos */
  const uint32_t* geo_has_nrm;   /* per geometry */
  const uint32_t* inst_material; /* per instance */
  const orc_material* materials;
  uint32_t n_materials;
  const float* lights;           /* 12 floats per quad light: corner, u, v, emission */
  uint32_t n_lights;
  double env_scale;
  const void* sdf_tape;          /* orc_sdf_node[], may be NULL */
  const void* sdf_drawables;     /* orc_sdf_drawable[] */
  uint32_t n_sdf;
  int sdf_steps;
  float sdf_eps;
  /* thin-lens camera (the product's generate pass): disabled when dof == 0 */
  int dof;
  float lens[2];
  float focus;
  const float* bokeh;            /* rgba floats, bokeh_w x bokeh_h, or NULL */
  int bokeh_w, bokeh_h;
  /* albedo textures and environment cube map */
  const float* uv;               /* per-vertex uv, same layout as nrm (2 floats), or NULL */
  const float* texels;           /* rgba floats of all textures */
  const uint32_t* tex_info;      /* per texture: offset (texels), w, h, 0 */
  uint32_t n_textures;
  const float* cubemap;          /* 6 faces of cube_size^2 rgba floats, or NULL */
  int cube_size;
} orc_pt_scene;

/* The product's wavefront estimator in binary64: camera (inverse(proj * view), with the
 * homogeneous divide), Fresnel-weighted GGX / Lambert / smooth-dielectric lobes, quad-light NEE
 * with power-heuristic MIS, Russian roulette from bounce 3, sky environment; the sampler is the
 * product's counter hash. out_rgb receives per-pixel SUMS over spp samples (w*h*3 doubles).
 * Traces through the oracle BVH `a` built over s->geo. Non-finite values are rejected as the
 * product rejects them (a non-finite geometric normal makes the hit a miss, a non-finite shading
 * normal falls back to the geometric one, non-finite contributions are dropped and a non-finite
 * throughput ends the path). Returns the number of rejections (>= 0), or -1 on failure. */
int orc_render(const orc_pt_scene* s, orc_accel a, const float view_colmajor[16], const float proj_colmajor[16],
               int width, int height, int spp, int sample_base, int max_depth, double* out_rgb, int threads);
/* CPU baseline of the reference's plain renderer (sequential_pathtracer.cpp:45-64 with
 * pt_generate/pt_trace/pt_color/pt_ray_filter and pbr.glsl restated, the reference's two sphere
 * lights, its stackless traversal) on rows [0, rows) of a width x height image at 1 spp with
 * max_depth bounces, each pass in 64-row jobs on a myrt::thread_pool clone of `threads` workers
 * (0 = hardware_concurrency). One untimed frame, then the median of `reps`. Returns samples/s;
 * *rays_per_s gets the median rays (closest + shadow) per second, *checksum the mean radiance
 * sum of a frame. */
double orc_reference_render_time(const orc_pt_scene* s, orc_accel a, const float view_colmajor[16],
                                 const float proj_colmajor[16], int width, int height, int rows, int max_depth,
                                 int threads, int reps, int* threads_used, double* rays_per_s, double* checksum);

/* The estimator's material model alone (n = shading normal on wo's side, all vectors unit),
 * batched: orc_bsdf_eval gives, for `count` directions wi (3 doubles each), the non-delta BSDF
 * value f (rgb) and the solid-angle density of the non-delta lobes as the renderer samples them;
 * orc_bsdf_sample draws one direction per uniform triple u (ul, u1, u2) as the renderer does, with
 * side 0 (path ends), 1 (reflected) or 2 (refracted), the throughput weight, the density (0 for
 * the delta transmission lobe) and a delta flag. */
int orc_bsdf_eval(const orc_material* m, const double n[3], const double wo[3], const double* wi, size_t count,
                  double* f, double* pdf);
int orc_bsdf_sample(const orc_material* m, const double n[3], const double wo[3], int entering, const double* u,
                    size_t count, double* wi, double* weight, double* pdf, int32_t* side, int32_t* is_delta);

/* A CPU replica of the product traversal kernel over the product's device-format arrays
 * (mrt_scene_copy_wide: 64-B wide nodes, 48-B triangle records, TLAS nodes and leaf slots, 64-B
 * instance records): per ray, counts[3i..3i+2] = wide nodes, triangle records and instance
 * records the kernel fetches; hits (may be NULL) get the closest hit (t, prim, inst) or, for
 * any_hit, t = 0 when occluded. single_instance = 0xFFFFFFFF unless the frame has one instance. */
int orc_fetch_counts(const void* blas_nodes, const void* tris, const void* tlas_nodes, const uint32_t* tlas_items,
                     const void* instances, int has_tlas, uint32_t single_instance, const orc_ray* rays, size_t n,
                     int any_hit, uint32_t* counts, orc_hit* hits, int threads);

/* ---- SDF sphere tracing -------------------------------------------------------------------- */
typedef struct orc_sdf_node {
  uint32_t op;
  uint32_t material;
  float p[6];
} orc_sdf_node;

typedef struct orc_sdf_drawable {
  float inv[12];
  float box_lo[3];
  uint32_t tape_offset;
  float box_hi[3];
  uint32_t tape_len;
} orc_sdf_drawable;

/* map(p) of a tape (the formulas of src/pathtracer/sdf.hpp:399-633 in postfix form), binary32. */
float orc_sdf_map(const orc_sdf_node* tape, uint32_t len, const float p[3], uint32_t* material);
/* The same with a ray seed for the stochastic material of smooth operators (raymarch.glsl:43-56),
 * and the seed of a world ray (what the closest-hit march uses at its hit point). */
float orc_sdf_map_seeded(const orc_sdf_node* tape, uint32_t len, const float p[3], uint32_t seed, uint32_t* material);
uint32_t orc_sdf_blend_seed(const float o[3], const float d[3]);
/* march_sdf over every drawable (src/pt/raymarch.glsl:116-171), starting at the drawable box's
 * entry, on top of prior hits: closest (hits[i] from a triangle query; a closer SDF surface
 * replaces it, inst = 0x80000000 | drawable, prim = material, normals[3i..3i+2] = world normal
 * when `normals` is given) or any (hits[i].t finite = occluded). */
void orc_sdf_trace(const orc_sdf_node* tape, const orc_sdf_drawable* drawables, uint32_t n_drawables,
                   const orc_ray* rays, size_t n, int any_hit, int steps, float eps, orc_hit* hits, float* normals,
                   int threads);

#ifdef __cplusplus
}
#endif
#endif
