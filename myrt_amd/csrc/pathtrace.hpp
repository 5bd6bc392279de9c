// Wavefront path tracer: records and kernel launchers shared by renderer.cpp and pathtrace.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.hpp"
#include "mlp.hpp"

namespace mrt {

// Material record, 32 B: the reference's wavefront material {albedo: vec4unorm8, ior, roughness,
// metallic, transmission, albedo_texture} (src/main.cpp:454-464) with the texture handle kept as a
// plain id (0 = none).
struct Material {
  uint32_t albedo_rgba8;
  float ior;
  float roughness;
  float metallic;
  float transmission;
  uint32_t texture;
  uint32_t pad[2];
};
static_assert(sizeof(Material) == 32, "material record is 32 bytes");

// One-sided quad area light: points corner + a*u + b*v, a, b in [0, 1], emitting Le towards
// normalize(cross(u, v)).
struct QuadLight {
  float corner[3];
  float u[3];
  float v[3];
  float emission[3];
};

// Per-geometry attribute bases into the concatenated vertex/index arrays.
struct GeoInfo {
  uint32_t idx_base;   // first index (3 per triangle)
  uint32_t vtx_base;   // first vertex
  uint32_t has_normals;
  uint32_t pad;
};

// Path state, 32 B per live path (ray in a separate 32-B mrt_ray array).
struct PathState {
  float thr[3];
  uint32_t pixel;
  uint32_t sample;
  uint32_t flags;     // bits 0-7: bounce; bit 8: previous vertex was the camera or a delta lobe
  float last_pdf;     // solid-angle pdf of the previous non-delta bounce (for MIS)
  uint32_t pad;
};
static_assert(sizeof(PathState) == 32, "path state is 32 bytes");

struct ShadowRecord {
  float contrib[3];
  uint32_t pixel;
  uint32_t train_slot;  // training path of the contribution, or 0xFFFFFFFF
  uint32_t pad[3];
};

// Radiance-cache training: one header per training path and up to max_vertices recorded
// vertices; the suffix radiance of vertex k is (T_final - prefix_k) / thr_k.
struct TrainPath {
  float T[3];   // radiance the path has contributed so far
  uint32_t nv;  // recorded vertices
};
struct TrainVertex {
  CacheQuery q;
  float thr[3];
  float prefix[3];
  float pad[2];
};
static_assert(sizeof(TrainVertex) == 96, "training vertex is 96 bytes");
struct CacheRecord {
  float thr[3];
  uint32_t pixel;
  uint32_t train_slot;  // training path that ended here (its tail radiance), or 0xFFFFFFFF
  uint32_t pad[3];
};
static_assert(sizeof(CacheRecord) == 32, "cache record is 32 bytes");

struct PtCache {
  uint32_t terminate;       // non-training paths end in the cache
  uint32_t train;           // training paths record vertices
  uint32_t train_group;     // one training path per train_group consecutive paths
  uint32_t max_train_paths;
  uint32_t max_vertices;
  uint32_t max_queries;
  uint32_t max_batch;
  float min_roughness;      // vertex min_bounce ends in the cache if roughness >= this; later ones always
  uint32_t min_bounce;      // first vertex index (0 = primary hit) that may end in the cache
  uint32_t train_end;       // every path ends in the cache at this vertex (0: training paths
                            // never do); training paths' earlier vertices then learn from the
                            // cache's own prediction
  float unbiased_fraction;  // training paths exempt from train_end
  TrainPath* tpaths;
  TrainVertex* tverts;
  CacheQuery* queries;
  CacheRecord* qrec;
  float* qout;
  CacheQuery* batch_q;
  float* batch_t;
};

struct Camera {
  float inv_view_proj[16];  // column-major inverse(projection * view)... stored as V^-1 * P^-1
  float eye[3];
  uint32_t width, height;
  // Depth of field (pt_generate.comp:38-52, bokeh.glsl:15-34): lens size, focus distance and an
  // optional bokeh mask (rgba, bilinear, clamped) sampled at 0.5 + lens offset.
  uint32_t dof;
  float lens[2];
  float focus;
  const float4* bokeh;
  uint32_t bokeh_w, bokeh_h;
};

// An albedo texture in the scene's texel array (rgba floats, row-major, row 0 at v = 0).
struct TexInfo {
  uint32_t offset;
  uint32_t w, h;
  uint32_t pad;
};

struct PtScene {
  const InstanceRecord* instances;
  const uint32_t* inst_material;
  const GeoInfo* geos;
  const uint32_t* idx;
  const float4* pos;
  const float4* nrm;
  const float2* uv;
  const Material* materials;
  const QuadLight* lights;
  const TexInfo* textures;
  const float4* texels;
  const float4* cubemap;   // 6 faces (+x, -x, +y, -y, +z, -z) of cube_size^2 texels, or null
  uint32_t cube_size;
  uint32_t num_lights;
  uint32_t num_materials;
  uint32_t num_textures;
  float env_scale;
};

struct PtBuffers {
  void* rays[2];          // mrt_ray queues
  PathState* paths[2];
  void* hits;             // mrt_hit
  void* shadow_rays;      // mrt_ray
  ShadowRecord* shadow;
  uint8_t* occluded;
  uint32_t* counts;       // [0], [1]: path queues; [2]: shadow queue; [3] training paths;
                          // [4] cache queries; [5] training samples; [8..9] rays traced (u64);
                          // [10..11] non-finite values rejected (u64)
  float4* accum;          // per pixel running sum (rgb); w unused
  float4* aux;            // per pixel 2 x float4 of primary-hit sums: albedo, t; normal, hits
  const uint32_t* perm;   // shading position -> queue entry (material order), or null
  uint32_t capacity;
  PtCache cache;
};

int launch_pt_generate(const Camera& cam, const PtBuffers& b, uint32_t cur, uint32_t n_paths, uint32_t sample_base,
                       hipStream_t stream);
int launch_pt_shade(const PtScene& sc, const PtBuffers& b, uint32_t cur, uint32_t max_depth, hipStream_t stream);
int launch_pt_resolve(const PtBuffers& b, hipStream_t stream);
int launch_pt_cache_resolve(const PtBuffers& b, hipStream_t stream);
int launch_pt_train_batch(const PtBuffers& b, hipStream_t stream);
int launch_pt_mean(const float4* accum, uint32_t spp, float4* out, size_t npix, hipStream_t stream);
int launch_pt_denoise_input(const float4* accum, const float4* aux, uint32_t spp, float* raw, size_t npix,
                            hipStream_t stream);

}  // namespace mrt
