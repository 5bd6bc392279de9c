// Scene registry, BVH preparation and the C ABI of libmyrt_amd.so.
//
// Mirrors myrt::scene (src/pathtracer/scene.hpp:60-255, scene.cpp:43-71, 255-409): geometries
// are copied and get their object BVH (split_aabbs) at push time; drawables are enqueued per
// frame and consumed by prepare(), which builds the instance BVH (dont_split) over the
// transformed geometry boxes. Storage is HIP device memory instead of GL buffers, and every BVH
// is collapsed to the 64-byte wide format the traversal kernels read.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <thread>
#include <atomic>
#include <functional>

#include "../../include/myrt.h"
#include "bvh_build.hpp"
#include "device_types.hpp"
#include "scene_internal.hpp"
#include "refit.hpp"
#include "sdf_grid.hpp"
#include "sdf_kernels.hpp"
#include "trace.hpp"
#include "wide_bvh.hpp"
#include "abi_guard.hpp"

namespace mrt {

int launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);
int launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream);
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(MRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Inverse of a column-major 4x4 matrix by cofactor expansion in binary64, rounded to binary32.
// Returns the first three rows of the inverse as row-major 3x4.
bool invert_affine_rows(const float m[16], float out[12]) {
  double a[4][4];  // a[r][c]
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) a[r][c] = double(m[c * 4 + r]);
  // 2x2 minors of rows 0-1 and rows 2-3.
  const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  if (!(det != 0.0) || !std::isfinite(det)) return false;
  const double id = 1.0 / det;
  double inv[3][4];
  inv[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id;
  inv[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id;
  inv[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id;
  inv[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id;
  inv[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id;
  inv[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id;
  inv[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id;
  inv[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id;
  inv[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id;
  inv[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id;
  inv[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id;
  inv[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) out[r * 4 + c] = float(inv[r][c]);
  return true;
}

namespace {

constexpr double kBlasPadRel = 1.0 / (1 << 20);
constexpr double kTlasPadRel = 1.0 / (1 << 16);

template <typename T>
int upload(T*& dptr, size_t& cap, const std::vector<T>& v) {
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  if (!dptr || cap < bytes) {
    if (dptr) hipFree(dptr);
    dptr = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dptr), bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    cap = bytes;
  }
  if (!v.empty()) {
    hipError_t e = hipMemcpy(dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  }
  return MRT_OK;
}

}  // namespace

// scene::enqueue's instance box: the eight corners of the geometry box through M (scene.cpp:256-261).
Box transform_box(const Box& b, const float m[16]) {
  Box out;
  for (unsigned i = 0; i < 8; ++i) {
    const float x = (i & 1u) ? b.mx.x : b.mn.x;
    const float y = (i & 2u) ? b.mx.y : b.mn.y;
    const float z = (i & 4u) ? b.mx.z : b.mn.z;
    vec3f p;
    for (int r = 0; r < 3; ++r) p[r] = m[0 * 4 + r] * x + m[1 * 4 + r] * y + m[2 * 4 + r] * z + m[3 * 4 + r];
    out.enclose(p);
  }
  return out;
}

// Wide TLAS nodes grouped by depth, deepest level first, for the level-by-level refit.
void Scene::tlas_levels() {
  std::vector<uint32_t> depth(tlas_nodes.size(), 0);
  uint32_t max_depth = 0;
  // Children are numbered after their parent, so one forward pass sets every depth.
  for (uint32_t i = 0; i < tlas_nodes.size(); ++i) {
    const WideNode& n = tlas_nodes[i];
    for (int c = 0; c < n.nchildren; ++c)
      if (!(n.entry[c] & 0x80000000u)) {
        const uint32_t j = n.entry[c];
        depth[j] = depth[i] + 1;
        max_depth = std::max(max_depth, depth[j]);
      }
  }
  std::vector<uint32_t> per(max_depth + 1, 0);
  for (uint32_t d : depth) ++per[d];
  tlas_level_start.assign(max_depth + 2, 0u);
  for (uint32_t l = 0; l <= max_depth; ++l) tlas_level_start[l + 1] = tlas_level_start[l] + per[max_depth - l];
  tlas_level_nodes.assign(tlas_nodes.size(), 0u);
  std::vector<uint32_t> fill(tlas_level_start.begin(), tlas_level_start.end() - 1);
  for (uint32_t i = 0; i < tlas_nodes.size(); ++i) tlas_level_nodes[fill[max_depth - depth[i]]++] = i;
}

int Scene::build_host() {
  try {
    if (geoms_dirty) {
      blas_nodes.clear();
      tris.clear();
      blas_root.assign(geoms.size(), 0);
      blas_depth = 0;
      for (size_t g = 0; g < geoms.size(); ++g) {
        const Geometry& geo = geoms[g];
        const uint32_t node_off = uint32_t(blas_nodes.size());
        const uint32_t item_off = uint32_t(tris.size());
        blas_root[g] = node_off;
        blas_depth = std::max(blas_depth, geo.wide.depth);
        for (WideNode n : geo.wide.nodes) {
          // Entries index the geometry's own arrays; move them to the concatenated ones.
          for (int c = 0; c < n.nchildren; ++c)
            n.entry[c] += (n.entry[c] & 0x80000000u) ? item_off : node_off;
          blas_nodes.push_back(n);
        }
        for (uint32_t item : geo.wide.items) {
          const uint32_t prim = geo.bvh.indices[item];
          TriRecord t{};
          const float* p0 = &geo.pos[size_t(geo.idx[3 * size_t(prim) + 0]) * 3];
          const float* p1 = &geo.pos[size_t(geo.idx[3 * size_t(prim) + 1]) * 3];
          const float* p2 = &geo.pos[size_t(geo.idx[3 * size_t(prim) + 2]) * 3];
          std::memcpy(t.v0, p0, 12);
          std::memcpy(t.v1, p1, 12);
          std::memcpy(t.v2, p2, 12);
          t.prim = prim;
          tris.push_back(t);
        }
      }
      if (tris.size() >= kMaxItemIndex) return fail(MRT_ERR_LIMIT, "scene exceeds 16M triangle references");
      geo_boxes.clear();
      for (const Geometry& geo : geoms)
        for (float v : {geo.box.mn.x, geo.box.mn.y, geo.box.mn.z, geo.box.mx.x, geo.box.mx.y, geo.box.mx.z})
          geo_boxes.push_back(v);
      geo_info.clear();
      all_idx.clear();
      all_pos.clear();
      all_nrm.clear();
      all_uv.clear();
      for (const Geometry& geo : geoms) {
        GeoInfo gi{};
        gi.idx_base = uint32_t(all_idx.size());
        gi.vtx_base = uint32_t(all_pos.size());
        gi.has_normals = geo.nrm.empty() ? 0u : 1u;
        geo_info.push_back(gi);
        all_idx.insert(all_idx.end(), geo.idx.begin(), geo.idx.end());
        const size_t nv = geo.pos.size() / 3;
        for (size_t v = 0; v < nv; ++v) {
          all_pos.push_back(make_float4(geo.pos[3 * v], geo.pos[3 * v + 1], geo.pos[3 * v + 2], 0.f));
          if (!geo.nrm.empty())
            all_nrm.push_back(make_float4(geo.nrm[3 * v], geo.nrm[3 * v + 1], geo.nrm[3 * v + 2], 0.f));
          else
            all_nrm.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
          all_uv.push_back(make_float2(geo.uv[2 * v], geo.uv[2 * v + 1]));
        }
      }
    }
    if (inst_dirty) {
      tlas_nodes.clear();
      tlas_items.clear();
      inst_records.clear();
      tlas_depth = 0;
      if (!current.empty()) {
        std::vector<Box> boxes;
        boxes.reserve(current.size());
        for (const Instance& in : current) boxes.push_back(in.world_box);
        tlas = sah_builder ? build_sah_bvh(boxes, tlas_leaf) : build_binary_bvh(boxes, false);
        WideBvh w = collapse_to_wide(tlas, kTlasPadRel, tlas_leaf);
        tlas_nodes = std::move(w.nodes);
        tlas_depth = w.depth;
        for (uint32_t item : w.items) tlas_items.push_back(tlas.indices[item]);
        if (tlas_items.size() >= kMaxItemIndex) return fail(MRT_ERR_LIMIT, "scene exceeds 16M instance references");
        tlas_levels();
        inst_material.clear();
        for (const Instance& in : current) inst_material.push_back(in.material);
        for (size_t i = 0; i < current.size(); ++i) {
          InstanceRecord r{};
          std::memcpy(r.inv, current[i].inv, sizeof(r.inv));
          r.geom = current[i].geom;
          r.inst = uint32_t(i);
          inst_records.push_back(r);
        }
      } else {
        tlas = BinaryBvh{};
        tlas_level_nodes.clear();
        tlas_level_start.assign(1, 0u);
      }
    }
    for (InstanceRecord& r : inst_records) r.blas_root = blas_root[r.geom];
    max_stack = 3 * (tlas_depth + blas_depth) + 2 + 4;
  } catch (const std::exception& ex) {
    return fail(MRT_ERR_INTERNAL, ex.what());
  }
  return MRT_OK;
}

int Scene::upload_all() {
  int rc;
  if (geoms_dirty) {
    if ((rc = upload(d_blas, cap_blas, blas_nodes)) != MRT_OK) return rc;
    if ((rc = upload(d_tris, cap_tris, tris)) != MRT_OK) return rc;
  }
  if (inst_dirty || geoms_dirty) {
    if ((rc = upload(d_tlas, cap_tlas, tlas_nodes)) != MRT_OK) return rc;
    if ((rc = upload(d_tlas_items, cap_items, tlas_items)) != MRT_OK) return rc;
    if ((rc = upload(d_inst, cap_inst, inst_records)) != MRT_OK) return rc;
    if ((rc = upload(d_inst_material, cap_inst_mat, inst_material)) != MRT_OK) return rc;
    if ((rc = upload(d_level_nodes, cap_level_nodes, tlas_level_nodes)) != MRT_OK) return rc;
    if ((rc = upload(d_level_start, cap_level_start, tlas_level_start)) != MRT_OK) return rc;
    tlas_host_stale = false;
  }
  if (geoms_dirty) {
    if ((rc = upload(d_geo_box, cap_geo_box, geo_boxes)) != MRT_OK) return rc;
  }
  if (geoms_dirty) {
    if ((rc = upload(d_geo_info, cap_geo_info, geo_info)) != MRT_OK) return rc;
    if ((rc = upload(d_idx, cap_idx, all_idx)) != MRT_OK) return rc;
    if ((rc = upload(d_pos, cap_pos, all_pos)) != MRT_OK) return rc;
    if ((rc = upload(d_nrm, cap_nrm, all_nrm)) != MRT_OK) return rc;
    if ((rc = upload(d_uv, cap_uv, all_uv)) != MRT_OK) return rc;
  }
  if (sdf_dirty) {
    // The device drawables address the compiled programs.
    std::vector<SdfDrawable> dev = sdf_drawables;
    for (size_t j = 0; j < dev.size(); ++j) {
      dev[j].tape_offset = sdf_prog_offset[sdf_current[j].sdf];
      dev[j].tape_len = sdf_prog_len[sdf_current[j].sdf];
    }
    std::vector<SdfGrid> grids(dev.size());
    sdf_chain_only = true;
    for (size_t j = 0; j < dev.size(); ++j) {
      grids[j] = sdf_grid[sdf_current[j].sdf];
      if (grids[j].dim == 0) sdf_chain_only = false;
    }
    if ((rc = upload(d_sdf_prog, cap_sdf_tape, sdf_prog)) != MRT_OK) return rc;
    if ((rc = upload(d_sdf_drawables, cap_sdf_dr, dev)) != MRT_OK) return rc;
    if ((rc = upload(d_sdf_grids, cap_sdf_grids, grids)) != MRT_OK) return rc;
    // Masks travel only when programs were pushed; grids of changed programs are rebuilt on the
    // GPU in place.
    if (sdf_masks_dirty || !d_sdf_masks) {
      if ((rc = sync_host_masks()) != MRT_OK) return rc;
      std::vector<uint64_t> masks = sdf_masks;
      if (masks.empty()) masks.push_back(~0ull);
      if ((rc = upload(d_sdf_masks, cap_sdf_masks, masks)) != MRT_OK) return rc;
      sdf_masks_dirty = false;
    }
    if (!sdf_grid_rebuild.empty()) {
      if (!d_sdf_coarse) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_sdf_coarse),
                                 sizeof(uint64_t) * kSdfGridCoarse * kSdfGridCoarse * kSdfGridCoarse);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc");
      }
      for (uint32_t id : sdf_grid_rebuild) {
        const SdfGrid& g = sdf_grid[id];
        if (!g.dim) continue;
        if (launch_sdf_grid(d_sdf_prog + sdf_prog_offset[id], sdf_prog_len[id], sdf_bounds[id], d_sdf_coarse,
                            d_sdf_masks + g.mask_offset, nullptr))
          return fail(MRT_ERR_HIP, "sdf grid rebuild");
        sdf_masks_host_stale = true;
      }
      sdf_grid_rebuild.clear();
    }
    sdf_dirty = false;
  }
  if (shading_dirty) {
    std::vector<Material> mats(this->mats.materials.size());
    for (size_t i = 0; i < mats.size(); ++i) mats[i] = this->mats.resolve(uint32_t(i));
    if (mats.empty()) {
      Material def{};
      def.albedo_rgba8 = 0xffccccccu;
      def.ior = 1.5f;
      def.roughness = 1.0f;
      mats.push_back(def);
    }
    if ((rc = upload(d_materials, cap_mat, mats)) != MRT_OK) return rc;
    if ((rc = upload(d_lights, cap_lights, lights)) != MRT_OK) return rc;
    std::vector<TexInfo> tex = textures;
    std::vector<float4> tx = texels;
    if (tex.empty()) tex.push_back(TexInfo{0, 1, 1, 0});
    if (tx.empty()) tx.push_back(make_float4(1.f, 1.f, 1.f, 1.f));
    if ((rc = upload(d_textures, cap_textures, tex)) != MRT_OK) return rc;
    if ((rc = upload(d_texels, cap_texels, tx)) != MRT_OK) return rc;
    shading_dirty = false;
  }
  if (!ctx[0].ctrl) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&ctx[0].ctrl), 256);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    num_cus = prop.multiProcessorCount;
  }
  return MRT_OK;
}

Scene::~Scene() {
  for (void* p : {(void*)d_sdf_drawables, (void*)d_sdf_prog, (void*)d_sdf_grids, (void*)d_sdf_masks, (void*)d_geo_info, (void*)d_idx, (void*)d_inst_material, (void*)d_pos, (void*)d_nrm,
                  (void*)d_materials, (void*)d_lights, (void*)d_uv, (void*)d_textures, (void*)d_texels})
    if (p) hipFree(p);
  if (d_blas) hipFree(d_blas);
  if (d_tris) hipFree(d_tris);
  for (TraceScratch& c : ctx)
    for (void* q : {(void*)c.ctrl, (void*)c.spill, (void*)c.sdf_side, (void*)c.sdf_hits})
      if (q) hipFree(q);
  for (void* q : {d_host_rays, d_host_out})
    if (q) hipFree(q);
  if (d_sdf_coarse) hipFree(d_sdf_coarse);
  for (void* q : {(void*)d_level_nodes, (void*)d_level_start, (void*)d_geo_box, (void*)d_inst_box, (void*)d_node_box, (void*)d_matrices,
                  (void*)d_refit_status})
    if (q) hipFree(q);
  if (d_tlas) hipFree(d_tlas);
  if (ctrl_host) hipHostFree(ctrl_host);
  if (d_tlas_items) hipFree(d_tlas_items);
  if (d_inst) hipFree(d_inst);
}

PtScene Scene::shading_view() const {
  PtScene v{};
  v.instances = d_inst;
  v.inst_material = d_inst_material;
  v.geos = d_geo_info;
  v.idx = d_idx;
  v.pos = d_pos;
  v.nrm = d_nrm;
  v.uv = d_uv;
  v.textures = d_textures;
  v.texels = d_texels;
  v.num_textures = uint32_t(textures.size());
  v.materials = d_materials;
  v.lights = d_lights;
  v.num_lights = uint32_t(lights.size());
  v.num_materials = uint32_t(std::max<size_t>(mats.materials.size(), 1));
  v.env_scale = env_scale;
  return v;
}

uint32_t Scene::spill_depth() const { return max_stack > lds_depth ? max_stack - lds_depth : 1; }

int Scene::full_grid(bool any, bool stats, size_t n) const {
  // Seven waves per SIMD pay off where rays change instances often (the per-iteration work is
  // larger and more divergent); a single mesh runs best at six. MRT_TRACE_WAVES overrides.
  const uint32_t waves = trace_waves ? trace_waves : (current.size() > 1 ? 7u : 6u);
  const int per_cu = trace_blocks_per_cu(any, stats, waves, lds_depth, current.size() == 1);
  const size_t max_grid = size_t(num_cus) * size_t(per_cu);
  const size_t want = (n + kTraceBlock - 1) / kTraceBlock;
  return int(std::max<size_t>(1, std::min(max_grid, want)));
}

int Scene::trace(const void* d_rays, size_t n, void* d_out, bool any, bool stats, hipStream_t stream,
                 const uint32_t* n_dev, int context) {
  if (!prepared) return fail(MRT_ERR_NOT_PREPARED, "trace before mrt_scene_prepare");
  if (n == 0) return MRT_OK;
  if (n >= 0x7fffffffu) return fail(MRT_ERR_LIMIT, "at most 2^31-1 rays per call");
  if (context < 0 || context >= kTraceContexts) return fail(MRT_ERR_INVALID, "trace context out of range");
  TraceScratch& cx = ctx[context];
  if (!cx.ctrl) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&cx.ctrl), 256);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc trace counters");
  }
  const size_t spill_entries = size_t(spill_depth()) * size_t(full_grid(any, stats, n)) * kTraceBlock;
  if (spill_entries > cx.spill_cap) {
    if (cx.spill) hipFree(cx.spill);
    cx.spill = nullptr;
    cx.spill_cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&cx.spill), spill_entries * 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc spill");
    cx.spill_cap = spill_entries;
  }
  if (!sdf_drawables.empty() && !any && cx.sdf_side_cap < n) {
    if (cx.sdf_side) hipFree(cx.sdf_side);
    if (cx.sdf_hits) hipFree(cx.sdf_hits);
    cx.sdf_side = nullptr;
    cx.sdf_hits = nullptr;
    cx.sdf_side_cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&cx.sdf_side), n * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&cx.sdf_hits), n * sizeof(uint32_t));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc sdf hit points");
    cx.sdf_side_cap = n;
  }
  const TraceArena a{cx.ctrl, cx.spill, cx.spill_cap, cx.sdf_side, cx.sdf_hits, cx.sdf_side_cap};
  return trace_in(a, d_rays, n, d_out, any, stats, stream, n_dev);
}

int Scene::trace_in(const TraceArena& a, const void* d_rays, size_t n, void* d_out, bool any, bool stats,
                    hipStream_t stream, const uint32_t* n_dev) {
  if (!prepared) return fail(MRT_ERR_NOT_PREPARED, "trace before mrt_scene_prepare");
  if (n == 0) return MRT_OK;
  if (n >= 0x7fffffffu) return fail(MRT_ERR_LIMIT, "at most 2^31-1 rays per call");
  const bool sdf = !sdf_drawables.empty();
  if (sdf && !any && a.sdf_cap < n) return fail(MRT_ERR_LIMIT, "SDF hit buffers smaller than the ray count");
  const uint32_t depth = spill_depth();
  // The persistent grid is cut to the spill stack the arena holds; each thread needs `depth` entries.
  const size_t fit = a.spill_entries / (size_t(depth) * kTraceBlock);
  if (fit == 0) return fail(MRT_ERR_LIMIT, "spill stack too small for one traversal workgroup");
  const int grid = int(std::min<size_t>(size_t(full_grid(any, stats, n)), fit));
  const uint32_t total_threads = uint32_t(grid) * kTraceBlock;
  if (launch_fill_u32(a.ctrl, 64, 0u, stream) != 0) return fail(MRT_ERR_HIP, "clearing trace counters");
  TraceParams p{};
  p.tlas_nodes = d_tlas;
  p.tlas_items = d_tlas_items;
  p.instances = d_inst;
  p.blas_nodes = d_blas;
  p.tris = d_tris;
  p.rays = d_rays;
  p.out = d_out;
  p.n = uint32_t(n);
  p.n_dev = n_dev;
  p.has_tlas = tlas_nodes.empty() ? 0u : 1u;
  p.counter = a.ctrl;
  p.flags = a.ctrl + 1;
  p.stats = reinterpret_cast<unsigned long long*>(a.ctrl + 8);
  p.spill = a.spill;
  p.spill_depth = depth;
  p.total_threads = total_threads;
  const bool single = current.size() == 1;
  p.refill_below = refill_below != kKnobUnset ? refill_below : (single ? 52u : 56u);
  p.grab = trace_grab;
  p.lds_depth = lds_depth;
  p.waves = trace_waves ? trace_waves : (current.size() > 1 ? 7u : 6u);
  p.single_instance = current.size() == 1 ? 0u : 0xFFFFFFFFu;
  p.leaf_wait = leaf_wait != kKnobUnset ? leaf_wait : (single ? 20u : 16u);
  if (launch_trace(p, any, stats, grid, stream) != 0) return fail(MRT_ERR_HIP, "trace kernel launch failed");
  if (sdf) {
    SdfParams sp{};
    sp.drawables = d_sdf_drawables;
    sp.prog = d_sdf_prog;
    sp.grids = d_sdf_grids;
    sp.masks = d_sdf_masks;
    sp.num_drawables = uint32_t(sdf_drawables.size());
    sp.rays = d_rays;
    sp.out = d_out;
    sp.n = uint32_t(n);
    sp.n_dev = n_dev;
    sp.steps = sdf_steps;
    sp.eps = sdf_eps;
    sp.side = any ? nullptr : a.sdf_side;
    sp.hit_list = any ? nullptr : a.sdf_hits;
    sp.hit_count = a.ctrl + 3;  // zeroed with the traversal counters above
    sp.prog_len = uint32_t(sdf_prog.size());
    sp.chain_only = sdf_chain_only ? 1u : 0u;
    if (launch_sdf(sp, any, num_cus, stream) != 0)
      return fail(MRT_ERR_HIP, "sdf kernel launch failed");
  }
  return MRT_OK;
}

int Scene::read_ctrl(uint32_t* flags, unsigned long long* stats, hipStream_t stream) {
  // Pinned staging: the copies are plain DMA on the stream and one synchronisation covers them
  // and whatever the caller queued before (a render reads its own counters the same way).
  hipError_t e = hipSuccess;
  if (!ctrl_host) e = hipHostMalloc(reinterpret_cast<void**>(&ctrl_host), kTraceContexts * 64, hipHostMallocDefault);
  if (e != hipSuccess) {
    ctrl_host = nullptr;
    return hip_fail(e, "hipHostMalloc trace counter staging");
  }
  uint32_t (*host)[16] = reinterpret_cast<uint32_t (*)[16]>(ctrl_host);
  std::memset(ctrl_host, 0, kTraceContexts * 64);
  for (int c = 0; c < kTraceContexts && e == hipSuccess; ++c)
    if (ctx[c].ctrl) e = hipMemcpyAsync(host[c], ctx[c].ctrl, 64, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "reading trace counters");
  *flags = 0;
  for (int c = 0; c < kTraceContexts; ++c) *flags |= host[c][1];
  if (stats) std::memcpy(stats, host[0] + 8, 4 * sizeof(unsigned long long));
  return MRT_OK;
}

}  // namespace mrt

using namespace mrt;

struct mrt_bvh_s {
  BinaryBvh b;
};

struct mrt_timer_s {
  hipEvent_t start = nullptr, stop = nullptr;
};


extern "C" {

const char* mrt_last_error(void) { return g_last_error.c_str(); }

int mrt_device_info(char* name_out, size_t name_cap, int* num_cus) {
  MRT_GUARD_BEGIN
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (name_out && name_cap) {
    std::snprintf(name_out, name_cap, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (num_cus) *num_cus = prop.multiProcessorCount;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_device_alloc(size_t bytes, void** out) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output pointer");
  hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 1));
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipMalloc");
  MRT_GUARD_END
}
int mrt_device_free(void* ptr) {
  MRT_GUARD_BEGIN
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipFree");
  MRT_GUARD_END
}
int mrt_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  MRT_GUARD_BEGIN
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipMemcpy H2D");
  MRT_GUARD_END
}
int mrt_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  MRT_GUARD_BEGIN
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipMemcpy D2H");
  MRT_GUARD_END
}
int mrt_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
  MRT_GUARD_BEGIN
  if (!dst || !src || bytes % 16 != 0) return fail(MRT_ERR_INVALID, "copy needs non-null pointers and a multiple of 16 bytes");
  return mrt::launch_copy(src, dst, bytes, static_cast<hipStream_t>(stream)) == 0 ? MRT_OK
                                                                                  : fail(MRT_ERR_HIP, "copy kernel launch failed");
  MRT_GUARD_END
}

int mrt_device_synchronize(void) {
  MRT_GUARD_BEGIN
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipDeviceSynchronize");
  MRT_GUARD_END
}

int mrt_timer_create(mrt_timer* out) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output pointer");
  auto* t = new mrt_timer_s;
  hipError_t e = hipEventCreate(&t->start);
  if (e == hipSuccess) e = hipEventCreate(&t->stop);
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "hipEventCreate");
  }
  *out = t;
  return MRT_OK;
  MRT_GUARD_END
}
int mrt_timer_destroy(mrt_timer t) {
  MRT_GUARD_BEGIN
  if (!t) return MRT_OK;
  hipEventDestroy(t->start);
  hipEventDestroy(t->stop);
  delete t;
  return MRT_OK;
  MRT_GUARD_END
}
int mrt_timer_start(mrt_timer t, void* stream) {
  MRT_GUARD_BEGIN
  if (!t) return fail(MRT_ERR_INVALID, "null timer");
  hipError_t e = hipEventRecord(t->start, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipEventRecord");
  MRT_GUARD_END
}
int mrt_timer_stop(mrt_timer t, void* stream) {
  MRT_GUARD_BEGIN
  if (!t) return fail(MRT_ERR_INVALID, "null timer");
  hipError_t e = hipEventRecord(t->stop, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipEventRecord");
  MRT_GUARD_END
}
int mrt_timer_elapsed_ms(mrt_timer t, float* ms) {
  MRT_GUARD_BEGIN
  if (!t || !ms) return fail(MRT_ERR_INVALID, "null argument");
  hipError_t e = hipEventSynchronize(t->stop);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, t->start, t->stop);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "hipEventElapsedTime");
  MRT_GUARD_END
}

static void box_to_c(const Box& b, mrt_aabb* o) {
  o->min[0] = b.mn.x; o->min[1] = b.mn.y; o->min[2] = b.mn.z;
  o->max[0] = b.mx.x; o->max[1] = b.mx.y; o->max[2] = b.mx.z;
  o->centroid[0] = b.wc.x; o->centroid[1] = b.wc.y; o->centroid[2] = b.wc.z;
  o->has_centroid = b.has_wc ? 1u : 0u;
}
static Box box_from_c(const mrt_aabb& a) {
  Box b;
  b.mn = {a.min[0], a.min[1], a.min[2]};
  b.mx = {a.max[0], a.max[1], a.max[2]};
  b.wc = {a.centroid[0], a.centroid[1], a.centroid[2]};
  b.has_wc = a.has_centroid != 0;
  return b;
}

static int check_mesh(const uint32_t* idx, size_t n_idx, const float* pos3, size_t n_vtx) {
  if (!idx || !pos3) return fail(MRT_ERR_INVALID, "null index or position pointer");
  if (n_idx == 0 || n_idx % 3 != 0) return fail(MRT_ERR_INVALID, "index count must be a positive multiple of 3");
  for (size_t i = 0; i < n_idx; ++i)
    if (idx[i] >= n_vtx) return fail(MRT_ERR_INVALID, "index out of range of the vertex array");
  return MRT_OK;
}

int mrt_generate_triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3, size_t n_vtx,
                                 mrt_aabb* out_boxes) {
  MRT_GUARD_BEGIN
  int rc = check_mesh(idx, n_idx, pos3, n_vtx);
  if (rc != MRT_OK) return rc;
  if (!out_boxes) return fail(MRT_ERR_INVALID, "null output");
  const std::vector<Box> boxes = triangle_bounds(idx, n_idx, pos3);
  for (size_t i = 0; i < boxes.size(); ++i) box_to_c(boxes[i], out_boxes + i);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_bvh_build(const mrt_aabb* boxes, size_t n, int split_mode, mrt_bvh* out) {
  MRT_GUARD_BEGIN
  if (!boxes || !out || n == 0) return fail(MRT_ERR_INVALID, "need at least one box and an output handle");
  if (split_mode != MRT_SPLIT_NONE && split_mode != MRT_SPLIT_AABBS) return fail(MRT_ERR_INVALID, "bad split mode");
  std::vector<Box> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = box_from_c(boxes[i]);
  auto* h = new mrt_bvh_s;
  try {
    h->b = build_binary_bvh(v, split_mode == MRT_SPLIT_AABBS);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_bvh_build_triangles(const uint32_t* idx, size_t n_idx, const float* pos3, size_t n_vtx, int split_mode,
                            mrt_bvh* out) {
  MRT_GUARD_BEGIN
  int rc = check_mesh(idx, n_idx, pos3, n_vtx);
  if (rc != MRT_OK) return rc;
  if (!out) return fail(MRT_ERR_INVALID, "null output handle");
  if (split_mode != MRT_SPLIT_NONE && split_mode != MRT_SPLIT_AABBS) return fail(MRT_ERR_INVALID, "bad split mode");
  auto* h = new mrt_bvh_s;
  try {
    h->b = build_binary_bvh(triangle_bounds(idx, n_idx, pos3), split_mode == MRT_SPLIT_AABBS);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_bvh_sizes(mrt_bvh b, size_t* n_nodes, size_t* n_indices) {
  MRT_GUARD_BEGIN
  if (!b) return fail(MRT_ERR_INVALID, "null bvh");
  if (n_nodes) *n_nodes = b->b.nodes.size();
  if (n_indices) *n_indices = b->b.indices.size();
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_bvh_copy(mrt_bvh b, mrt_bvh_node* nodes, uint32_t* indices) {
  MRT_GUARD_BEGIN
  if (!b) return fail(MRT_ERR_INVALID, "null bvh");
  static_assert(sizeof(mrt_bvh_node) == sizeof(BinNode), "node layouts differ");
  if (nodes) std::memcpy(nodes, b->b.nodes.data(), b->b.nodes.size() * sizeof(BinNode));
  if (indices) std::memcpy(indices, b->b.indices.data(), b->b.indices.size() * 4);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_bvh_destroy(mrt_bvh b) {
  MRT_GUARD_BEGIN
  delete b;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_create(mrt_scene* out) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output handle");
  Scene* s = new Scene;
  // Launch-shape knobs for tuning sweeps (the defaults are the measured best).
  if (const char* v = std::getenv("MRT_TRACE_LDS_STACK")) {
    const long d = std::strtol(v, nullptr, 10);
    if (d >= 4 && d <= 64 && (d & (d - 1)) == 0) s->lds_depth = uint32_t(d);
  }
  if (const char* v = std::getenv("MRT_BVH_LEAF")) {
    const long l = std::strtol(v, nullptr, 10);
    if (l >= 1 && l <= 127) s->blas_leaf = uint32_t(l);
  }
  if (const char* v = std::getenv("MRT_TLAS_LEAF")) {
    const long l = std::strtol(v, nullptr, 10);
    if (l >= 1 && l <= 127) s->tlas_leaf = uint32_t(l);
  }
  if (const char* v = std::getenv("MRT_TRACE_REFILL")) {
    const long r = std::strtol(v, nullptr, 10);
    if (r >= 1 && r <= 64) s->refill_below = uint32_t(r);
  }
  if (const char* v = std::getenv("MRT_TRACE_GRAB")) {
    const long g = std::strtol(v, nullptr, 10);
    if (g >= 1 && g <= 4096) s->trace_grab = uint32_t(g);
  }
  if (const char* v = std::getenv("MRT_TRACE_LEAF_WAIT")) {
    const long w = std::strtol(v, nullptr, 10);
    if (w >= 0 && w <= 64) s->leaf_wait = uint32_t(w);
  }
  if (const char* v = std::getenv("MRT_BVH_BUILDER")) s->sah_builder = std::strtol(v, nullptr, 10) != 0;
  if (const char* v = std::getenv("MRT_TRACE_WAVES")) {
    const long w = std::strtol(v, nullptr, 10);
    if (w == 6 || w == 7) s->trace_waves = uint32_t(w);
  }
  *out = reinterpret_cast<mrt_scene>(s);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_destroy(mrt_scene s) {
  MRT_GUARD_BEGIN
  delete reinterpret_cast<Scene*>(s);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_push_geometry(mrt_scene sh, const uint32_t* idx, size_t n_idx, const float* pos3, const float* nrm3,
                            const float* uv2, size_t n_vtx, uint32_t* out_geom) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  int rc = check_mesh(idx, n_idx, pos3, n_vtx);
  if (rc != MRT_OK) return rc;
  Geometry g;
  g.idx.assign(idx, idx + n_idx);
  g.pos.assign(pos3, pos3 + 3 * n_vtx);
  if (nrm3) g.nrm.assign(nrm3, nrm3 + 3 * n_vtx);
  if (uv2) g.uv.assign(uv2, uv2 + 2 * n_vtx);
  else g.uv.assign(2 * n_vtx, 0.f);  // scene.cpp:55-58 zero-fills missing uvs
  g.bvh = s->sah_builder ? build_sah_bvh(triangle_bounds(idx, n_idx, pos3), s->blas_leaf)
                         : build_binary_bvh(triangle_bounds(idx, n_idx, pos3), true);
  g.box.mn = {g.bvh.nodes[0].mn[0], g.bvh.nodes[0].mn[1], g.bvh.nodes[0].mn[2]};
  g.box.mx = {g.bvh.nodes[0].mx[0], g.bvh.nodes[0].mx[1], g.bvh.nodes[0].mx[2]};
  g.wide = collapse_to_wide(g.bvh, 1.0 / (1 << 20), s->blas_leaf);
  s->geoms.push_back(std::move(g));
  s->geoms_dirty = true;
  if (out_geom) *out_geom = uint32_t(s->geoms.size() - 1);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_enqueue_geometry(mrt_scene sh, uint32_t geom, uint32_t material, const float m4x4_colmajor[16],
                               uint32_t* out_inst) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  if (geom >= s->geoms.size()) return fail(MRT_ERR_INVALID, "unknown geometry");
  if (!m4x4_colmajor) return fail(MRT_ERR_INVALID, "null matrix");
  Instance in;
  in.geom = geom;
  in.material = material;
  std::memcpy(in.m, m4x4_colmajor, sizeof(in.m));
  if (!invert_affine_rows(in.m, in.inv)) return fail(MRT_ERR_INVALID, "transformation is not invertible");
  in.world_box = transform_box(s->geoms[geom].box, in.m);
  s->pending.push_back(in);
  if (out_inst) *out_inst = uint32_t(s->pending.size() - 1);
  return MRT_OK;
  MRT_GUARD_END
}

static bool same_drawables(const std::vector<Instance>& a, const std::vector<Instance>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i].geom != b[i].geom || a[i].material != b[i].material || std::memcmp(a[i].m, b[i].m, sizeof(a[i].m)) != 0)
      return false;
  return true;
}

// World-space box of an SDF drawable: the tape's object bounds through M, widened by a margin
// for the marching tolerance.
static void build_sdf_drawables(Scene* s) {
  s->sdf_drawables.clear();
  for (const SdfInstance& in : s->sdf_current) {
    SdfDrawable d{};
    std::memcpy(d.inv, in.inv, sizeof(d.inv));
    d.tape_offset = s->sdf_offset[in.sdf];
    d.tape_len = s->sdf_len[in.sdf];
    const Box wb = transform_box(s->sdf_bounds[in.sdf], in.m);
    double mag = 0;
    for (int a = 0; a < 3; ++a) mag = std::max(mag, std::max(std::fabs(double(wb.mn[a])), std::fabs(double(wb.mx[a]))));
    const float pad = float(1e-4 * (1.0 + mag)) + 1e-3f;
    for (int a = 0; a < 3; ++a) {
      d.box_lo[a] = wb.mn[a] - pad;
      d.box_hi[a] = wb.mx[a] + pad;
    }
    s->sdf_drawables.push_back(d);
  }
  s->sdf_dirty = true;
}

// Same drawables but for their matrices: the prepared TLAS topology still fits.
static bool same_but_transforms(const std::vector<Instance>& a, const std::vector<Instance>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i].geom != b[i].geom || a[i].material != b[i].material) return false;
  return true;
}

static int prepare_impl(Scene* s, mrt_prepare_result* out, bool upload_too) {
  if (!s->sdf_pending.empty()) {
    s->sdf_current = std::move(s->sdf_pending);
    s->sdf_pending.clear();
    build_sdf_drawables(s);
  }
  // The instance BVH is rebuilt only when this frame's drawables differ from the prepared ones
  // (the reference compares drawable hashes, scene.cpp:336-344). When only matrices differ and
  // refits are enabled, the prepared TLAS is refitted on the GPU instead.
  bool refit = false;
  if (!s->pending.empty()) {
    if (!s->prepared || s->device_newer || !same_drawables(s->pending, s->current)) {
      refit = upload_too && s->prepared && !s->geoms_dirty && !s->inst_dirty && s->refits_done < s->refit_limit &&
              !s->tlas_nodes.empty() && same_but_transforms(s->pending, s->current);
      s->current = std::move(s->pending);
      if (!refit) s->inst_dirty = true;
    }
    s->pending.clear();
  }
  const bool gd = s->geoms_dirty, id = s->inst_dirty || s->sdf_dirty || refit;
  if (s->inst_dirty) s->refits_done = 0;
  int rc = s->build_host();
  if (rc != MRT_OK) return rc;
  if (upload_too) {
    rc = s->upload_all();
    if (rc != MRT_OK) return rc;
    if (refit) {
      std::vector<float> mats(s->current.size() * 16);
      for (size_t i = 0; i < s->current.size(); ++i) std::memcpy(&mats[16 * i], s->current[i].m, 64);
      if ((rc = upload(s->d_matrices, s->cap_matrices, mats)) != MRT_OK) return rc;
      if ((rc = s->refit(s->d_matrices, nullptr)) != MRT_OK) return rc;
      // The host already holds these frames' inverses (computed at enqueue, bit-equal to the
      // device's); the refitted nodes are read back only if asked for.
      for (size_t i = 0; i < s->current.size(); ++i) std::memcpy(s->inst_records[i].inv, s->current[i].inv, 48);
      s->device_newer = false;
      ++s->refits_done;
    }
    s->geoms_dirty = false;
    s->inst_dirty = false;
    s->prepared = true;
  }
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->geometries_changed = gd ? 1 : 0;
    out->drawables_changed = id ? 1 : 0;
    out->tlas_refitted = refit ? 1 : 0;
    out->num_geometries = uint32_t(s->geoms.size());
    out->num_instances = uint32_t(s->current.size());
    out->blas_wide_nodes = uint32_t(s->blas_nodes.size());
    out->blas_tri_records = uint32_t(s->tris.size());
    out->tlas_wide_nodes = uint32_t(s->tlas_nodes.size());
    out->max_stack_entries = s->max_stack;
    out->device_bytes = s->blas_nodes.size() * sizeof(WideNode) + s->tris.size() * sizeof(TriRecord) +
                        s->tlas_nodes.size() * sizeof(WideNode) + s->tlas_items.size() * 4 +
                        s->inst_records.size() * sizeof(InstanceRecord);
  }
  return MRT_OK;
}

// Brings the host copy of the culling masks up to the GPU-rebuilt ones.
int Scene::sync_host_masks() {
  if (!sdf_masks_host_stale) return MRT_OK;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess && !sdf_masks.empty())
    e = hipMemcpy(sdf_masks.data(), d_sdf_masks, sdf_masks.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the sdf culling masks");
  sdf_masks_host_stale = false;
  return MRT_OK;
}

int Scene::refit(const float* d_mats, hipStream_t stream) {
  const size_t n = current.size();
  int rc;
  auto grow = [](float*& p, size_t& cap, size_t bytes) -> int {
    if (p && cap >= bytes) return MRT_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    cap = bytes;
    return MRT_OK;
  };
  if ((rc = grow(d_inst_box, cap_inst_box, std::max<size_t>(n, 1) * 24)) != MRT_OK) return rc;
  if ((rc = grow(d_node_box, cap_node_box, std::max<size_t>(tlas_nodes.size(), 1) * 24)) != MRT_OK) return rc;
  if (!d_refit_status) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_refit_status), 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  }
  RefitParams p{};
  p.matrices = d_mats;
  p.n_inst = uint32_t(n);
  p.inst = d_inst;
  p.geo_box = d_geo_box;
  p.inst_box = d_inst_box;
  p.tlas = d_tlas;
  p.items = d_tlas_items;
  p.level_nodes = d_level_nodes;
  p.node_box = d_node_box;
  p.status = d_refit_status;
  p.pad_rel = kTlasPadRel;
  const uint32_t n_levels = uint32_t(tlas_level_start.size() - 1);
  if (tlas_nodes.size() <= kRefitOneBlockNodes) {
    if (launch_refit_one_block(p, d_level_start, n_levels, stream)) return fail(MRT_ERR_HIP, "refit launch");
  } else {
    if (launch_fill_u32(d_refit_status, 1, 0u, stream)) return fail(MRT_ERR_HIP, "refit status clear");
    if (launch_refit(p, tlas_level_start.data(), n_levels, stream)) return fail(MRT_ERR_HIP, "refit launch");
  }
  tlas_host_stale = true;
  return MRT_OK;
}

// Brings tlas_nodes and inst_records up to the device after refits.
int Scene::sync_host_tlas() {
  if (!tlas_host_stale) return MRT_OK;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess && !tlas_nodes.empty())
    e = hipMemcpy(tlas_nodes.data(), d_tlas, tlas_nodes.size() * sizeof(WideNode), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !inst_records.empty())
    e = hipMemcpy(inst_records.data(), d_inst, inst_records.size() * sizeof(InstanceRecord), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the refitted instance BVH");
  if (device_newer)
    for (size_t i = 0; i < current.size() && i < inst_records.size(); ++i)
      std::memcpy(current[i].inv, inst_records[i].inv, 48);
  tlas_host_stale = false;
  return MRT_OK;
}

// Object-space bounds of a tape; also checks the stack discipline. Each operand carries a list of
// boxes whose union holds every point where its distance is below zero. The bounds follow from
// the operators: smooth union r >= min(a, b) - k/4 (equality at a = b), so its zero set lies where
// a or b is below k/4, i.e. within the operands' boxes grown by k/4 (primitive distances are at
// least the distance to their boxes); hard union r = min(a, b); subtraction r >= max(-a, b) >= b
// keeps the right operand's boxes; intersection r >= max(a, b) keeps the left operand's.
static bool sdf_bounds(const SdfNode* t, size_t n, Box& out, std::string& err) {
  std::vector<std::vector<Box>> st;
  std::vector<size_t> mod_depth;  // stack size when each modifier opened
  std::vector<SdfNode> open;
  auto apply_mods = [&](Box b) {
    for (size_t k = open.size(); k-- > 0;) {
      const SdfNode& m = open[k];
      if (m.op == kSdfPushTranslate) {
        for (int a = 0; a < 3; ++a) {
          b.mn[a] += m.p[a];
          b.mx[a] += m.p[a];
        }
      } else {
        const int a = int(m.p[0]);
        const float r = std::max(std::fabs(b.mn[a]), std::fabs(b.mx[a]));
        b.mn[a] = -r;
        b.mx[a] = r;
      }
    }
    return std::vector<Box>{b};
  };
  for (size_t i = 0; i < n; ++i) {
    const SdfNode& nd = t[i];
    Box b;
    switch (nd.op) {
      case kSdfSphere:
        if (!(nd.p[0] > 0.f)) return err = "sphere radius must be positive", false;
        b.mn = {-nd.p[0], -nd.p[0], -nd.p[0]};
        b.mx = {nd.p[0], nd.p[0], nd.p[0]};
        st.push_back(apply_mods(b));
        break;
      case kSdfBox:
        b.mn = {-std::fabs(nd.p[0]), -std::fabs(nd.p[1]), -std::fabs(nd.p[2])};
        b.mx = {std::fabs(nd.p[0]), std::fabs(nd.p[1]), std::fabs(nd.p[2])};
        st.push_back(apply_mods(b));
        break;
      case kSdfTorus: {
        const float r = std::fabs(nd.p[0]) + std::fabs(nd.p[1]);
        b.mn = {-r, -std::fabs(nd.p[1]), -r};
        b.mx = {r, std::fabs(nd.p[1]), r};
        st.push_back(apply_mods(b));
        break;
      }
      case kSdfCapsule:
        if (!(nd.p[0] >= 0.f)) return err = "capsule height must be >= 0", false;
        b.mn = {-std::fabs(nd.p[1]), -std::fabs(nd.p[1]), -std::fabs(nd.p[1])};
        b.mx = {std::fabs(nd.p[1]), nd.p[0] + std::fabs(nd.p[1]), std::fabs(nd.p[1])};
        st.push_back(apply_mods(b));
        break;
      case kSdfMenger:
        b.mn = {-1.f, -1.f, -1.f};
        b.mx = {1.f, 1.f, 1.f};
        st.push_back(apply_mods(b));
        break;
      case kSdfUnion:
      case kSdfSubtract:
      case kSdfIntersect:
      case kSdfSmoothUnion:
      case kSdfSmoothSubtract:
      case kSdfSmoothIntersect: {
        if (st.size() < 2 || (!mod_depth.empty() && st.size() < mod_depth.back() + 2))
          return err = "operator needs two operands", false;
        std::vector<Box> rhs = std::move(st.back());
        st.pop_back();
        std::vector<Box> lhs = std::move(st.back());
        st.pop_back();
        const bool smooth = nd.op >= kSdfSmoothUnion;
        if (smooth && !(nd.p[0] > 0.f)) return err = "smooth operator needs k > 0", false;
        std::vector<Box> r;
        if (nd.op == kSdfUnion || nd.op == kSdfSmoothUnion) {
          r = std::move(lhs);
          r.insert(r.end(), rhs.begin(), rhs.end());
          if (smooth) {
            const float g = 0.25f * nd.p[0];
            for (Box& x : r)
              for (int a = 0; a < 3; ++a) {
                x.mn[a] -= g;
                x.mx[a] += g;
              }
          }
        } else if (nd.op == kSdfSubtract || nd.op == kSdfSmoothSubtract) {
          r = std::move(rhs);
        } else {
          r = std::move(lhs);
        }
        st.push_back(std::move(r));
        break;
      }
      case kSdfPushTranslate:
      case kSdfPushMirror:
        if (nd.op == kSdfPushMirror && !(nd.p[0] == 0.f || nd.p[0] == 1.f || nd.p[0] == 2.f))
          return err = "mirror axis must be 0, 1 or 2", false;
        if (open.size() >= size_t(kSdfPosStack)) return err = "more than 4 nested modifiers", false;
        open.push_back(nd);
        mod_depth.push_back(st.size());
        break;
      case kSdfPop:
        if (open.empty()) return err = "pop without a modifier", false;
        if (st.size() != mod_depth.back() + 1) return err = "a modifier must wrap exactly one operand", false;
        open.pop_back();
        mod_depth.pop_back();
        break;
      default:
        return err = "unknown sdf instruction " + std::to_string(nd.op), false;
    }
    if (st.size() > size_t(kSdfStack)) return err = "operand stack deeper than 8", false;
  }
  if (!open.empty()) return err = "unclosed modifier", false;
  if (st.size() != 1) return err = "the program must leave exactly one distance", false;
  out = Box{};
  for (const Box& x : st[0]) out.enclose(x);
  return true;
}

// Compiles a validated tape into executable instructions: PUSH_TRANSLATE, primitive, POP becomes a
// translated primitive, and a binary operator right after a primitive is fused into it.
static void compile_sdf(const SdfNode* t, size_t n, std::vector<SdfInsn>& out) {
  auto is_prim = [](uint32_t op) { return op >= kSdfSphere && op <= kSdfMenger; };
  auto is_binary = [](uint32_t op) { return op >= kSdfUnion && op <= kSdfSmoothIntersect; };
  size_t i = 0;
  while (i < n) {
    SdfInsn in{};
    if (t[i].op == kSdfPushTranslate && i + 2 < n && is_prim(t[i + 1].op) && t[i + 2].op == kSdfPop) {
      in.op = t[i + 1].op | kSdfFuseTranslate;
      in.material = t[i + 1].material;
      for (int a = 0; a < 3; ++a) in.t[a] = t[i].p[a];
      std::memcpy(in.p, t[i + 1].p, sizeof(in.p));
      i += 3;
    } else {
      in.op = t[i].op;
      in.material = t[i].material;
      std::memcpy(in.p, t[i].p, sizeof(in.p));
      i += 1;
    }
    if (is_prim(in.op & 0xffu) && i < n && is_binary(t[i].op)) {  // validation guarantees an lhs
      in.op |= t[i].op << 8;
      in.k = t[i].p[0];
      i += 1;
    }
    out.push_back(in);
  }
}

int mrt_scene_push_sdf(mrt_scene sh, const mrt_sdf_node* nodes, size_t n, uint32_t* out) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !nodes || n == 0) return fail(MRT_ERR_INVALID, "need a non-empty sdf program");
  static_assert(sizeof(mrt_sdf_node) == sizeof(SdfNode), "sdf node layouts differ");
  const SdfNode* t = reinterpret_cast<const SdfNode*>(nodes);
  Box b;
  std::string err;
  if (!sdf_bounds(t, n, b, err)) return fail(MRT_ERR_INVALID, "invalid sdf program: " + err);
  s->sdf_offset.push_back(uint32_t(s->sdf_tape.size()));
  s->sdf_len.push_back(uint32_t(n));
  s->sdf_tape.insert(s->sdf_tape.end(), t, t + n);
  s->sdf_bounds.push_back(b);
  s->sdf_prog_offset.push_back(uint32_t(s->sdf_prog.size()));
  compile_sdf(t, n, s->sdf_prog);
  s->sdf_prog_len.push_back(uint32_t(s->sdf_prog.size()) - s->sdf_prog_offset.back());
  SdfGrid g;
  int rc = s->sync_host_masks();
  if (rc != MRT_OK) return rc;
  const size_t masks_before = s->sdf_masks.size();
  build_sdf_grid(s->sdf_prog.data() + s->sdf_prog_offset.back(), s->sdf_prog_len.back(), b, g, s->sdf_masks);
  s->sdf_grid.push_back(g);
  if (s->sdf_masks.size() != masks_before) s->sdf_masks_dirty = true;
  s->sdf_dirty = true;
  if (out) *out = uint32_t(s->sdf_len.size() - 1);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_sdf_param(mrt_scene sh, uint32_t sdf, uint32_t node, uint32_t slot, const float* values,
                            size_t count) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (count && !values)) return fail(MRT_ERR_INVALID, "null argument");
  if (sdf >= s->sdf_len.size()) return fail(MRT_ERR_INVALID, "unknown sdf");
  const uint32_t n = s->sdf_len[sdf];
  if (node >= n) return fail(MRT_ERR_INVALID, "node index out of range");
  if (count == 0 || slot >= 6 || count > 6 - size_t(slot)) return fail(MRT_ERR_INVALID, "parameter slots out of range");
  for (size_t k = 0; k < count; ++k)
    if (!std::isfinite(values[k])) return fail(MRT_ERR_INVALID, "sdf parameters must be finite");
  SdfNode* t = s->sdf_tape.data() + s->sdf_offset[sdf];
  const SdfNode old = t[node];
  for (size_t k = 0; k < count; ++k) t[node].p[slot + k] = values[k];
  Box b;
  std::string err;
  if (!sdf_bounds(t, n, b, err)) {
    t[node] = old;
    return fail(MRT_ERR_INVALID, "invalid sdf program: " + err);
  }
  std::vector<SdfInsn> prog;
  compile_sdf(t, n, prog);
  if (prog.size() != s->sdf_prog_len[sdf]) {
    t[node] = old;
    return fail(MRT_ERR_INTERNAL, "recompiled program changed length");
  }
  std::copy(prog.begin(), prog.end(), s->sdf_prog.begin() + s->sdf_prog_offset[sdf]);
  s->sdf_bounds[sdf] = b;
  SdfGrid& g = s->sdf_grid[sdf];
  const uint32_t np = uint32_t(prog.size());
  auto& rb = s->sdf_grid_rebuild;
  rb.erase(std::remove(rb.begin(), rb.end(), sdf), rb.end());
  if (!sdf_grid_applies(prog.data(), np, b)) {
    g.dim = 0;
  } else if (g.dim == uint32_t(sdf_grid_dim(np))) {
    const uint32_t off = g.mask_offset;
    sdf_grid_frame(b, np, g);
    g.mask_offset = off;
    rb.push_back(sdf);  // the masks are rebuilt on the GPU at the next prepare
  } else {
    int rc = s->sync_host_masks();
    if (rc != MRT_OK) return rc;
    build_sdf_grid(prog.data(), np, b, g, s->sdf_masks);
    s->sdf_masks_dirty = true;
  }
  bool used = false;
  for (const SdfInstance& in : s->sdf_current) used = used || in.sdf == sdf;
  if (used) build_sdf_drawables(s);
  s->sdf_dirty = true;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_copy_sdf_grid(mrt_scene sh, uint32_t sdf, float lo[3], float inv[3], uint32_t* dim, uint64_t* masks,
                            size_t n) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !dim) return fail(MRT_ERR_INVALID, "null argument");
  if (sdf >= s->sdf_len.size()) return fail(MRT_ERR_INVALID, "unknown sdf");
  const SdfGrid& g = s->sdf_grid[sdf];
  *dim = g.dim;
  for (int a = 0; a < 3; ++a) {
    if (lo) lo[a] = g.lo[a];
    if (inv) inv[a] = g.inv[a];
  }
  if (!masks || !g.dim) return MRT_OK;
  const size_t cells = size_t(g.dim) * g.dim * g.dim;
  if (n < cells) return fail(MRT_ERR_INVALID, "mask buffer too small");
  if (std::find(s->sdf_grid_rebuild.begin(), s->sdf_grid_rebuild.end(), sdf) != s->sdf_grid_rebuild.end())
    return fail(MRT_ERR_NOT_PREPARED, "the grid is rebuilt at the next prepare");
  const int rc = s->sync_host_masks();
  if (rc != MRT_OK) return rc;
  std::memcpy(masks, s->sdf_masks.data() + g.mask_offset, cells * sizeof(uint64_t));
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_enqueue_sdf(mrt_scene sh, uint32_t sdf, const float m4x4_colmajor[16], uint32_t* out) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !m4x4_colmajor) return fail(MRT_ERR_INVALID, "null argument");
  if (sdf >= s->sdf_len.size()) return fail(MRT_ERR_INVALID, "unknown sdf");
  SdfInstance in;
  in.sdf = sdf;
  std::memcpy(in.m, m4x4_colmajor, sizeof(in.m));
  if (!invert_affine_rows(in.m, in.inv)) return fail(MRT_ERR_INVALID, "transformation is not invertible");
  s->sdf_pending.push_back(in);
  if (out) *out = uint32_t(s->sdf_pending.size() - 1);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_sdf_drawables(mrt_scene sh, void* out, size_t n) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && !out)) return fail(MRT_ERR_INVALID, "null argument");
  if (n < s->sdf_drawables.size()) return fail(MRT_ERR_INVALID, "output too small");
  if (!s->sdf_drawables.empty()) std::memcpy(out, s->sdf_drawables.data(), s->sdf_drawables.size() * sizeof(SdfDrawable));
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_sdf_marching(mrt_scene sh, int steps, float eps) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || steps < 1 || !(eps > 0.f)) return fail(MRT_ERR_INVALID, "steps must be >= 1 and eps > 0");
  s->sdf_steps = steps;
  s->sdf_eps = eps;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_prepare(mrt_scene sh, mrt_prepare_result* out) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  return prepare_impl(s, out, true);
  MRT_GUARD_END
}

int mrt_scene_instance_inverses(mrt_scene sh, float* out, size_t n_inst) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !out) return fail(MRT_ERR_INVALID, "null argument");
  if (n_inst < s->current.size()) return fail(MRT_ERR_INVALID, "output too small");
  const int rc = s->sync_host_tlas();
  if (rc != MRT_OK) return rc;
  for (size_t i = 0; i < s->current.size(); ++i) std::memcpy(out + 12 * i, s->current[i].inv, 48);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_build_host(mrt_scene sh, mrt_prepare_result* out) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  return prepare_impl(s, out, false);
  MRT_GUARD_END
}

int mrt_scene_copy_wide(mrt_scene sh, void* blas_nodes, void* tris, void* tlas_nodes, uint32_t* tlas_items,
                        void* instances) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  const int rc = s->sync_host_tlas();
  if (rc != MRT_OK) return rc;
  if (blas_nodes) std::memcpy(blas_nodes, s->blas_nodes.data(), s->blas_nodes.size() * sizeof(WideNode));
  if (tris) std::memcpy(tris, s->tris.data(), s->tris.size() * sizeof(TriRecord));
  if (tlas_nodes) std::memcpy(tlas_nodes, s->tlas_nodes.data(), s->tlas_nodes.size() * sizeof(WideNode));
  if (tlas_items) std::memcpy(tlas_items, s->tlas_items.data(), s->tlas_items.size() * 4);
  if (instances) std::memcpy(instances, s->inst_records.data(), s->inst_records.size() * sizeof(InstanceRecord));
  return MRT_OK;
  MRT_GUARD_END
}

// Host-buffer queries stage through the scene's grow-only device buffers (no allocation once
// they are large enough, which matters for per-click picks).
static int trace_host(mrt_scene sh, const mrt_ray* rays, size_t n, void* out, size_t out_elem, bool any) {
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && (!rays || !out))) return fail(MRT_ERR_INVALID, "null argument");
  if (!s->prepared) return fail(MRT_ERR_NOT_PREPARED, "trace before mrt_scene_prepare");
  if (n == 0) return MRT_OK;
  const auto grow = [](void*& p, size_t& cap, size_t bytes) -> hipError_t {
    if (p && cap >= bytes) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  };
  hipError_t e = grow(s->d_host_rays, s->cap_host_rays, n * sizeof(mrt_ray));
  if (e == hipSuccess) e = grow(s->d_host_out, s->cap_host_out, n * sizeof(mrt_hit));
  if (e == hipSuccess) e = hipMemcpy(s->d_host_rays, rays, n * sizeof(mrt_ray), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "staging rays");
  int rc = s->trace(s->d_host_rays, n, s->d_host_out, any, false, nullptr);
  uint32_t flags = 0;
  if (rc == MRT_OK) rc = s->read_ctrl(&flags, nullptr, nullptr);
  if (rc == MRT_OK) {
    e = hipMemcpy(out, s->d_host_out, n * out_elem, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "reading results");
  }
  if (rc == MRT_OK && (flags & 1u)) rc = fail(MRT_ERR_INTERNAL, "traversal stack overflow");
  return rc;
}

int mrt_trace_closest(mrt_scene s, const mrt_ray* rays, size_t n, mrt_hit* hits) {
  MRT_GUARD_BEGIN
  return trace_host(s, rays, n, hits, sizeof(mrt_hit), false);
  MRT_GUARD_END
}
int mrt_trace_any(mrt_scene s, const mrt_ray* rays, size_t n, uint8_t* occluded) {
  MRT_GUARD_BEGIN
  return trace_host(s, rays, n, occluded, 1, true);
  MRT_GUARD_END
}
int mrt_trace_closest_device(mrt_scene sh, const mrt_ray* d_rays, size_t n, mrt_hit* d_hits, void* stream) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && (!d_rays || !d_hits))) return fail(MRT_ERR_INVALID, "null argument");
  return s->trace(d_rays, n, d_hits, false, false, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}
int mrt_trace_any_device(mrt_scene sh, const mrt_ray* d_rays, size_t n, uint8_t* d_occluded, void* stream) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && (!d_rays || !d_occluded))) return fail(MRT_ERR_INVALID, "null argument");
  return s->trace(d_rays, n, d_occluded, true, false, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}
int mrt_trace_stats_device(mrt_scene sh, const mrt_ray* d_rays, size_t n, int any_hit, mrt_trace_stats* out,
                           void* stream) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !out || (n && !d_rays)) return fail(MRT_ERR_INVALID, "null argument");
  void* d_o = nullptr;
  hipError_t e = hipMalloc(&d_o, std::max<size_t>(n, 1) * (any_hit ? 1 : sizeof(mrt_hit)));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = s->trace(d_rays, n, d_o, any_hit != 0, true, st);
  uint32_t flags = 0;
  unsigned long long st4[4] = {0, 0, 0, 0};
  if (rc == MRT_OK) rc = s->read_ctrl(&flags, st4, st);
  hipFree(d_o);
  if (rc != MRT_OK) return rc;
  if (flags & 1u) return fail(MRT_ERR_INTERNAL, "traversal stack overflow");
  out->rays = st4[0];
  out->node_visits = st4[1];
  out->tri_tests = st4[2];
  out->instance_visits = st4[3];
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_tlas_refit(mrt_scene sh, uint32_t max_refits) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  s->refit_limit = max_refits;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_update_transforms_device(mrt_scene sh, const float* d_m4x4_colmajor, size_t n, void* stream) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !d_m4x4_colmajor) return fail(MRT_ERR_INVALID, "null argument");
  if (!s->prepared) return fail(MRT_ERR_NOT_PREPARED, "scene not prepared");
  if (n != s->current.size()) return fail(MRT_ERR_INVALID, "one matrix per prepared instance is required");
  if (s->tlas_nodes.empty()) return fail(MRT_ERR_INVALID, "the prepared frame has no instance BVH");
  const int rc = s->refit(d_m4x4_colmajor, static_cast<hipStream_t>(stream));
  if (rc != MRT_OK) return rc;
  s->device_newer = true;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_update_status(mrt_scene sh, uint32_t* status) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !status) return fail(MRT_ERR_INVALID, "null argument");
  *status = 0;
  if (!s->d_refit_status) return MRT_OK;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(status, s->d_refit_status, 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading the refit status");
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_pick(mrt_scene s, const mrt_ray* ray, mrt_hit* hit, int* hit_found) {
  MRT_GUARD_BEGIN
  if (!ray || !hit || !hit_found) return fail(MRT_ERR_INVALID, "null argument");
  int rc = trace_host(s, ray, 1, hit, sizeof(mrt_hit), false);
  if (rc != MRT_OK) return rc;
  *hit_found = hit->inst != 0xFFFFFFFFu ? 1 : 0;
  return MRT_OK;
  MRT_GUARD_END
}

}  // extern "C"
