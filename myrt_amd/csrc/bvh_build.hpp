// Host-side binned-SAH BVH builder whose output is bit-identical to the reference builder
// (src/pathtracer/bvh.cpp:43-116, 136-161, 356-512).
#pragma once
#include <cstdint>
#include <vector>
#include <limits>

namespace mrt {

struct vec3f {
  float x, y, z;
  float operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
  float& operator[](int a) { return a == 0 ? x : (a == 1 ? y : z); }
};

// The reference's vector min/max (rnu) are taken with std::min / std::max semantics:
// min(a, b) = (b < a) ? b : a, max(a, b) = (a < b) ? b : a, component-wise.
inline float ref_min(float a, float b) { return (b < a) ? b : a; }
inline float ref_max(float a, float b) { return (a < b) ? b : a; }

// aabb_t (src/pathtracer/bvh.hpp:25-86). An empty box is (+FLT_MAX, lowest).
struct Box {
  vec3f mn{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
  vec3f mx{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
           std::numeric_limits<float>::lowest()};
  vec3f wc{0, 0, 0};
  bool has_wc = false;

  vec3f dim() const { return {mx.x - mn.x, mx.y - mn.y, mx.z - mn.z}; }
  vec3f center() const { return {(mx.x + mn.x) * 0.5f, (mx.y + mn.y) * 0.5f, (mx.z + mn.z) * 0.5f}; }
  vec3f centroid() const { return has_wc ? wc : center(); }
  float area() const {
    const vec3f d = dim();
    return ((d.x + d.y) * d.z + (d.x * d.y)) * 2.f;
  }
  void enclose(vec3f p) {
    mn = {ref_min(mn.x, p.x), ref_min(mn.y, p.y), ref_min(mn.z, p.z)};
    mx = {ref_max(mx.x, p.x), ref_max(mx.y, p.y), ref_max(mx.z, p.z)};
  }
  void enclose(const Box& b) {
    mn = {ref_min(mn.x, b.mn.x), ref_min(mn.y, b.mn.y), ref_min(mn.z, b.mn.z)};
    mx = {ref_max(mx.x, b.mx.x), ref_max(mx.y, b.mx.y), ref_max(mx.z, b.mx.z)};
  }
};

// Node record identical to aligned_node_t (src/pathtracer/bvh.hpp:98-139), 48 bytes.
struct BinNode {
  float mn[3];
  uint32_t type_and_parent;
  float mx[3];
  uint32_t first;
  uint32_t second;
  uint32_t pad[3];
  bool is_leaf() const { return (type_and_parent & 0x80000000u) != 0; }
};
static_assert(sizeof(BinNode) == 48, "binary node must be 48 bytes");

struct BinaryBvh {
  std::vector<BinNode> nodes;       // breadth-first (queue) order, root = 0
  std::vector<uint32_t> indices;    // reordered primitive indices (duplicates from spatial splits)
  std::vector<Box> ref_boxes;       // final (possibly clipped) box of each entry of `indices`
  uint32_t max_depth = 0;
};

// generate_triangle_bounds (src/pathtracer/bvh.cpp:136-152).
std::vector<Box> triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3);

// bvh::bvh(aabbs, split) (src/pathtracer/bvh.cpp:154-161). split_aabbs: box-clip wide references.
BinaryBvh build_binary_bvh(const std::vector<Box>& boxes, bool split_aabbs, unsigned threads = 0);

// Traversal-quality BVH (sah_build.cpp): binned/sweep SAH with leaves of at most max_leaf
// references; no reference duplication. Used for the trees the GPU walks.
BinaryBvh build_sah_bvh(const std::vector<Box>& boxes, uint32_t max_leaf, unsigned threads = 0);

}  // namespace mrt
