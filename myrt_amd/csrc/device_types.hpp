// Records shared by the host collapse code and the HIP traversal kernels.
#pragma once
#include <cstdint>

namespace mrt {

// 4-wide BVH node quantised to one 64-byte line.
//   Child c's box on axis a is [fma(qlo[a][c], 2^e[a], origin[a]), fma(qhi[a][c], 2^e[a], origin[a])];
//   the host rounds the codes outward and checks the decoded floats, so every decoded box
//   strictly contains the child's exact box.
//   entry[c] is the child's stack entry, so the traversal pushes it without decoding:
//   bit 31 clear: inner child, node index entry[c];
//   bit 31 set:   leaf child of (entry[c] >> 24) & 0x7f items starting at entry[c] & 0xffffff
//                 (triangle records in a BLAS, instance indices in the TLAS).
//   Children 0 .. nchildren-1 are valid. An unused slot holds an inverted box (low codes 255,
//   high codes 0), which the slab test rejects unless the node is smaller than about 2^-19 of
//   its distance from the ray origin, and the entry of a leaf of no items, which is harmless if
//   visited; so the traversal needs no child count.
struct alignas(64) WideNode {
  float origin[3];
  int8_t exp[3];
  uint8_t nchildren;
  uint8_t qlo[3][4];
  uint8_t qhi[3][4];
  uint32_t pad[2];
  uint32_t entry[4];  // each child's stack entry: an inner child's node index, or for a leaf
                      // 0x80000000 | item count << 24 | first item (the traversal pushes it as is)
};
static_assert(sizeof(WideNode) == 64, "wide node must be one 64-byte line");

// Triangle record, 48 bytes: three float4 with the triangle's index in its geometry in v0.w.
struct alignas(16) TriRecord {
  float v0[3];
  uint32_t prim;
  float v1[3];
  uint32_t pad1;
  float v2[3];
  uint32_t pad2;
};
static_assert(sizeof(TriRecord) == 48, "triangle record must be 48 bytes");

// Instance record, 64 bytes: world -> object transform (rows of the inverse 3x4 matrix) and
// the wide-node index of the instance's geometry BVH root.
struct alignas(16) InstanceRecord {
  float inv[12];
  uint32_t blas_root;
  uint32_t geom;
  uint32_t inst;
  uint32_t pad;
};
static_assert(sizeof(InstanceRecord) == 64, "instance record must be 64 bytes");

constexpr uint32_t kMaxLeafItems = 127;          // 7-bit count in a stack entry
constexpr uint32_t kMaxItemIndex = (1u << 24) - 1;  // 24-bit item start in a stack entry
constexpr uint32_t kStackReturn = 0xFFFFFFFFu;   // leave the instance, restore the world ray

}  // namespace mrt
