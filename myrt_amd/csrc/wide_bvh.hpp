// Collapse of the reference-exact binary BVH into 4-wide, 64-byte quantised nodes.
#pragma once
#include <vector>

#include "bvh_build.hpp"
#include "device_types.hpp"

namespace mrt {

struct WideBvh {
  std::vector<WideNode> nodes;   // root = 0
  std::vector<uint32_t> items;   // leaf item slots -> position in BinaryBvh::indices
  uint32_t depth = 0;            // wide nodes on the longest root-to-leaf path
};

// pad_rel widens each child box by pad_rel * (largest |coordinate| of the box) on every side
// before quantisation; decoded boxes additionally end strictly outside the padded box. Binary
// leaves with more than max_leaf items are split further at centroid medians, so wide leaves hold
// at most max_leaf items (the binary BVH itself is left as built).
WideBvh collapse_to_wide(const BinaryBvh& b, double pad_rel, uint32_t max_leaf = 127);

// Decoded child box, exactly as the kernels compute it (fmaf(code, 2^e, origin)).
void decode_child_box(const WideNode& n, int c, float lo[3], float hi[3]);

}  // namespace mrt
