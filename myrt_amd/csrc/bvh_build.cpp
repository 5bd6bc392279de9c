// Binned-SAH BVH builder, bit-identical in node order, node boxes and reordered indices to the
// reference builder (src/pathtracer/bvh.cpp:43-116, 154-161, 356-512).
//
// The reference keeps one global reference array, inserts spatial-split pieces into it and
// shifts the ranges of every later leaf after each split. Each node's references always form a
// contiguous range that its two children subdivide, so the final array is the left-to-right
// concatenation of the leaves. This builder therefore gives every node its own reference list,
// splits nodes independently (in parallel across subtrees), numbers the nodes breadth-first
// exactly as the reference's FIFO processing does, and assigns leaf ranges by an in-order walk.
#include "bvh_build.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <thread>

namespace mrt {
namespace {

constexpr int kBins = 16;               // bvh::binned_sah_bin_count (bvh.hpp:167)
constexpr float kCostTraverse = 7.f;    // bvh.cpp:84
constexpr float kCostIntersect = 2.f;   // bvh.cpp:85

// float -> int as the x86 truncating conversion does it: NaN and out-of-range give INT_MIN.
// The reference converts centroid offsets with rnu::vec3i(float3) (bvh.cpp:61-62); a zero
// centroid extent makes the quotient NaN or infinite, which then clamps to bin 0.
inline int trunc_i32(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
  return static_cast<int>(f);
}

struct Ref {
  Box box;
  uint32_t idx;
};

struct Node {
  Box box;
  std::vector<Ref> refs;  // leaves only, after the build
  std::unique_ptr<Node> left, right;
  uint32_t depth = 0;
};

struct SplitChoice {
  int axis;
  float plane;
  bool split;
};

// bvh::compute_split_axis (bvh.cpp:43-116), including its quirks: the second-side count is the
// prefix count accum_counts[15 - i], the second-side box includes bin i, ties go to the last
// (axis, bin), and the split plane is min + step * (i + 1).
SplitChoice choose_split(const std::vector<Ref>& refs, const Box& node_box) {
  Box sub;
  for (const Ref& r : refs) sub.enclose(r.box.centroid());
  const vec3f size = sub.dim();
  const vec3f step{size.x / float(kBins), size.y / float(kBins), size.z / float(kBins)};

  Box bins[3][kBins];
  int counts[3][kBins] = {};
  for (const Ref& r : refs) {
    const vec3f c = r.box.centroid();
    for (int a = 0; a < 3; ++a) {
      const float rel = c[a] - sub.mn[a];
      const float q = rel / step[a];
      const int b = std::min(std::max(trunc_i32(q), 0), kBins - 1);
      bins[a][b].enclose(r.box);
      counts[a][b]++;
    }
  }

  Box acc[3][kBins], accr[3][kBins];
  int acc_counts[3][kBins];
  for (int a = 0; a < 3; ++a) {
    acc[a][0] = bins[a][0];
    acc_counts[a][0] = counts[a][0];
    for (int i = 1; i < kBins; ++i) {
      acc[a][i] = acc[a][i - 1];
      acc[a][i].enclose(bins[a][i]);
      acc_counts[a][i] = acc_counts[a][i - 1] + counts[a][i];
    }
    accr[a][kBins - 1] = bins[a][kBins - 1];
    for (int i = kBins - 2; i >= 0; --i) {
      accr[a][i] = accr[a][i + 1];
      accr[a][i].enclose(bins[a][i]);
    }
  }

  const float total_area = node_box.area();
  float best_cost = std::numeric_limits<float>::max();
  int best_axis = 0, best_part = 0;
  for (int a = 0; a < 3; ++a) {
    for (int i = 0; i < kBins - 1; ++i) {
      const int cf = acc_counts[a][i];
      const int cs = acc_counts[a][kBins - 1 - i];
      const float af = acc[a][i].area();
      const float as = accr[a][i].area();
      const float lhs = float(cf) * (af / total_area);
      const float rhs = float(cs) * (as / total_area);
      const float cost = kCostTraverse + kCostIntersect * (lhs + rhs);
      if (cost <= best_cost) {
        best_axis = a;
        best_part = i;
        best_cost = cost;
      }
    }
  }
  const float total_count = float(acc_counts[best_axis][kBins - 1]);
  const float no_split_cost = kCostIntersect * total_count;
  const float plane = sub.mn[best_axis] + step[best_axis] * float(best_part + 1);
  return {best_axis, plane, best_cost < no_split_cost};
}

// bvh::split (bvh.cpp:408-512) on a node-local reference list. Returns false for a leaf.
bool split_node(Node& node, const Box& full, bool split_aabbs) {
  const size_t count = node.refs.size();
  if (count < 2) return false;  // second_child - first_child < min_leaf_primitives
  const SplitChoice ch = choose_split(node.refs, node.box);
  if (!ch.split) return false;
  const int axis = ch.axis;
  const float plane = ch.plane;
  const float thres = 0.05f * full.dim()[axis];

  // Spatial "box split" of wide references (bvh.cpp:431-459, split_aabbs at :33-41).
  std::vector<Ref> pieces;
  for (Ref& r : node.refs) {
    std::optional<Box> lower, higher;
    if (r.box.dim()[axis] > thres) {
      if (!split_aabbs) {
        lower = r.box;
      } else if (plane <= r.box.mn[axis]) {
        higher = r.box;
      } else if (plane >= r.box.mx[axis]) {
        lower = r.box;
      } else {
        Box lo, hi;
        lo.mn = r.box.mn;
        lo.mx = r.box.mx;
        lo.mx[axis] = plane;
        hi.mn = r.box.mn;
        hi.mn[axis] = plane;
        hi.mx = r.box.mx;
        lower = lo;
        higher = hi;
      }
    } else {
      lower = r.box;
    }
    if ((higher && higher->dim()[axis] < thres) || (lower && lower->dim()[axis] < thres)) {
      higher.reset();
      lower = r.box;
    }
    if (lower && !higher) {
      r.box = *lower;
    } else if (higher && !lower) {
      r.box = *higher;
    } else if (higher && lower) {
      r.box = *higher;
      pieces.push_back(Ref{*lower, r.idx});
    }
  }

  // The lower pieces go in front of the node's range (bvh.cpp:460-465).
  std::vector<Ref> all;
  all.reserve(pieces.size() + count);
  all.insert(all.end(), pieces.begin(), pieces.end());
  all.insert(all.end(), node.refs.begin(), node.refs.end());
  node.refs.clear();
  node.refs.shrink_to_fit();

  // Partition (bvh.cpp:477-486). The scan stops at the node's old end, so when every element
  // up to it passes, the inserted tail is left unpartitioned, as in the reference.
  const auto goes_left = [&](const Ref& r) { return r.box.centroid()[axis] < plane; };
  const size_t old_last = count;
  const size_t end = all.size();
  size_t first = 0;
  while (first != old_last && goes_left(all[first])) ++first;
  if (first != old_last) {
    for (size_t i = first; i != end; ++i) {
      if (goes_left(all[i])) {
        std::swap(all[i], all[first]);
        ++first;
      }
    }
  }
  const size_t split_index = first == 0 ? 1 : first;  // bvh.cpp:494
  if (split_index >= end) {
    // Every reference went left. This requires a split plane above the largest centroid,
    // which the plane formula cannot produce; refuse rather than emit an empty leaf.
    throw std::runtime_error("bvh build: degenerate split with an empty right child");
  }

  node.left = std::make_unique<Node>();
  node.right = std::make_unique<Node>();
  node.left->refs.assign(all.begin(), all.begin() + split_index);
  node.right->refs.assign(all.begin() + split_index, all.end());
  for (const Ref& r : node.left->refs) node.left->box.enclose(r.box);
  for (const Ref& r : node.right->refs) node.right->box.enclose(r.box);
  node.left->depth = node.right->depth = node.depth + 1;
  return true;
}

// Splits every node of the tree below `root`, with a small pool of threads popping pending nodes.
void build_tree(Node* root, const Box& full, bool split_aabbs, unsigned threads) {
  std::vector<Node*> work{root};
  std::mutex m;
  std::condition_variable cv;
  size_t busy = 0;
  std::exception_ptr error;

  const auto worker = [&]() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      cv.wait(lk, [&] { return !work.empty() || busy == 0 || error; });
      if (error || (work.empty() && busy == 0)) return;
      Node* n = work.back();
      work.pop_back();
      ++busy;
      lk.unlock();
      bool did = false;
      try {
        did = split_node(*n, full, split_aabbs);
      } catch (...) {
        lk.lock();
        if (!error) error = std::current_exception();
        --busy;
        cv.notify_all();
        return;
      }
      lk.lock();
      if (did) {
        work.push_back(n->right.get());
        work.push_back(n->left.get());
      }
      --busy;
      cv.notify_all();
    }
  };

  if (threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
  }
  if (error) std::rethrow_exception(error);
}

}  // namespace

std::vector<Box> triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3) {
  std::vector<Box> out(n_idx / 3);
  for (size_t i = 0; i + 2 < n_idx; i += 3) {
    vec3f p[3];
    for (int k = 0; k < 3; ++k) {
      const float* q = pos3 + size_t(idx[i + k]) * 3;
      p[k] = {q[0], q[1], q[2]};
    }
    Box& b = out[i / 3];
    b.enclose(p[0]);
    b.enclose(p[1]);
    b.enclose(p[2]);
    b.mn = {b.mn.x - 1e-5f, b.mn.y - 1e-5f, b.mn.z - 1e-5f};
    b.mx = {b.mx.x + 1e-5f, b.mx.y + 1e-5f, b.mx.z + 1e-5f};
    b.wc = {(p[0].x + p[1].x + p[2].x) / 3.0f, (p[0].y + p[1].y + p[2].y) / 3.0f,
            (p[0].z + p[1].z + p[2].z) / 3.0f};
    b.has_wc = true;
  }
  return out;
}

BinaryBvh build_binary_bvh(const std::vector<Box>& boxes, bool split_aabbs, unsigned threads) {
  if (boxes.empty()) throw std::invalid_argument("bvh build: no primitives");
  if (boxes.size() >= 0x7fffffffu) throw std::invalid_argument("bvh build: too many primitives");
  Box full;
  for (const Box& b : boxes) full.enclose(b);

  auto root = std::make_unique<Node>();
  root->box.mn = full.mn;
  root->box.mx = full.mx;
  root->refs.resize(boxes.size());
  for (size_t i = 0; i < boxes.size(); ++i) root->refs[i] = Ref{boxes[i], uint32_t(i)};

  if (threads == 0) threads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  if (boxes.size() < 4096) threads = 1;
  build_tree(root.get(), full, split_aabbs, threads);

  // Breadth-first numbering: the reference processes nodes in creation order and appends both
  // children of a split node at the end (bvh.cpp:380-400).
  BinaryBvh out;
  std::vector<Node*> order{root.get()};
  std::vector<uint32_t> parent{0};
  for (size_t i = 0; i < order.size(); ++i) {
    Node* n = order[i];
    if (n->left) {
      order.push_back(n->left.get());
      order.push_back(n->right.get());
      parent.push_back(uint32_t(i));
      parent.push_back(uint32_t(i));
    }
  }
  out.nodes.resize(order.size());
  std::vector<uint32_t> index_of_child(order.size());
  {
    size_t next = 1;
    for (size_t i = 0; i < order.size(); ++i) {
      Node* n = order[i];
      BinNode& bn = out.nodes[i];
      bn = BinNode{};
      bn.mn[0] = n->box.mn.x; bn.mn[1] = n->box.mn.y; bn.mn[2] = n->box.mn.z;
      bn.mx[0] = n->box.mx.x; bn.mx[1] = n->box.mx.y; bn.mx[2] = n->box.mx.z;
      const bool leaf = !n->left;
      bn.type_and_parent = (uint32_t(leaf) << 31) | (parent[i] & 0x7fffffffu);
      if (!leaf) {
        bn.first = uint32_t(next);
        bn.second = uint32_t(next + 1);
        next += 2;
      }
      out.max_depth = std::max(out.max_depth, n->depth);
    }
  }
  // In-order walk assigns the leaf ranges of the reordered index array.
  size_t total = 0;
  for (Node* n : order)
    if (!n->left) total += n->refs.size();
  out.indices.reserve(total);
  out.ref_boxes.reserve(total);
  std::vector<std::pair<Node*, uint32_t>> stack{{root.get(), 0u}};
  while (!stack.empty()) {
    auto [n, id] = stack.back();
    stack.pop_back();
    if (n->left) {
      const BinNode& bn = out.nodes[id];
      stack.push_back({n->right.get(), bn.second});
      stack.push_back({n->left.get(), bn.first});
    } else {
      BinNode& bn = out.nodes[id];
      bn.first = uint32_t(out.indices.size());
      bn.second = uint32_t(out.indices.size() + n->refs.size() - 1);
      for (const Ref& r : n->refs) {
        out.indices.push_back(r.idx);
        out.ref_boxes.push_back(r.box);
      }
    }
  }
  return out;
}

}  // namespace mrt
