// Traversal-quality binary BVH for the wide collapse: binned SAH (32 bins per axis) on large
// nodes, an exact sweep over sorted centroids on small ones, and SAH leaf termination. The
// reference builder (bvh_build.cpp) stays the bit-exact mrt_bvh_build; this one only decides the
// shape of the tree the GPU walks, which does not change any hit.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "bvh_build.hpp"

namespace mrt {
namespace {

constexpr int kSahBins = 32;
constexpr uint32_t kSweepBelow = 256;  // nodes with fewer references are split by a full sweep
constexpr float kTraverseCost = 1.0f;  // SAH: cost of a node visit relative to a triangle test
constexpr float kTriCost = 1.0f;

struct Ref {
  Box box;
  vec3f c;
  uint32_t prim;
};

struct BuildNode {
  Box box;
  uint32_t left = 0, right = 0;  // children (inner) or [first, first + count) in refs (leaf)
  uint32_t first = 0, count = 0;
  bool leaf = false;
};

inline float half_area(const Box& b) {
  if (!(b.mn.x <= b.mx.x)) return 0.f;
  const vec3f d = b.dim();
  return d.x * d.y + d.y * d.z + d.z * d.x;
}

struct Builder {
  std::vector<Ref>& refs;
  uint32_t max_leaf;
  std::vector<BuildNode> nodes;

  Builder(std::vector<Ref>& r, uint32_t ml) : refs(r), max_leaf(ml) {}

  // Best split of refs[first, first + count): axis, and the position in a partition by `pred`.
  // Returns false when no split beats a leaf (only allowed for count <= max_leaf).
  bool find_split(uint32_t first, uint32_t count, const Box& box, int& axis, float& pos, bool& sweep_split,
                  uint32_t& sweep_mid) {
    Box cb;
    for (uint32_t i = first; i < first + count; ++i) cb.enclose(refs[i].c);
    const float leaf_cost = kTriCost * float(count);
    const float parent = half_area(box);
    float best = INFINITY;
    axis = -1;
    if (count <= kSweepBelow) {
      // Exact sweep: sort by centroid on each axis, evaluate every split position.
      std::vector<Ref> tmp(refs.begin() + first, refs.begin() + first + count);
      std::vector<float> right_area(count);
      for (int a = 0; a < 3; ++a) {
        if (!(cb.mx[a] > cb.mn[a])) continue;
        std::sort(tmp.begin(), tmp.end(), [a](const Ref& x, const Ref& y) {
          return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.prim < y.prim);
        });
        Box rb;
        for (uint32_t i = count; i-- > 1;) {
          rb.enclose(tmp[i].box);
          right_area[i] = half_area(rb);
        }
        Box lb;
        for (uint32_t i = 1; i < count; ++i) {
          lb.enclose(tmp[i - 1].box);
          const float cost = half_area(lb) * float(i) + right_area[i] * float(count - i);
          if (cost < best) {
            best = cost;
            axis = a;
            sweep_mid = i;
          }
        }
      }
      sweep_split = true;
    } else {
      sweep_split = false;
      for (int a = 0; a < 3; ++a) {
        const float lo = cb.mn[a], ext = cb.mx[a] - cb.mn[a];
        if (!(ext > 0.f)) continue;
        const float k = float(kSahBins) / ext;
        Box bins[kSahBins];
        uint32_t cnt[kSahBins] = {};
        for (uint32_t i = first; i < first + count; ++i) {
          const int b = std::min(kSahBins - 1, std::max(0, int((refs[i].c[a] - lo) * k)));
          bins[b].enclose(refs[i].box);
          ++cnt[b];
        }
        float ra[kSahBins];
        uint32_t rc[kSahBins];
        Box rb;
        uint32_t rn = 0;
        for (int b = kSahBins - 1; b > 0; --b) {
          rb.enclose(bins[b]);
          rn += cnt[b];
          ra[b] = half_area(rb);
          rc[b] = rn;
        }
        Box lb;
        uint32_t ln = 0;
        for (int b = 1; b < kSahBins; ++b) {
          lb.enclose(bins[b - 1]);
          ln += cnt[b - 1];
          if (ln == 0 || rc[b] == 0) continue;
          const float cost = half_area(lb) * float(ln) + ra[b] * float(rc[b]);
          if (cost < best) {
            best = cost;
            axis = a;
            pos = lo + float(b) / k;
          }
        }
      }
    }
    if (axis < 0) return false;
    const float split_cost = kTraverseCost + kTriCost * best / std::max(parent, 1e-30f);
    return !(count <= max_leaf && leaf_cost <= split_cost);
  }

  uint32_t build(uint32_t first, uint32_t count, int depth) {
    Box box;
    for (uint32_t i = first; i < first + count; ++i) box.enclose(refs[i].box);
    uint32_t id;
    {
      id = uint32_t(nodes.size());
      nodes.emplace_back();
    }
    nodes[id].box = box;
    int axis = -1;
    float pos = 0.f;
    bool sweep = false;
    uint32_t mid = 0;
    // Past depth 48 (degenerate inputs) the node is halved at its median centroid instead, which
    // bounds the recursion.
    bool split = depth < 48 && count > 1 && find_split(first, count, box, axis, pos, sweep, mid);
    if (depth >= 48 && count > max_leaf) {
      Box cb;
      for (uint32_t i = first; i < first + count; ++i) cb.enclose(refs[i].c);
      const vec3f d = cb.dim();
      const int a = (d.x >= d.y && d.x >= d.z) ? 0 : (d.y >= d.z ? 1 : 2);
      mid = count / 2;
      std::nth_element(refs.begin() + first, refs.begin() + first + mid, refs.begin() + first + count,
                       [a](const Ref& x, const Ref& y) { return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.prim < y.prim); });
      split = true;
      sweep = true;  // already partitioned
      axis = -1;
    }
    if (split) {
      if (sweep && axis >= 0) {
        std::sort(refs.begin() + first, refs.begin() + first + count, [axis](const Ref& x, const Ref& y) {
          return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.prim < y.prim);
        });
      } else if (!sweep) {
        const auto it = std::partition(refs.begin() + first, refs.begin() + first + count,
                                       [axis, pos](const Ref& r) { return r.c[axis] < pos; });
        mid = uint32_t(it - (refs.begin() + first));
        if (mid == 0 || mid == count) split = false;
      }
    }
    if (!split && count > max_leaf) {
      // Coincident centroids (or no split found): halve by index.
      mid = count / 2;
      split = true;
    }
    if (!split) {
      nodes[id].leaf = true;
      nodes[id].first = first;
      nodes[id].count = count;
      return id;
    }
    const uint32_t l = build(first, mid, depth + 1);
    const uint32_t r = build(first + mid, count - mid, depth + 1);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
};

}  // namespace

BinaryBvh build_sah_bvh(const std::vector<Box>& boxes, uint32_t max_leaf, unsigned threads) {
  BinaryBvh out;
  if (boxes.empty()) return out;
  std::vector<Ref> refs;
  refs.reserve(boxes.size());
  for (uint32_t i = 0; i < boxes.size(); ++i) {
    const Box& b = boxes[i];
    if (!(b.mn.x <= b.mx.x && b.mn.y <= b.mx.y && b.mn.z <= b.mx.z)) continue;  // empty: never hit
    refs.push_back(Ref{b, b.center(), i});
  }
  if (refs.empty()) {
    // One empty leaf keeps the layout valid.
    BinNode n{};
    n.type_and_parent = 0x80000000u;
    n.first = 0;
    n.second = 0;
    for (int a = 0; a < 3; ++a) {
      n.mn[a] = 0.f;
      n.mx[a] = 0.f;
    }
    out.nodes.push_back(n);
    out.indices.push_back(0);
    out.ref_boxes.push_back(boxes[0]);
    return out;
  }
  if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // Split the top of the tree serially into up to 4 * threads independent ranges, then build
  // those subtrees on a thread pool and splice them in.
  Builder top(refs, max_leaf);
  struct Pending {
    uint32_t id, first, count;
  };
  std::vector<Pending> work;
  {
    // Serial top: breadth-first until enough independent ranges exist.
    std::vector<Pending> q;
    top.nodes.emplace_back();
    q.push_back({0, 0, uint32_t(refs.size())});
    const size_t want = size_t(threads) * 4;
    size_t head = 0;
    while (head < q.size() && q.size() - head < want) {
      const Pending p = q[head++];
      Box box;
      for (uint32_t i = p.first; i < p.first + p.count; ++i) box.enclose(refs[i].box);
      top.nodes[p.id].box = box;
      int axis = -1;
      float pos = 0.f;
      bool sweep = false;
      uint32_t mid = 0;
      bool split = p.count > std::max<uint32_t>(max_leaf, 4096) &&
                   top.find_split(p.first, p.count, box, axis, pos, sweep, mid);
      if (split && !sweep) {
        const auto it = std::partition(refs.begin() + p.first, refs.begin() + p.first + p.count,
                                       [axis, pos](const Ref& r) { return r.c[axis] < pos; });
        mid = uint32_t(it - (refs.begin() + p.first));
        if (mid == 0 || mid == p.count) split = false;
      }
      if (!split) {
        work.push_back(p);
        continue;
      }
      const uint32_t l = uint32_t(top.nodes.size());
      top.nodes.emplace_back();
      top.nodes.emplace_back();
      top.nodes[p.id].left = l;
      top.nodes[p.id].right = l + 1;
      q.push_back({l, p.first, mid});
      q.push_back({l + 1, p.first + mid, p.count - mid});
    }
    for (size_t i = head; i < q.size(); ++i) work.push_back(q[i]);
  }
  // Parallel subtrees.
  std::vector<Builder> sub;
  sub.reserve(work.size());
  for (size_t i = 0; i < work.size(); ++i) sub.emplace_back(refs, max_leaf);
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= work.size()) break;
      sub[k].build(work[k].first, work[k].count, 0);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  // Splice: map every subtree node to a global id; the subtree root replaces its placeholder.
  std::vector<BuildNode> all = top.nodes;
  std::vector<uint32_t> root_of(work.size());
  for (size_t k = 0; k < work.size(); ++k) {
    const uint32_t base = uint32_t(all.size());
    const std::vector<BuildNode>& sn = sub[k].nodes;
    // Subtree node 0 is its root: copy it into the placeholder, the rest after `base`.
    auto remap = [&](uint32_t j) { return j == 0 ? work[k].id : base + j - 1; };
    for (size_t j = 1; j < sn.size(); ++j) all.emplace_back();
    for (size_t j = 0; j < sn.size(); ++j) {
      BuildNode n = sn[j];
      if (!n.leaf) {
        n.left = remap(n.left);
        n.right = remap(n.right);
      }
      all[remap(uint32_t(j))] = n;
    }
  }
  // Emit BinaryBvh: leaves reference contiguous ranges of `indices` in depth-first order.
  out.nodes.resize(all.size());
  out.indices.reserve(refs.size());
  out.ref_boxes.reserve(refs.size());
  std::vector<uint32_t> parent(all.size(), 0);
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};  // node, depth
  while (!stack.empty()) {
    const auto [id, depth] = stack.back();
    stack.pop_back();
    out.max_depth = std::max(out.max_depth, depth);
    const BuildNode& n = all[id];
    BinNode& b = out.nodes[id];
    for (int a = 0; a < 3; ++a) {
      b.mn[a] = n.box.mn[a];
      b.mx[a] = n.box.mx[a];
    }
    b.pad[0] = b.pad[1] = b.pad[2] = 0;
    if (n.leaf) {
      b.type_and_parent = 0x80000000u | (parent[id] & 0x7fffffffu);
      b.first = uint32_t(out.indices.size());
      for (uint32_t i = n.first; i < n.first + n.count; ++i) {
        out.indices.push_back(refs[i].prim);
        out.ref_boxes.push_back(refs[i].box);
      }
      b.second = uint32_t(out.indices.size()) - 1;
    } else {
      b.type_and_parent = parent[id] & 0x7fffffffu;
      b.first = n.left;
      b.second = n.right;
      parent[n.left] = id;
      parent[n.right] = id;
      stack.push_back({n.right, depth + 1});
      stack.push_back({n.left, depth + 1});
    }
  }
  return out;
}

}  // namespace mrt
