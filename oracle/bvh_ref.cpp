// Oracle (test infrastructure only): literal restatement of the reference BVH builder.
//
// Follows src/pathtracer/bvh.cpp step by step: one global box/index array, lower pieces of
// spatial splits inserted at the node's first slot (bvh.cpp:460-465), every leaf's range shifted
// after a split (bvh.cpp:467-488), nodes processed in creation order (bvh.cpp:380-400).
// Vector min/max follow std::min/std::max; float->int follows the x86 truncating conversion.
#include <algorithm>
#include <climits>
#include <cstring>
#include <limits>
#include <optional>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

struct V3 {
  float v[3];
};
inline float mn_(float a, float b) { return (b < a) ? b : a; }
inline float mx_(float a, float b) { return (a < b) ? b : a; }

// aabb_t (bvh.hpp:25-86)
struct AB {
  V3 lo{{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()}};
  V3 hi{{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
         std::numeric_limits<float>::lowest()}};
  std::optional<V3> wc;

  V3 dimension() const { return {{hi.v[0] - lo.v[0], hi.v[1] - lo.v[1], hi.v[2] - lo.v[2]}}; }
  V3 center() const {
    return {{(hi.v[0] + lo.v[0]) * 0.5f, (hi.v[1] + lo.v[1]) * 0.5f, (hi.v[2] + lo.v[2]) * 0.5f}};
  }
  V3 centroid() const { return wc ? *wc : center(); }
  float surface_area() const {
    V3 d = dimension();
    return ((d.v[0] + d.v[1]) * d.v[2] + (d.v[0] * d.v[1])) * 2.f;
  }
  void enclose_point(const V3& p) {
    for (int i = 0; i < 3; ++i) {
      lo.v[i] = mn_(lo.v[i], p.v[i]);
      hi.v[i] = mx_(hi.v[i], p.v[i]);
    }
  }
  void enclose_box(const AB& b) {
    for (int i = 0; i < 3; ++i) {
      lo.v[i] = mn_(lo.v[i], b.lo.v[i]);
      hi.v[i] = mx_(hi.v[i], b.hi.v[i]);
    }
  }
  // aabb_t::split (bvh.hpp:53-67)
  std::pair<std::optional<AB>, std::optional<AB>> split(int axis, float position) const {
    if (position <= lo.v[axis]) return {std::nullopt, *this};
    if (position >= hi.v[axis]) return {*this, std::nullopt};
    AB a, b;
    a.lo = lo;
    a.hi = hi;
    a.hi.v[axis] = position;
    b.lo = lo;
    b.lo.v[axis] = position;
    b.hi = hi;
    return {a, b};
  }
};

int x86_trunc(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
  return (int)f;
}

struct Node {  // bvh_node_t (bvh.hpp:98-134)
  V3 lo, hi;
  uint32_t tp = 0x80000000u;
  uint32_t first = 0, second = 0;
  bool leaf() const { return (tp & 0x80000000u) != 0; }
  void set_parent(uint32_t p) { tp = (tp & 0x80000000u) | (p & 0x7fffffffu); }
  void make_leaf(bool l) { tp = (tp & 0x7fffffffu) | (uint32_t(l) << 31); }
};

struct State {  // build_state_t (bvh.cpp:18-27)
  AB full;
  std::vector<AB> boxes;
  std::vector<uint32_t> indices;
};

struct Builder {
  std::vector<Node> nodes;
  std::vector<uint32_t> reordered;
  bool split_aabbs = false;

  // bvh::compute_split_axis (bvh.cpp:43-116)
  void split_axis(const Node& node, State& st, int& axis_out, float& plane_out, bool& should) {
    AB sub;
    for (uint32_t i = node.first; i <= node.second; ++i) sub.enclose_point(st.boxes[i].centroid());
    V3 size = sub.dimension();
    V3 step{{size.v[0] / 16.0f, size.v[1] / 16.0f, size.v[2] / 16.0f}};
    AB bins[3][16];
    int cnt[3][16];
    std::memset(cnt, 0, sizeof(cnt));
    for (uint32_t i = node.first; i <= node.second; ++i) {
      V3 c = st.boxes[i].centroid();
      int id[3];
      for (int k = 0; k < 3; ++k) {
        float rel = c.v[k] - sub.lo.v[k];
        float q = rel / step.v[k];
        int b = x86_trunc(q);
        id[k] = b < 0 ? 0 : (b > 15 ? 15 : b);
      }
      for (int k = 0; k < 3; ++k) {
        bins[k][id[k]].enclose_box(st.boxes[i]);
        cnt[k][id[k]]++;
      }
    }
    AB acc[3][16], accr[3][16];
    int acc_cnt[3][16];
    for (int a = 0; a < 3; ++a) {
      AB run = bins[a][0];
      acc[a][0] = run;
      for (int i = 1; i < 16; ++i) {
        run.enclose_box(bins[a][i]);
        acc[a][i] = run;
      }
      AB runr = bins[a][15];
      accr[a][15] = runr;
      for (int i = 14; i >= 0; --i) {
        runr.enclose_box(bins[a][i]);
        accr[a][i] = runr;
      }
      int s = 0;
      for (int i = 0; i < 16; ++i) {
        s += cnt[a][i];
        acc_cnt[a][i] = s;
      }
    }
    AB nb;
    nb.lo = node.lo;
    nb.hi = node.hi;
    const float total_area = nb.surface_area();
    float best = std::numeric_limits<float>::max();
    int best_axis = 0, best_i = 0;
    for (int a = 0; a < 3; ++a) {
      for (int i = 0; i < 15; ++i) {
        int c1 = acc_cnt[a][i];
        int c2 = acc_cnt[a][15 - i];
        float a1 = acc[a][i].surface_area();
        float a2 = accr[a][i].surface_area();
        float cost = 7.f + 2.f * (float(c1) * (a1 / total_area) + float(c2) * (a2 / total_area));
        if (cost <= best) {
          best_axis = a;
          best_i = i;
          best = cost;
        }
      }
    }
    const float total = float(acc_cnt[best_axis][15]);
    axis_out = best_axis;
    plane_out = sub.lo.v[best_axis] + step.v[best_axis] * float(best_i + 1);
    should = best < 2.f * total;
  }

  // bvh::split (bvh.cpp:408-512)
  bool split(uint32_t node_index, State& st, Node& left, Node& right) {
    const Node cur = nodes[node_index];
    if (cur.second - cur.first < 1u) return false;
    int axis;
    float plane;
    bool should;
    split_axis(cur, st, axis, plane, should);
    if (!should) return false;
    auto test = [&](size_t i) { return st.boxes[i].centroid().v[axis] < plane; };
    uint32_t first = cur.first;
    const uint32_t last = cur.second + 1;
    std::vector<AB> ins_boxes;
    std::vector<uint32_t> ins_idx;
    int nsplit = 0;
    for (uint32_t i = first; i != last; ++i) {
      const float thres = 0.05f * st.full.dimension().v[axis];
      std::optional<AB> lower, higher;
      if (st.boxes[i].dimension().v[axis] > thres) {
        if (split_aabbs) {
          auto pr = st.boxes[i].split(axis, plane);
          lower = pr.first;
          higher = pr.second;
        } else {
          lower = st.boxes[i];
        }
      } else {
        lower = st.boxes[i];
      }
      if ((higher && higher->dimension().v[axis] < thres) || (lower && lower->dimension().v[axis] < thres)) {
        higher.reset();
        lower = st.boxes[i];
      }
      if (lower && !higher) {
        st.boxes[i] = *lower;
      } else if (higher && !lower) {
        st.boxes[i] = *higher;
      } else if (higher && lower) {
        uint32_t index = st.indices[i];
        st.boxes[i] = *higher;
        ins_boxes.push_back(*lower);
        ins_idx.push_back(index);
        ++nsplit;
      }
    }
    if (!ins_boxes.empty()) {
      st.boxes.insert(st.boxes.begin() + first, ins_boxes.begin(), ins_boxes.end());
      st.indices.insert(st.indices.begin() + first, ins_idx.begin(), ins_idx.end());
    }
    while (first != last && test(first)) first++;
    if (first != last) {
      for (uint32_t i = first; i != last + uint32_t(nsplit); ++i) {
        if (test(i)) {
          std::swap(st.indices[i], st.indices[first]);
          std::swap(st.boxes[i], st.boxes[first]);
          ++first;
        }
      }
    }
    if (nsplit > 0) {
      for (Node& n : nodes) {
        if (n.leaf()) {
          if (n.first >= last) n.first += uint32_t(nsplit);
          if (n.second >= last) n.second += uint32_t(nsplit);
        }
      }
    }
    const uint32_t split_index = first == cur.first ? first + 1 : first;
    left = Node{};
    right = Node{};
    left.first = cur.first;
    left.second = split_index - 1;
    right.first = split_index;
    right.second = cur.second + uint32_t(nsplit);
    AB la, ra;
    for (uint32_t i = left.first; i <= left.second; ++i) la.enclose_box(st.boxes[i]);
    for (uint32_t i = right.first; i <= right.second; ++i) ra.enclose_box(st.boxes[i]);
    left.lo = la.lo;
    left.hi = la.hi;
    right.lo = ra.lo;
    right.hi = ra.hi;
    return true;
  }

  // bvh::bvh + bvh::create (bvh.cpp:154-161, 356-407)
  void build(const std::vector<AB>& input, State& st) {
    st.boxes = input;
    for (const AB& b : input) st.full.enclose_box(b);
    reordered.resize(input.size());
    for (size_t i = 0; i < input.size(); ++i) reordered[i] = uint32_t(i);
    st.indices = reordered;
    Node root;
    root.lo = st.full.lo;
    root.hi = st.full.hi;
    root.tp = 0x80000000u;
    root.first = 0;
    root.second = uint32_t(input.size()) - 1;
    nodes.push_back(root);
    uint32_t current = 0;
    long active = 1;
    while (active--) {
      Node l, r;
      if (split(current, st, l, r)) {
        active += 2;
        nodes[current].first = uint32_t(nodes.size());
        nodes[current].second = uint32_t(nodes.size() + 1);
        nodes[current].make_leaf(false);
        l.set_parent(current);
        r.set_parent(current);
        nodes.push_back(l);
        nodes.push_back(r);
      }
      ++current;
    }
    reordered = st.indices;
  }
};

}  // namespace

struct orc_bvh_s {
  std::vector<Node> nodes;
  std::vector<uint32_t> indices;
  std::vector<AB> boxes;
};

extern "C" {

void orc_triangle_bounds(const uint32_t* idx, size_t n_idx, const float* pos3, orc_aabb* out) {
  for (size_t i = 0; i + 2 < n_idx; i += 3) {
    AB b;
    V3 p[3];
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) p[k].v[c] = pos3[3 * size_t(idx[i + k]) + c];
    b.enclose_point(p[0]);
    b.enclose_point(p[1]);
    b.enclose_point(p[2]);
    for (int c = 0; c < 3; ++c) {
      b.lo.v[c] -= 1e-5f;
      b.hi.v[c] += 1e-5f;
    }
    orc_aabb& o = out[i / 3];
    for (int c = 0; c < 3; ++c) {
      o.min[c] = b.lo.v[c];
      o.max[c] = b.hi.v[c];
      o.centroid[c] = (p[0].v[c] + p[1].v[c] + p[2].v[c]) / 3.0f;
    }
    o.has_centroid = 1;
  }
}

int orc_bvh_build(const orc_aabb* boxes, size_t n, int split_aabbs, orc_bvh* out) {
  if (!boxes || n == 0 || !out) return -1;
  std::vector<AB> in(n);
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      in[i].lo.v[c] = boxes[i].min[c];
      in[i].hi.v[c] = boxes[i].max[c];
    }
    if (boxes[i].has_centroid) in[i].wc = V3{{boxes[i].centroid[0], boxes[i].centroid[1], boxes[i].centroid[2]}};
  }
  Builder b;
  b.split_aabbs = split_aabbs != 0;
  State st;
  b.build(in, st);
  auto* h = new orc_bvh_s;
  h->nodes = std::move(b.nodes);
  h->indices = std::move(b.reordered);
  h->boxes = std::move(st.boxes);
  *out = h;
  return 0;
}

void orc_bvh_sizes(orc_bvh b, size_t* n_nodes, size_t* n_indices) {
  if (n_nodes) *n_nodes = b->nodes.size();
  if (n_indices) *n_indices = b->indices.size();
}

void orc_bvh_copy(orc_bvh b, orc_node* nodes, uint32_t* indices, orc_aabb* final_boxes) {
  if (nodes) {
    for (size_t i = 0; i < b->nodes.size(); ++i) {
      orc_node& o = nodes[i];
      std::memset(&o, 0, sizeof(o));
      for (int c = 0; c < 3; ++c) {
        o.min[c] = b->nodes[i].lo.v[c];
        o.max[c] = b->nodes[i].hi.v[c];
      }
      o.type_and_parent = b->nodes[i].tp;
      o.first = b->nodes[i].first;
      o.second = b->nodes[i].second;
    }
  }
  if (indices) std::memcpy(indices, b->indices.data(), b->indices.size() * 4);
  if (final_boxes) {
    for (size_t i = 0; i < b->boxes.size(); ++i) {
      orc_aabb& o = final_boxes[i];
      for (int c = 0; c < 3; ++c) {
        o.min[c] = b->boxes[i].lo.v[c];
        o.max[c] = b->boxes[i].hi.v[c];
        o.centroid[c] = b->boxes[i].wc ? b->boxes[i].wc->v[c] : 0.f;
      }
      o.has_centroid = b->boxes[i].wc ? 1u : 0u;
    }
  }
}

void orc_bvh_destroy(orc_bvh b) { delete b; }

}  // extern "C"
