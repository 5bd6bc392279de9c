// Collapse of a binary BVH into 4-wide nodes of one 64-byte line each, with 8-bit quantised
// child boxes that always decode to a box strictly containing the child.
#include "wide_bvh.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mrt {
namespace {

struct Child {
  bool from_bin = false;          // a binary inner node; otherwise a list of reference entries
  uint32_t bin = 0;
  std::vector<uint32_t> items;    // positions in BinaryBvh::indices (leaf lists)
  Box box;
  bool internal = false;          // needs a wide node of its own
};

Child list_child(const BinaryBvh& b, std::vector<uint32_t> items, uint32_t max_leaf) {
  Child c;
  c.items = std::move(items);
  for (uint32_t i : c.items) c.box.enclose(b.ref_boxes[i]);
  c.internal = c.items.size() > max_leaf;
  return c;
}

Child bin_child(const BinaryBvh& b, uint32_t idx, uint32_t max_leaf) {
  const BinNode& n = b.nodes[idx];
  if (n.is_leaf()) {
    std::vector<uint32_t> items;
    for (uint32_t i = n.first; i <= n.second; ++i) items.push_back(i);
    return list_child(b, std::move(items), max_leaf);
  }
  Child c;
  c.from_bin = true;
  c.bin = idx;
  c.box.mn = {n.mn[0], n.mn[1], n.mn[2]};
  c.box.mx = {n.mx[0], n.mx[1], n.mx[2]};
  c.internal = true;
  return c;
}

// Splits a leaf list at the median centroid of its widest centroid axis (index order when all
// centroids coincide).
void split_list(const BinaryBvh& b, const Child& c, Child out[2], uint32_t max_leaf) {
  std::vector<uint32_t> items = c.items;
  Box cb;
  for (uint32_t i : items) cb.enclose(b.ref_boxes[i].center());
  const vec3f d = cb.dim();
  const int axis = (d.x >= d.y && d.x >= d.z) ? 0 : (d.y >= d.z ? 1 : 2);
  const size_t half = items.size() / 2;
  if (d[axis] > 0.f) {
    std::nth_element(items.begin(), items.begin() + half, items.end(), [&](uint32_t a, uint32_t bb) {
      const float ca = b.ref_boxes[a].center()[axis], cbv = b.ref_boxes[bb].center()[axis];
      return ca < cbv || (ca == cbv && a < bb);
    });
  }
  out[0] = list_child(b, std::vector<uint32_t>(items.begin(), items.begin() + half), max_leaf);
  out[1] = list_child(b, std::vector<uint32_t>(items.begin() + half, items.end()), max_leaf);
}

void expand(const BinaryBvh& b, const Child& c, Child out[2], uint32_t max_leaf) {
  if (c.from_bin) {
    out[0] = bin_child(b, b.nodes[c.bin].first, max_leaf);
    out[1] = bin_child(b, b.nodes[c.bin].second, max_leaf);
  } else {
    split_list(b, c, out, max_leaf);
  }
}

// Opens the largest-area inner child until the node has four children or only leaves remain.
std::vector<Child> gather_children(const BinaryBvh& b, const Child& parent, uint32_t max_leaf) {
  std::vector<Child> ch(2);
  expand(b, parent, ch.data(), max_leaf);
  while (ch.size() < 4) {
    int best = -1;
    float best_area = 0.f;
    for (int i = 0; i < int(ch.size()); ++i) {
      if (!ch[i].internal) continue;
      const float a = ch[i].box.area();
      if (best < 0 || a > best_area) {
        best = i;
        best_area = a;
      }
    }
    if (best < 0) break;
    Child e[2];
    expand(b, ch[best], e, max_leaf);
    ch[best] = e[0];
    ch.insert(ch.begin() + best + 1, e[1]);
  }
  return ch;
}

inline double clamp_coord(double v) {
  if (std::isnan(v)) return 0.0;
  return std::fmin(std::fmax(v, -1e37), 1e37);
}

// Largest float <= v and smallest float >= v.
inline float float_down(double v) {
  float f = float(v);
  if (double(f) > v) f = std::nextafterf(f, -INFINITY);
  return f;
}
inline float float_up(double v) {
  float f = float(v);
  if (double(f) < v) f = std::nextafterf(f, INFINITY);
  return f;
}

void quantize(WideNode& n, const std::vector<Child>& ch, double pad_rel) {
  const int nc = int(ch.size());
  for (int a = 0; a < 3; ++a) {
    float lreq[4], hreq[4];
    for (int c = 0; c < nc; ++c) {
      const Box& bx = ch[c].box;
      double m = 0.0;
      for (int k = 0; k < 3; ++k) {
        m = std::fmax(m, std::fabs(clamp_coord(bx.mn[k])));
        m = std::fmax(m, std::fabs(clamp_coord(bx.mx[k])));
      }
      const double pad = pad_rel * m;
      const float cmn = bx.mn[a], cmx = bx.mx[a];
      if (!(cmn <= cmx)) {  // empty or NaN box: nothing in it can be hit; keep a point box
        lreq[c] = hreq[c] = 0.f;
        continue;
      }
      const double lo = clamp_coord(double(cmn) - pad), hi = clamp_coord(double(cmx) + pad);
      lreq[c] = std::fmin(float_down(lo), std::nextafterf(float(clamp_coord(cmn)), -INFINITY));
      hreq[c] = std::fmax(float_up(hi), std::nextafterf(float(clamp_coord(cmx)), INFINITY));
    }
    float origin = lreq[0];
    double top = hreq[0];
    for (int c = 1; c < nc; ++c) {
      origin = std::fmin(origin, lreq[c]);
      top = std::fmax(top, double(hreq[c]));
    }
    const double span = top - double(origin);
    int e = span > 0 ? int(std::ceil(std::log2(span / 255.0))) - 1 : -126;
    if (e < -126) e = -126;
    for (;; ++e) {
      if (e > 127) throw std::runtime_error("wide bvh: box extent too large to quantise");
      const float scale = std::ldexp(1.0f, e);
      bool ok = true;
      for (int c = 0; c < nc && ok; ++c) {
        double ql = std::floor((double(lreq[c]) - double(origin)) / double(scale));
        int qlo = int(std::fmin(std::fmax(ql, 0.0), 255.0));
        while (qlo > 0 && std::fmaf(float(qlo), scale, origin) > lreq[c]) --qlo;
        double qh = std::ceil((double(hreq[c]) - double(origin)) / double(scale));
        int qhi = int(std::fmin(std::fmax(qh, 0.0), 255.0));
        while (std::fmaf(float(qhi), scale, origin) < hreq[c]) {
          if (qhi == 255) {
            ok = false;
            break;
          }
          ++qhi;
        }
        n.qlo[a][c] = uint8_t(qlo);
        n.qhi[a][c] = uint8_t(qhi);
      }
      if (ok) break;
    }
    n.origin[a] = origin;
    n.exp[a] = int8_t(e);
    for (int c = nc; c < 4; ++c) {  // unused slot: an inverted box (low code above high)
      n.qlo[a][c] = 255;
      n.qhi[a][c] = 0;
    }
  }
}

}  // namespace

void decode_child_box(const WideNode& n, int c, float lo[3], float hi[3]) {
  for (int a = 0; a < 3; ++a) {
    const float scale = std::ldexp(1.0f, n.exp[a]);
    lo[a] = std::fmaf(float(n.qlo[a][c]), scale, n.origin[a]);
    hi[a] = std::fmaf(float(n.qhi[a][c]), scale, n.origin[a]);
  }
}

WideBvh collapse_to_wide(const BinaryBvh& b, double pad_rel, uint32_t max_leaf) {
  if (max_leaf < 1) max_leaf = 1;
  if (max_leaf > kMaxLeafItems) max_leaf = kMaxLeafItems;
  WideBvh w;
  if (b.nodes.empty()) return w;
  struct Pending {
    uint32_t node;
    Child c;
    uint32_t depth;
  };
  w.nodes.emplace_back();
  std::vector<Pending> stack{{0u, bin_child(b, 0, max_leaf), 1u}};
  while (!stack.empty()) {
    Pending p = stack.back();
    stack.pop_back();
    w.depth = std::max(w.depth, p.depth);
    const std::vector<Child> ch = p.c.internal ? gather_children(b, p.c, max_leaf) : std::vector<Child>{p.c};
    uint32_t n_inner = 0;
    for (const Child& c : ch) n_inner += c.internal ? 1 : 0;
    const uint32_t node_base = uint32_t(w.nodes.size());
    w.nodes.resize(node_base + n_inner);
    WideNode& n = w.nodes[p.node];
    n = WideNode{};
    n.nchildren = uint8_t(ch.size());
    for (int k = 0; k < 4; ++k) n.entry[k] = 0x80000000u;  // unused slots: a leaf of no items
    uint32_t inner = 0;
    for (size_t k = 0; k < ch.size(); ++k) {
      if (ch[k].internal) {
        n.entry[k] = node_base + inner++;
      } else {
        n.entry[k] = 0x80000000u | (uint32_t(ch[k].items.size()) << 24) | uint32_t(w.items.size());
        w.items.insert(w.items.end(), ch[k].items.begin(), ch[k].items.end());
      }
    }
    quantize(n, ch, pad_rel);
    inner = n_inner;
    for (size_t k = ch.size(); k-- > 0;) {
      if (ch[k].internal) stack.push_back({node_base + --inner, ch[k], p.depth + 1});
    }
  }
  return w;
}

}  // namespace mrt
