// Collapse of a binary BVH into 4-wide nodes of one 64-byte line each, with 8-bit quantised
// child boxes that always decode to a box strictly containing the child.
#include "wide_bvh.hpp"

#include <cmath>
#include <stdexcept>

namespace mrt {
namespace {

struct Child {
  bool from_bin = false;  // a binary node; otherwise a range of reference entries
  uint32_t bin = 0;
  uint32_t start = 0, count = 0;
  Box box;
  bool internal = false;  // needs a wide node of its own
};

Child bin_child(const BinaryBvh& b, uint32_t idx) {
  const BinNode& n = b.nodes[idx];
  Child c;
  c.from_bin = true;
  c.bin = idx;
  c.box.mn = {n.mn[0], n.mn[1], n.mn[2]};
  c.box.mx = {n.mx[0], n.mx[1], n.mx[2]};
  if (n.is_leaf()) {
    c.start = n.first;
    c.count = n.second - n.first + 1;
    c.internal = c.count > kMaxLeafItems;
  } else {
    c.internal = true;
  }
  return c;
}

Child range_child(const BinaryBvh& b, uint32_t start, uint32_t count) {
  Child c;
  c.start = start;
  c.count = count;
  for (uint32_t i = start; i < start + count; ++i) c.box.enclose(b.ref_boxes[i]);
  c.internal = count > kMaxLeafItems;
  return c;
}

void expand(const BinaryBvh& b, const Child& c, Child out[2]) {
  if (c.from_bin && !b.nodes[c.bin].is_leaf()) {
    out[0] = bin_child(b, b.nodes[c.bin].first);
    out[1] = bin_child(b, b.nodes[c.bin].second);
  } else {  // an over-full leaf: halve its range
    const uint32_t half = c.count / 2;
    out[0] = range_child(b, c.start, half);
    out[1] = range_child(b, c.start + half, c.count - half);
  }
}

// Opens the largest-area inner child until the node has four children or only leaves remain.
std::vector<Child> gather_children(const BinaryBvh& b, const Child& parent) {
  std::vector<Child> ch(2);
  expand(b, parent, ch.data());
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
    expand(b, ch[best], e);
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
    for (int c = nc; c < 4; ++c) {
      n.qlo[a][c] = 0;
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

WideBvh collapse_to_wide(const BinaryBvh& b, double pad_rel) {
  WideBvh w;
  if (b.nodes.empty()) return w;
  struct Pending {
    uint32_t node;
    Child c;
    uint32_t depth;
  };
  w.nodes.emplace_back();
  std::vector<Pending> stack{{0u, bin_child(b, 0), 1u}};
  while (!stack.empty()) {
    Pending p = stack.back();
    stack.pop_back();
    w.depth = std::max(w.depth, p.depth);
    const std::vector<Child> ch = p.c.internal ? gather_children(b, p.c) : std::vector<Child>{p.c};
    uint32_t n_inner = 0;
    for (const Child& c : ch) n_inner += c.internal ? 1 : 0;
    const uint32_t node_base = uint32_t(w.nodes.size());
    w.nodes.resize(node_base + n_inner);
    WideNode& n = w.nodes[p.node];
    n = WideNode{};
    n.nchildren = uint8_t(ch.size());
    n.node_base = node_base;
    n.item_base = uint32_t(w.items.size());
    uint32_t inner = 0;
    for (size_t k = 0; k < ch.size(); ++k) {
      if (ch[k].internal) {
        n.count[k] = 0;
        n.offset[k] = uint16_t(inner++);
      } else {
        n.count[k] = uint8_t(ch[k].count);
        n.offset[k] = uint16_t(w.items.size() - n.item_base);
        for (uint32_t i = 0; i < ch[k].count; ++i) w.items.push_back(ch[k].start + i);
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
