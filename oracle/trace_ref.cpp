// Oracle (test infrastructure only): ray/triangle tests, brute-force and BVH queries, and the
// reference's own stackless traversal used as the CPU baseline.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "oracle.h"
#include "ray_math.hpp"

namespace {
using namespace orc_math;

inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
inline float fmax_std(float a, float b) { return (a < b) ? b : a; }


const float* vtx(const orc_scene* s, uint32_t g, uint32_t tri, int k) {
  const uint32_t vi = s->idx[s->geo_idx_off[g] + 3 * tri + k];
  return s->pos + 3 * (size_t(s->geo_vtx_off[g]) + vi);
}

struct Best {
  float t;
  uint32_t inst, prim;
  float b1, b2;
  bool found;
};

inline void consider(Best& b, float t, uint32_t inst, uint32_t prim, float b1, float b2) {
  if (t < b.t || (t == b.t && (inst < b.inst || (inst == b.inst && prim < b.prim)))) {
    b.t = t;
    b.inst = inst;
    b.prim = prim;
    b.b1 = b1;
    b.b2 = b2;
    b.found = true;
  }
}

inline void write_hit(const Best& b, bool any, orc_hit& h) {
  if (any) {
    h.t = b.found ? 1.f : INFINITY;
    h.prim = h.inst = 0xFFFFFFFFu;
    h.b1 = h.b2 = 0.f;
    return;
  }
  h.t = b.found ? b.t : INFINITY;
  h.prim = b.found ? b.prim : 0xFFFFFFFFu;
  h.inst = b.found ? b.inst : 0xFFFFFFFFu;
  h.b1 = b.found ? b.b1 : 0.f;
  h.b2 = b.found ? b.b2 : 0.f;
}

template <typename F>
void parallel_for(size_t n, int threads, F f) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  if (threads == 1 || n < 64) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([=, &f] {
      for (size_t i = b; i < e; ++i) f(i);
    });
  }
  for (auto& th : pool) th.join();
}

inline bool dir_zero(const float d[3]) { return !(d[0] != 0.f || d[1] != 0.f || d[2] != 0.f); }

}  // namespace

// ---- accelerated oracle: binary BVHs from the restated reference builder ---------------------
struct orc_accel_s {
  const orc_scene* s = nullptr;
  std::vector<std::vector<orc_node>> blas;
  std::vector<std::vector<uint32_t>> blas_idx;
  std::vector<orc_node> tlas;
  std::vector<uint32_t> tlas_idx;
  // Reference-baseline layout: all BLAS nodes concatenated, per geometry node/index bases.
  std::vector<orc_node> all_nodes;
  std::vector<uint32_t> all_idx;
  std::vector<uint32_t> node_base, idx_base;
  std::vector<orc_node> tlas_ref;  // instance BVH over unpadded boxes (scene.cpp:338)
  std::vector<uint32_t> tlas_ref_idx;
};

namespace {
using namespace orc_math;

// Robust slab test (far distance scaled by 1 + 2*gamma(3)) against a box padded by a relative
// epsilon, so the oracle BVH never culls a box containing a watertight hit.
inline bool slab(const orc_node& n, const float o[3], const float inv[3], float tmin, float tcut, float& tnear) {
  float tn = tmin, tf = tcut;
  for (int a = 0; a < 3; ++a) {
    const float pad = 1e-6f * (std::fabs(n.min[a]) + std::fabs(n.max[a])) + 1e-30f;
    const float lo = n.min[a] - pad, hi = n.max[a] + pad;
    float t0 = (lo - o[a]) * inv[a], t1 = (hi - o[a]) * inv[a];
    if (std::isnan(t0)) t0 = -INFINITY;  // origin on a face of a slab parallel to the ray
    if (std::isnan(t1)) t1 = INFINITY;
    if (t0 > t1) std::swap(t0, t1);
    tn = std::fmax(tn, t0);
    tf = std::fmin(tf, t1 * 1.0000005f);
  }
  tnear = tn;
  return tn <= tf * 1.0000005f;
}

void accel_query(const orc_accel_s* A, const orc_ray& r, bool any, Best& best) {
  const orc_scene* s = A->s;
  best = Best{r.tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.f, 0.f, false};
  if (dir_zero(r.d) || A->tlas.empty()) return;
  const float winv[3] = {1.0f / r.d[0], 1.0f / r.d[1], 1.0f / r.d[2]};
  std::vector<uint32_t> st;
  st.reserve(128);
  st.push_back(0);
  while (!st.empty()) {
    const uint32_t ni = st.back();
    st.pop_back();
    const orc_node& n = A->tlas[ni];
    float tn;
    if (!slab(n, r.o, winv, r.tmin, best.t, tn)) continue;
    if (n.type_and_parent & 0x80000000u) {
      for (uint32_t k = n.first; k <= n.second; ++k) {
        const uint32_t inst = A->tlas_idx[k];
        const uint32_t g = s->inst_geom[inst];
        float oo[3], od[3];
        to_object(s->inst_inv + 12 * inst, r.o, r.d, oo, od);
        Pre pre;
        make_pre(oo, od, pre);
        const float oinv[3] = {1.0f / od[0], 1.0f / od[1], 1.0f / od[2]};
        const std::vector<orc_node>& bn = A->blas[g];
        const std::vector<uint32_t>& bi = A->blas_idx[g];
        std::vector<uint32_t> st2{0u};
        while (!st2.empty()) {
          const orc_node& m = bn[st2.back()];
          st2.pop_back();
          float tn2;
          if (!slab(m, oo, oinv, r.tmin, best.t, tn2)) continue;
          if (m.type_and_parent & 0x80000000u) {
            for (uint32_t q = m.first; q <= m.second; ++q) {
              const uint32_t prim = bi[q];
              float t, b1, b2;
              if (!watertight(pre, vtx(s, g, prim, 0), vtx(s, g, prim, 1), vtx(s, g, prim, 2), t, b1, b2)) continue;
              if (!(t > r.tmin && t < r.tmax)) continue;
              if (any) {
                best.found = true;
                return;
              }
              consider(best, t, inst, prim, b1, b2);
            }
          } else {
            st2.push_back(m.second);
            st2.push_back(m.first);
          }
        }
      }
    } else {
      st.push_back(n.second);
      st.push_back(n.first);
    }
  }
}

// ---- the reference traversal (CPU baseline) --------------------------------------------------
// bvh_ray_aabb_intersect (src/pt/bvh.glsl:31-45)
inline bool ref_aabb(const orc_node& n, const float o[3], const float id[3], float mt, float& t) {
  float t135[3], t246[3];
  for (int a = 0; a < 3; ++a) {
    t135[a] = (n.min[a] - o[a]) * id[a];
    t246[a] = (n.max[a] - o[a]) * id[a];
  }
  const float minv[3] = {fmin_std(t135[0], t246[0]), fmin_std(t135[1], t246[1]), fmin_std(t135[2], t246[2])};
  const float maxv[3] = {fmax_std(t135[0], t246[0]), fmax_std(t135[1], t246[1]), fmax_std(t135[2], t246[2])};
  const float tmin = fmax_std(fmax_std(minv[0], minv[1]), minv[2]);
  const float tmax = fmin_std(fmin_std(maxv[0], maxv[1]), maxv[2]);
  if (tmax >= 0 && tmin <= tmax && tmin <= mt) {
    t = tmin;
    return true;
  }
  t = mt;
  return false;
}

// intersect_triangle (src/pt/intersect.glsl:3-28)
inline bool ref_mt(const float o[3], const float d[3], const float* v1, const float* v2, const float* v3, float& t,
                   float& bx, float& by) {
  const float eps = 1e-23f, border = 1e-6f;
  const float e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  const float e2[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
  const float P[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
  const float det = e1[0] * P[0] + e1[1] * P[1] + e1[2] * P[2];
  if (det > -eps && det < eps) return false;
  const float inv_det = 1.0f / det;
  const float T[3] = {o[0] - v1[0], o[1] - v1[1], o[2] - v1[2]};
  bx = (T[0] * P[0] + T[1] * P[1] + T[2] * P[2]) * inv_det;
  if (bx < -border || bx > 1.0f + border) return false;
  const float Q[3] = {T[1] * e1[2] - T[2] * e1[1], T[2] * e1[0] - T[0] * e1[2], T[0] * e1[1] - T[1] * e1[0]};
  by = (d[0] * Q[0] + d[1] * Q[1] + d[2] * Q[2]) * inv_det;
  if (by < -border || bx + by > 1.0f + border) return false;
  const float tt = (e2[0] * Q[0] + e2[1] * Q[1] + e2[2] * Q[2]) * inv_det;
  if (tt > eps) {
    t = tt;
    return true;
  }
  return false;
}

// traversal_state_t (src/pt/pt_trace.comp:229-257)
struct RefState {
  bool any_hit;
  float t, global_t;
  uint32_t hit_geometry, hit_index, global_hit_index;
  uint32_t geom;
};

// MYRT_BVH_TRAVERSE (src/pt/bvh.impl.glsl:8-72) with a visitor.
template <typename Visit>
bool ref_traverse(const orc_node* nodes, uint32_t node_base, const uint32_t* indices, uint32_t prim_base,
                  const float ro[3], const float rd[3], float maxt, Visit&& visit) {
  bool hits_primitive = false;
  const float inv_rd[3] = {1.0f / rd[0], 1.0f / rd[1], 1.0f / rd[2]};
  uint32_t visited = 0, lr = 0;
  uint32_t node_index = node_base;
  float tmp = INFINITY;
  const bool hits_any = ref_aabb(nodes[node_base], ro, inv_rd, maxt, tmp);
  while (hits_any) {
    orc_node current = nodes[node_index];
    while (!(current.type_and_parent & 0x80000000u)) {
      float tl, tr;
      const orc_node& fc = nodes[current.first + node_base];
      const orc_node& sc = nodes[current.second + node_base];
      const bool hl = ref_aabb(fc, ro, inv_rd, maxt, tl);
      const bool hr = ref_aabb(sc, ro, inv_rd, maxt, tr);
      if (!hl && !hr) break;
      const bool use_right = !hl || (hr && tl >= tr);
      lr = uint32_t(use_right) | (lr << 1u);
      visited = uint32_t(hl && hr) | (visited << 1u);
      node_index = (use_right ? current.second : current.first) + node_base;
      current = nodes[node_index];
    }
    if (current.type_and_parent & 0x80000000u) {
      for (uint32_t tri = current.first; tri <= current.second; ++tri) {
        bool h = false;
        if (visit(ro, rd, indices[prim_base + tri], maxt, h)) return true;
        if (h) hits_primitive = true;
      }
    }
    while ((visited & 1u) != 1u) {
      if (visited == 0) return hits_primitive;
      node_index = (nodes[node_index].type_and_parent & 0x7fffffffu) + node_base;
      visited >>= 1u;
      lr >>= 1u;
    }
    const orc_node& parent = nodes[(nodes[node_index].type_and_parent & 0x7fffffffu) + node_base];
    node_index = node_base + (((lr & 1u) == 1u) ? parent.first : parent.second);
    visited ^= 1u;
  }
  return false;
}

// nearest_hit / any_hit (src/pt/pt_trace.comp:311-340) without the SDF part.
void ref_query(const orc_accel_s* A, const orc_ray& r, bool any, orc_hit& out) {
  const orc_scene* s = A->s;
  RefState st{any, r.tmax, r.tmax, 0, 0, 0, 0};
  bool hits = false;
  if (!A->tlas_ref.empty()) {
    auto visit_triangle = [&](const float* ro, const float* rd, uint32_t index, float& max_d, bool& h) {
      const uint32_t g = st.geom;
      float t = 0, bx, by;
      h = ref_mt(ro, rd, vtx(s, g, index, 0), vtx(s, g, index, 1), vtx(s, g, index, 2), t, bx, by) && t < max_d &&
          st.t > t;
      if (h) {
        st.t = t;
        st.hit_index = index;
        max_d = st.t;
      }
      return st.any_hit && h;
    };
    auto visit_object = [&](const float* ro, const float* rd, uint32_t i, float& max_d, bool& h) {
      const uint32_t g = s->inst_geom[i];
      st.geom = g;
      float oo[3], od[3];
      to_object(s->inst_inv + 12 * i, ro, rd, oo, od);
      st.t = max_d;
      bool cur = ref_traverse(A->all_nodes.data(), A->node_base[g], A->all_idx.data(), A->idx_base[g], oo, od, max_d,
                              visit_triangle);
      cur = cur && st.global_t > st.t;
      if (cur) {
        st.global_t = st.t;
        st.global_hit_index = st.hit_index;
        st.hit_geometry = i;
        max_d = st.global_t;
      }
      h = cur;
      return st.any_hit && h;
    };
    hits = ref_traverse(A->tlas_ref.data(), 0, A->tlas_ref_idx.data(), 0, r.o, r.d, r.tmax, visit_object);
  }
  if (any) {
    out.t = hits ? 1.f : INFINITY;
    out.prim = out.inst = 0xFFFFFFFFu;
  } else {
    out.t = hits ? st.global_t : INFINITY;
    out.prim = hits ? st.global_hit_index : 0xFFFFFFFFu;
    out.inst = hits ? st.hit_geometry : 0xFFFFFFFFu;
  }
  out.b1 = out.b2 = 0.f;
}

void copy_bvh(orc_bvh b, std::vector<orc_node>& nodes, std::vector<uint32_t>& idx) {
  size_t nn, ni;
  orc_bvh_sizes(b, &nn, &ni);
  nodes.resize(nn);
  idx.resize(ni);
  orc_bvh_copy(b, nodes.data(), idx.data(), nullptr);
}

}  // namespace

extern "C" {

int orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                   float* t, float* b1, float* b2) {
  Pre p;
  make_pre(o, d, p);
  float tt = 0, u = 0, v = 0;
  const bool hit = watertight(p, v0, v1, v2, tt, u, v);
  if (t) *t = tt;
  if (b1) *b1 = u;
  if (b2) *b2 = v;
  return hit ? 1 : 0;
}

int orc_moller_trumbore(const float o[3], const float d[3], const float v1[3], const float v2[3], const float v3[3],
                        float* t, float* bx, float* by) {
  float tt = 0, x = 0, y = 0;
  const bool hit = ref_mt(o, d, v1, v2, v3, tt, x, y);
  if (t) *t = tt;
  if (bx) *bx = x;
  if (by) *by = y;
  return hit ? 1 : 0;
}

// Same cofactor expansion, in the same order, as the product's instance inverse.
int orc_invert(const float m[16], float out[12]) {
  double a[4][4];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) a[r][c] = double(m[c * 4 + r]);
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
  if (!(det != 0.0) || !std::isfinite(det)) return 0;
  const double id = 1.0 / det;
  const double r[3][4] = {
      {(a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id, (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id,
       (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id, (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id},
      {(-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id, (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id,
       (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id, (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id},
      {(a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id, (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id,
       (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id, (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = float(r[i][j]);
  return 1;
}

void orc_transform_box(const float bmin[3], const float bmax[3], const float m[16], float omin[3], float omax[3]) {
  for (int k = 0; k < 3; ++k) {
    omin[k] = std::numeric_limits<float>::max();
    omax[k] = std::numeric_limits<float>::lowest();
  }
  for (unsigned i = 0; i < 8; ++i) {
    const float p[3] = {(i & 1u) ? bmax[0] : bmin[0], (i & 2u) ? bmax[1] : bmin[1], (i & 4u) ? bmax[2] : bmin[2]};
    for (int r = 0; r < 3; ++r) {
      const float v = m[0 * 4 + r] * p[0] + m[1 * 4 + r] * p[1] + m[2 * 4 + r] * p[2] + m[3 * 4 + r];
      omin[r] = fmin_std(omin[r], v);
      omax[r] = fmax_std(omax[r], v);
    }
  }
}

void orc_trace_brute(const orc_scene* s, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits, int threads) {
  parallel_for(n, threads, [&](size_t i) {
    const orc_ray& r = rays[i];
    Best best{r.tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.f, 0.f, false};
    if (!dir_zero(r.d)) {
      for (uint32_t inst = 0; inst < s->n_inst && !(any_hit && best.found); ++inst) {
        const uint32_t g = s->inst_geom[inst];
        float oo[3], od[3];
        to_object(s->inst_inv + 12 * inst, r.o, r.d, oo, od);
        Pre pre;
        make_pre(oo, od, pre);
        for (uint32_t tri = 0; tri < s->geo_tri_cnt[g]; ++tri) {
          float t, b1, b2;
          if (!watertight(pre, vtx(s, g, tri, 0), vtx(s, g, tri, 1), vtx(s, g, tri, 2), t, b1, b2)) continue;
          if (!(t > r.tmin && t < r.tmax)) continue;
          if (any_hit) {
            best.found = true;
            break;
          }
          consider(best, t, inst, tri, b1, b2);
        }
      }
    }
    write_hit(best, any_hit != 0, hits[i]);
  });
}

int orc_accel_build(const orc_scene* s, orc_accel* out) {
  auto* A = new orc_accel_s;
  A->s = s;
  A->blas.resize(s->n_geoms);
  A->blas_idx.resize(s->n_geoms);
  std::vector<orc_aabb> geo_box(s->n_geoms);
  for (uint32_t g = 0; g < s->n_geoms; ++g) {
    const uint32_t ntri = s->geo_tri_cnt[g];
    std::vector<orc_aabb> boxes(ntri);
    std::vector<uint32_t> local(s->idx + s->geo_idx_off[g], s->idx + s->geo_idx_off[g] + 3 * ntri);
    orc_triangle_bounds(local.data(), local.size(), s->pos + 3 * size_t(s->geo_vtx_off[g]), boxes.data());
    orc_bvh b;
    if (orc_bvh_build(boxes.data(), boxes.size(), 1, &b) != 0) {
      delete A;
      return -1;
    }
    copy_bvh(b, A->blas[g], A->blas_idx[g]);
    orc_bvh_destroy(b);
    A->node_base.push_back(uint32_t(A->all_nodes.size()));
    A->idx_base.push_back(uint32_t(A->all_idx.size()));
    A->all_nodes.insert(A->all_nodes.end(), A->blas[g].begin(), A->blas[g].end());
    A->all_idx.insert(A->all_idx.end(), A->blas_idx[g].begin(), A->blas_idx[g].end());
    for (int c = 0; c < 3; ++c) {
      geo_box[g].min[c] = A->blas[g][0].min[c];
      geo_box[g].max[c] = A->blas[g][0].max[c];
    }
  }
  if (s->n_inst > 0) {
    // Instance boxes: the object box's corners through M = inverse of the stored inverse rows
    // (in double), widened by a relative margin for the oracle BVH; the reference-baseline TLAS
    // uses the same boxes without the margin.
    std::vector<orc_aabb> padded(s->n_inst), plain(s->n_inst);
    for (uint32_t i = 0; i < s->n_inst; ++i) {
      const float* iv = s->inst_inv + 12 * i;
      float m4[16] = {iv[0], iv[4], iv[8], 0, iv[1], iv[5], iv[9], 0, iv[2], iv[6], iv[10], 0, iv[3], iv[7], iv[11], 1};
      float fwd[12];
      orc_invert(m4, fwd);
      float mcol[16] = {fwd[0], fwd[4], fwd[8], 0, fwd[1], fwd[5], fwd[9], 0,
                        fwd[2], fwd[6], fwd[10], 0, fwd[3], fwd[7], fwd[11], 1};
      const orc_aabb& gb = geo_box[s->inst_geom[i]];
      float lo[3], hi[3];
      orc_transform_box(gb.min, gb.max, mcol, lo, hi);
      double mag = 0, ext = 0;
      for (int c = 0; c < 3; ++c) {
        mag = std::max(mag, std::max(std::fabs(double(lo[c])), std::fabs(double(hi[c]))));
        ext = std::max(ext, double(hi[c]) - double(lo[c]));
      }
      const float pad = float(1e-4 * (mag + ext));
      for (int c = 0; c < 3; ++c) {
        plain[i].min[c] = lo[c];
        plain[i].max[c] = hi[c];
        padded[i].min[c] = lo[c] - pad;
        padded[i].max[c] = hi[c] + pad;
      }
      plain[i].has_centroid = padded[i].has_centroid = 0;
    }
    orc_bvh b;
    if (orc_bvh_build(padded.data(), padded.size(), 0, &b) != 0) {
      delete A;
      return -1;
    }
    copy_bvh(b, A->tlas, A->tlas_idx);
    orc_bvh_destroy(b);
    if (orc_bvh_build(plain.data(), plain.size(), 0, &b) != 0) {
      delete A;
      return -1;
    }
    copy_bvh(b, A->tlas_ref, A->tlas_ref_idx);
    orc_bvh_destroy(b);
  }
  *out = A;
  return 0;
}

void orc_accel_trace(orc_accel a, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits, int threads) {
  parallel_for(n, threads, [&](size_t i) {
    Best best;
    accel_query(a, rays[i], any_hit != 0, best);
    write_hit(best, any_hit != 0, hits[i]);
  });
}

void orc_accel_destroy(orc_accel a) { delete a; }

void orc_reference_trace(orc_accel a, const orc_ray* rays, size_t n, int any_hit, orc_hit* hits) {
  for (size_t i = 0; i < n; ++i) ref_query(a, rays[i], any_hit != 0, hits[i]);
}

}  // extern "C"
