// Instance-BVH refit for frames that only move instances.
//
// The reference rebuilds its instance BVH on the host whenever the frame's drawable hash changes
// (src/pathtracer/scene.cpp:255-271, 336-344), a transform included. Here a frame that keeps
// every drawable's geometry and material and changes only matrices keeps the wide TLAS topology
// and refits it on the GPU:
//   1. instance_kernel: per instance, the world -> object rows (the double-precision cofactor
//      inverse the host computes at enqueue) and the world box (the geometry box's 8 corners
//      through M, float, the host's operation order), written to the instance records and a box
//      array;
//   2. level_kernel, deepest level first: per wide node, the union of its children's boxes
//      (inner children from the level below, leaf children from their instances) and the
//      requantisation of the child boxes, the host collapse's rule step for step, so a refit with
//      unchanged matrices reproduces the host-built nodes bit for bit.
// Everything is fixed-order float/double arithmetic with contraction off (-ffp-contract=off), so
// host and device agree to the bit.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "refit.hpp"

namespace mrt {
namespace {

struct DBox {
  float mn[3], mx[3];
};

__device__ inline float ref_min(float a, float b) { return (b < a) ? b : a; }
__device__ inline float ref_max(float a, float b) { return (a < b) ? b : a; }

__device__ inline void empty_box(DBox& b) {
  for (int a = 0; a < 3; ++a) {
    b.mn[a] = FLT_MAX;
    b.mx[a] = -FLT_MAX;
  }
}
__device__ inline void enclose(DBox& b, const float* mn, const float* mx) {
  for (int a = 0; a < 3; ++a) {
    b.mn[a] = ref_min(b.mn[a], mn[a]);
    b.mx[a] = ref_max(b.mx[a], mx[a]);
  }
}

// Rows 0-2 of the inverse of a column-major 4x4 matrix by 2x2 cofactors in double.
__device__ bool invert_rows(const float* m, float* out) {
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
  if (!(det != 0.0) || !isfinite(det)) return false;
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

__global__ void instance_kernel(RefitParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n_inst) return;
  float m[16];
  for (int k = 0; k < 16; ++k) m[k] = p.matrices[size_t(i) * 16 + k];
  InstanceRecord& rec = p.inst[i];
  float inv[12];
  DBox wb;
  empty_box(wb);
  if (invert_rows(m, inv)) {
    const float* g = p.geo_box + size_t(rec.geom) * 6;
    for (unsigned c = 0; c < 8; ++c) {
      const float x = (c & 1u) ? g[3] : g[0];
      const float y = (c & 2u) ? g[4] : g[1];
      const float z = (c & 4u) ? g[5] : g[2];
      float q[3];
      for (int r = 0; r < 3; ++r) q[r] = m[0 * 4 + r] * x + m[1 * 4 + r] * y + m[2 * 4 + r] * z + m[3 * 4 + r];
      enclose(wb, q, q);
    }
  } else {
    // A singular matrix: the instance can no longer be entered (its box stays empty and the
    // rows are zero) and the call's status word records it.
    for (int k = 0; k < 12; ++k) inv[k] = 0.f;
    atomicOr(p.status, 1u);
  }
  for (int k = 0; k < 12; ++k) rec.inv[k] = inv[k];
  float* o = p.inst_box + size_t(i) * 6;
  for (int a = 0; a < 3; ++a) {
    o[a] = wb.mn[a];
    o[3 + a] = wb.mx[a];
  }
}

__device__ inline double clamp_coord(double v) {
  if (isnan(v)) return 0.0;
  return fmin(fmax(v, -1e37), 1e37);
}
__device__ inline float next_toward(float f, bool up) {
  // nextafterf(f, +-inf) for finite f.
  if (f == 0.f) return up ? __uint_as_float(1u) : -__uint_as_float(1u);
  uint32_t u = __float_as_uint(f);
  u = ((f > 0.f) == up) ? u + 1u : u - 1u;
  return __uint_as_float(u);
}
__device__ inline float float_down(double v) {
  float f = float(v);
  if (double(f) > v) f = next_toward(f, false);
  return f;
}
__device__ inline float float_up(double v) {
  float f = float(v);
  if (double(f) < v) f = next_toward(f, true);
  return f;
}

// The host collapse's quantisation of child boxes (wide_bvh.cpp `quantize`), for one node.
__device__ void quantize(WideNode& n, const DBox* ch, int nc, double pad_rel) {
  for (int a = 0; a < 3; ++a) {
    float lreq[4], hreq[4];
    for (int c = 0; c < nc; ++c) {
      const DBox& bx = ch[c];
      double m = 0.0;
      for (int k = 0; k < 3; ++k) {
        m = fmax(m, fabs(clamp_coord(bx.mn[k])));
        m = fmax(m, fabs(clamp_coord(bx.mx[k])));
      }
      const double pad = pad_rel * m;
      const float cmn = bx.mn[a], cmx = bx.mx[a];
      if (!(cmn <= cmx)) {
        lreq[c] = hreq[c] = 0.f;
        continue;
      }
      const double lo = clamp_coord(double(cmn) - pad), hi = clamp_coord(double(cmx) + pad);
      lreq[c] = fminf(float_down(lo), next_toward(float(clamp_coord(cmn)), false));
      hreq[c] = fmaxf(float_up(hi), next_toward(float(clamp_coord(cmx)), true));
    }
    float origin = lreq[0];
    double top = hreq[0];
    for (int c = 1; c < nc; ++c) {
      origin = fminf(origin, lreq[c]);
      top = fmax(top, double(hreq[c]));
    }
    const double span = top - double(origin);
    // Start at or below the host's first candidate (ceil(log2(span / 255)) - 1): every exponent
    // below that cannot hold the span in 255 steps, so both searches end on the same exponent.
    int e = span > 0 ? ilogb(span / 255.0) - 1 : -126;
    if (e < -126) e = -126;
    uint8_t qlo_out[4] = {0, 0, 0, 0}, qhi_out[4] = {0, 0, 0, 0};
    for (; e <= 127; ++e) {
      const float scale = ldexpf(1.0f, e);
      bool ok = true;
      for (int c = 0; c < nc && ok; ++c) {
        const double ql = floor((double(lreq[c]) - double(origin)) / double(scale));
        int qlo = int(fmin(fmax(ql, 0.0), 255.0));
        while (qlo > 0 && fmaf(float(qlo), scale, origin) > lreq[c]) --qlo;
        const double qh = ceil((double(hreq[c]) - double(origin)) / double(scale));
        int qhi = int(fmin(fmax(qh, 0.0), 255.0));
        while (fmaf(float(qhi), scale, origin) < hreq[c]) {
          if (qhi == 255) {
            ok = false;
            break;
          }
          ++qhi;
        }
        qlo_out[c] = uint8_t(qlo);
        qhi_out[c] = uint8_t(qhi);
      }
      if (ok) break;
    }
    // Coordinates are clamped to +-1e37, so a 255-step exponent <= 127 always exists.
    n.origin[a] = origin;
    n.exp[a] = int8_t(e);
    for (int c = 0; c < 4; ++c) {
      n.qlo[a][c] = c < nc ? qlo_out[c] : 0;
      n.qhi[a][c] = c < nc ? qhi_out[c] : 0;
    }
  }
}

__global__ void level_kernel(RefitParams p, uint32_t first, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t id = p.level_nodes[first + t];
  WideNode n = p.tlas[id];
  const int nc = n.nchildren;
  DBox ch[4];
  DBox all;
  empty_box(all);
  for (int c = 0; c < nc; ++c) {
    empty_box(ch[c]);
    if (n.count[c] == 0) {
      const float* b = p.node_box + size_t(n.node_base + n.offset[c]) * 6;
      enclose(ch[c], b, b + 3);
    } else {
      const uint32_t s = n.item_base + n.offset[c];
      for (uint32_t k = 0; k < n.count[c]; ++k) {
        const float* b = p.inst_box + size_t(p.items[s + k]) * 6;
        enclose(ch[c], b, b + 3);
      }
    }
    enclose(all, ch[c].mn, ch[c].mx);
  }
  quantize(n, ch, nc, p.pad_rel);
  p.tlas[id] = n;
  float* o = p.node_box + size_t(id) * 6;
  for (int a = 0; a < 3; ++a) {
    o[a] = all.mn[a];
    o[3 + a] = all.mx[a];
  }
}

}  // namespace

int launch_refit(const RefitParams& p, const uint32_t* level_start, uint32_t n_levels, hipStream_t s) {
  if (p.n_inst) {
    instance_kernel<<<(p.n_inst + 255) / 256, 256, 0, s>>>(p);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  // level_start[l] .. level_start[l + 1] are the nodes of level l, deepest level first.
  for (uint32_t l = 0; l < n_levels; ++l) {
    const uint32_t first = level_start[l], count = level_start[l + 1] - first;
    if (!count) continue;
    level_kernel<<<(count + 63) / 64, 64, 0, s>>>(p, first, count);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  return 0;
}

}  // namespace mrt
