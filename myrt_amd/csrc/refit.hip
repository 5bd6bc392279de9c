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
// A TLAS of up to kRefitOneBlockNodes nodes runs both passes in one workgroup and one launch
// (refit_one_block_kernel), with a barrier between levels: at a few thousand instances the
// per-level launches would cost more than their work. Everything is fixed-order float/double
// arithmetic with contraction off (-ffp-contract=off), so host and device agree to the bit.
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

__device__ void update_instance(const RefitParams& p, uint32_t i) {
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

__global__ void instance_kernel(RefitParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.n_inst) update_instance(p, i);
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

// The host collapse's quantisation of child boxes (wide_bvh.cpp `quantize`) on one axis, given
// the children's extents on that axis and their paddings. Loops over children are unrolled over
// the 4 slots with a guard, so per-child values stay in registers (dynamically indexed arrays
// would live in scratch, one memory round trip per access). Writes origin, exponent and the
// child codes of the axis straight to the node in memory.
__device__ void quantize_axis(WideNode* out, int a, const float (&cmn)[4], const float (&cmx)[4],
                              const double (&pad)[4], int nc) {
  float lreq[4], hreq[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    lreq[c] = hreq[c] = 0.f;
    if (c < nc && cmn[c] <= cmx[c]) {  // an empty or NaN box keeps a point box at 0
      const double lo = clamp_coord(double(cmn[c]) - pad[c]), hi = clamp_coord(double(cmx[c]) + pad[c]);
      lreq[c] = fminf(float_down(lo), next_toward(float(clamp_coord(cmn[c])), false));
      hreq[c] = fmaxf(float_up(hi), next_toward(float(clamp_coord(cmx[c])), true));
    }
  }
  float origin = lreq[0];
  double top = hreq[0];
#pragma unroll
  for (int c = 1; c < 4; ++c) {
    if (c < nc) {
      origin = fminf(origin, lreq[c]);
      top = fmax(top, double(hreq[c]));
    }
  }
  const double span = top - double(origin);
  // The host's first candidate, ceil(log2(x)) - 1 with x = span / 255, exactly: ilogb(x), less one
  // when x is a power of two. Every smaller exponent leaves half the span or more outside 255
  // steps, so both searches try the same exponents from here on.
  int e = -126;
  if (span > 0) {
    const double x = span / 255.0;
    e = ilogb(x);
    if (ldexp(1.0, e) == x) --e;
  }
  if (e < -126) e = -126;
  uint32_t lo_codes = 0, hi_codes = 0;
  for (; e <= 127; ++e) {
    const float scale = ldexpf(1.0f, e);
    bool ok = true;
    lo_codes = hi_codes = 0;
    for (int c = nc; c < 4; ++c) lo_codes |= 255u << (8 * c);  // unused slots keep an inverted box
    // All children are evaluated even after one fails (a failed exponent is discarded anyway),
    // so their independent chains overlap.
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c < nc) {
        const double ql = floor((double(lreq[c]) - double(origin)) / double(scale));
        int lo = int(fmin(fmax(ql, 0.0), 255.0));
        while (lo > 0 && fmaf(float(lo), scale, origin) > lreq[c]) --lo;
        const double qh = ceil((double(hreq[c]) - double(origin)) / double(scale));
        int hi = int(fmin(fmax(qh, 0.0), 255.0));
        while (fmaf(float(hi), scale, origin) < hreq[c] && hi < 255) ++hi;
        ok = ok && !(fmaf(float(hi), scale, origin) < hreq[c]);
        lo_codes |= uint32_t(lo) << (8 * c);
        hi_codes |= uint32_t(hi) << (8 * c);
      }
    }
    if (ok) break;
  }
  // Coordinates are clamped to +-1e37, so a 255-step exponent <= 127 always exists.
  out->origin[a] = origin;
  out->exp[a] = int8_t(e);
  *reinterpret_cast<uint32_t*>(out->qlo[a]) = lo_codes;
  *reinterpret_cast<uint32_t*>(out->qhi[a]) = hi_codes;
}

// One axis of one node: the union of the node's children on that axis (for its parent) and the
// requantisation of the children's codes on that axis. The paddings need each child's whole box.
__device__ void refit_node_axis(const RefitParams& p, uint32_t id, int a) {
  WideNode* node = p.tlas + id;
  const int nc = node->nchildren;
  const uint4 ents = *reinterpret_cast<const uint4*>(node->entry);
  float mn[4], mx[4];
  double pad[4];
  float all_mn = FLT_MAX, all_mx = -FLT_MAX;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    DBox b;
    empty_box(b);
    if (c < nc) {
      const uint32_t e = c == 0 ? ents.x : (c == 1 ? ents.y : (c == 2 ? ents.z : ents.w));
      if (!(e & 0x80000000u)) {
        const float* q = p.node_box + size_t(e) * 6;
        enclose(b, q, q + 3);
      } else {
        const uint32_t count = (e >> 24) & 0x7fu, first = e & 0xffffffu;
        for (uint32_t k = 0; k < count; ++k) {
          const float* q = p.inst_box + size_t(p.items[first + k]) * 6;
          enclose(b, q, q + 3);
        }
      }
    }
    double m = 0.0;
    for (int k = 0; k < 3; ++k) {
      m = fmax(m, fabs(clamp_coord(b.mn[k])));
      m = fmax(m, fabs(clamp_coord(b.mx[k])));
    }
    pad[c] = p.pad_rel * m;
    mn[c] = a == 0 ? b.mn[0] : (a == 1 ? b.mn[1] : b.mn[2]);
    mx[c] = a == 0 ? b.mx[0] : (a == 1 ? b.mx[1] : b.mx[2]);
    if (c < nc) {
      all_mn = ref_min(all_mn, mn[c]);
      all_mx = ref_max(all_mx, mx[c]);
    }
  }
  quantize_axis(node, a, mn, mx, pad, nc);
  p.node_box[size_t(id) * 6 + a] = all_mn;
  p.node_box[size_t(id) * 6 + 3 + a] = all_mx;
}

// Three threads per node, one per axis.
__global__ void __launch_bounds__(64) level_kernel(RefitParams p, uint32_t first, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * count) refit_node_axis(p, p.level_nodes[first + t / 3], int(t % 3));
}

// A small TLAS in one workgroup: the status word, the instances and every level, with a barrier
// between the passes, in one launch (the multi-launch form is bound by launch latency there).
__global__ void __launch_bounds__(1024) refit_one_block_kernel(RefitParams p, const uint32_t* level_start,
                                                               uint32_t n_levels) {
  if (threadIdx.x == 0) *p.status = 0u;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < p.n_inst; i += blockDim.x) update_instance(p, i);
  __syncthreads();
  for (uint32_t l = 0; l < n_levels; ++l) {
    const uint32_t first = level_start[l], end = level_start[l + 1];
    for (uint32_t t = 3 * first + threadIdx.x; t < 3 * end; t += blockDim.x)
      refit_node_axis(p, p.level_nodes[t / 3], int(t % 3));
    __syncthreads();
  }
}

}  // namespace

int launch_refit_one_block(const RefitParams& p, const uint32_t* d_level_start, uint32_t n_levels, hipStream_t s) {
  refit_one_block_kernel<<<1, 1024, 0, s>>>(p, d_level_start, n_levels);
  return hipGetLastError() != hipSuccess ? 1 : 0;
}

int launch_refit(const RefitParams& p, const uint32_t* level_start, uint32_t n_levels, hipStream_t s) {
  if (p.n_inst) {
    instance_kernel<<<(p.n_inst + 255) / 256, 256, 0, s>>>(p);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  // level_start[l] .. level_start[l + 1] are the nodes of level l, deepest level first.
  for (uint32_t l = 0; l < n_levels; ++l) {
    const uint32_t first = level_start[l], count = level_start[l + 1] - first;
    if (!count) continue;
    level_kernel<<<(3 * count + 63) / 64, 64, 0, s>>>(p, first, count);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  return 0;
}

}  // namespace mrt
