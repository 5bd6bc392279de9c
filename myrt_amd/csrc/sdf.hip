// Sphere tracing of SDF drawables for gfx950 (src/pt/raymarch.glsl:95-171 over the tape format of
// sdf.hpp). Runs after the BVH pass on the same rays: a closer SDF surface replaces the hit
// record (instance id with bit 31 set, triangle id = material, barycentrics = octahedral normal);
// for shadow rays any SDF surface inside the segment occludes.
//
// The interpreter keeps its distance, material and position stacks in registers and shifts them
// on push/pop, so there is no dynamically indexed private array (no scratch memory).
// All binary32 arithmetic is written out; the oracle repeats it operation for operation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "sdf.hpp"
#include "sdf_kernels.hpp"

namespace mrt {
namespace sdfk {

constexpr int kBlock = 256;

struct Stack {
  float d0, d1, d2, d3, d4, d5, d6, d7;
  uint32_t m0, m1, m2, m3, m4, m5, m6, m7;
  __device__ void push(float d, uint32_t m) {
    d7 = d6; d6 = d5; d5 = d4; d4 = d3; d3 = d2; d2 = d1; d1 = d0; d0 = d;
    m7 = m6; m6 = m5; m5 = m4; m4 = m3; m3 = m2; m2 = m1; m1 = m0; m0 = m;
  }
  // Replaces the top two entries (lhs = d1, rhs = d0) by one.
  __device__ void reduce(float d, uint32_t m) {
    d0 = d; d1 = d2; d2 = d3; d3 = d4; d4 = d5; d5 = d6; d6 = d7;
    m0 = m; m1 = m2; m2 = m3; m3 = m4; m4 = m5; m5 = m6; m6 = m7;
  }
};

struct PosStack {
  float x0, y0, z0, x1, y1, z1, x2, y2, z2, x3, y3, z3;
  __device__ void push(float x, float y, float z) {
    x3 = x2; y3 = y2; z3 = z2; x2 = x1; y2 = y1; z2 = z1; x1 = x0; y1 = y0; z1 = z0;
    x0 = x; y0 = y; z0 = z;
  }
  __device__ void pop(float& x, float& y, float& z) {
    x = x0; y = y0; z = z0;
    x0 = x1; y0 = y1; z0 = z1; x1 = x2; y1 = y2; z1 = z2; x2 = x3; y2 = y3; z2 = z3;
  }
};

__device__ inline float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
__device__ inline float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ inline float mixf(float a, float b, float h) { return a * (1.0f - h) + b * h; }

__device__ inline float menger(float px, float py, float pz) {
  float x = px, y = py, z = pz;
  for (int n = 0; n < 4; ++n) {
    x = fabsf(x);
    y = fabsf(y);
    z = fabsf(z);
    if (x < y) { const float t = x; x = y; y = t; }
    if (x < z) { const float t = x; x = z; z = t; }
    if (y < z) { const float t = y; y = z; z = t; }
    z = z - 1.0f / 3.0f;
    z = -fabsf(z);
    z = z + 1.0f / 3.0f;
    x = x * 3.0f;
    y = y * 3.0f;
    z = z * 3.0f;
    x = x - 2.0f;
    y = y - 2.0f;
  }
  const float dx = fabsf(x) - 1.0f, dy = fabsf(y) - 1.0f, dz = fabsf(z) - 1.0f;
  const float inner = fminf(fmaxf(dx, fmaxf(dy, dz)), 0.0f);
  const float dis = inner + len3(fmaxf(dx, 0.0f), fmaxf(dy, 0.0f), fmaxf(dz, 0.0f));
  return dis * (1.0f / 81.0f);
}

__device__ inline float prim_distance(uint32_t kind, const SdfInsn& in, float x, float y, float z) {
  switch (kind) {
    case kSdfSphere:
      return len3(x, y, z) - in.p[0];
    case kSdfBox: {
      const float qx = fabsf(x) - in.p[0], qy = fabsf(y) - in.p[1], qz = fabsf(z) - in.p[2];
      return len3(fmaxf(qx, 0.0f), fmaxf(qy, 0.0f), fmaxf(qz, 0.0f)) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
    }
    case kSdfTorus: {
      const float qx = sqrtf(x * x + z * z) - in.p[0];
      return sqrtf(qx * qx + y * y) - in.p[1];
    }
    case kSdfCapsule: {
      const float yy = y - fminf(fmaxf(y, 0.0f), in.p[0]);
      return len3(x, yy, z) - in.p[1];
    }
    default:
      return menger(x, y, z);
  }
}

// Material of a smooth operator (raymarch.glsl:43-56, `_mxm` with t = 1 - h, sdf.hpp:462-522):
// the right operand's material with probability 1 - h, else the left one's, drawn with one
// uniform number per ray and operator. The number hashes the world ray's bits and the operator's
// ordinal among its program's binary operators (MurmurHash3's 32-bit finaliser, Appleby), so it
// does not depend on the ray's slot in a launch, and the culled chain evaluation draws the same
// number for an operator as the full interpreter. Hard operators stay deterministic, as in the
// reference (t is 0 or 1).
__device__ inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ inline uint32_t blend_seed(const float* o, const float* d) {
  uint32_t h = 0x2545F491u;
  for (int a = 0; a < 3; ++a) h = fmix32(h ^ __float_as_uint(o[a]));
  for (int a = 0; a < 3; ++a) h = fmix32(h ^ __float_as_uint(d[a]));
  return h;
}
__device__ inline float blend_u(uint32_t seed, uint32_t ordinal) {
  return float(fmix32(seed ^ (ordinal * 0x9E3779B9u + 0x7F4A7C15u)) >> 8) * (1.0f / 16777216.0f);
}

// Binary operator on (lhs d1, m1) and (rhs d2, m2); u is the operator's uniform number.
__device__ inline float combine(uint32_t op, float k, float d1, uint32_t m1, float d2, uint32_t m2, uint32_t& m,
                                float u) {
  switch (op) {
    case kSdfUnion: {
      const float r = fminf(d1, d2);
      m = r == d2 ? m2 : m1;
      return r;
    }
    case kSdfSubtract: {
      const float r = fmaxf(-d1, d2);
      m = r == d2 ? m2 : m1;
      return r;
    }
    case kSdfIntersect: {
      const float r = fmaxf(d1, d2);
      m = r == d2 ? m2 : m1;
      return r;
    }
    case kSdfSmoothUnion: {
      const float h = clamp01(0.5f + 0.5f * (d2 - d1) / k);
      m = u < 1.0f - h ? m2 : m1;
      return mixf(d2, d1, h) - k * h * (1.0f - h);
    }
    case kSdfSmoothSubtract: {
      const float h = clamp01(0.5f - 0.5f * (d2 + d1) / k);
      m = u < 1.0f - h ? m2 : m1;
      return mixf(d2, -d1, h) + k * h * (1.0f - h);
    }
    default: {  // kSdfSmoothIntersect
      const float h = clamp01(0.5f - 0.5f * (d2 - d1) / k);
      m = u < 1.0f - h ? m2 : m1;
      return mixf(d2, d1, h) + k * h * (1.0f - h);
    }
  }
}

// map(p): distance and material of a compiled program at object-space point p. The program is
// uniform across the wave, so its instructions come through scalar loads; the next one is
// fetched before the current one is evaluated.
__device__ inline float map(const SdfInsn* prog, uint32_t len, float px, float py, float pz, uint32_t& mat,
                           uint32_t seed) {
  Stack st{};
  PosStack ps{};
  float x = px, y = py, z = pz;
  uint32_t ordinal = 0;  // binary operators so far
  SdfInsn nxt = prog[0];
  for (uint32_t i = 0; i < len; ++i) {
    const SdfInsn in = nxt;
    if (i + 1 < len) nxt = prog[i + 1];
    const uint32_t kind = in.op & 0xffu;
    if (kind <= kSdfMenger) {
      float qx = x, qy = y, qz = z;
      if (in.op & kSdfFuseTranslate) {
        qx = x - in.t[0];
        qy = y - in.t[1];
        qz = z - in.t[2];
      }
      const float d = prim_distance(kind, in, qx, qy, qz);
      const uint32_t fused = (in.op >> 8) & 0xffu;
      if (fused) {
        uint32_t m;
        st.d0 = combine(fused, in.k, st.d0, st.m0, d, in.material, m, blend_u(seed, ordinal++));
        st.m0 = m;
      } else {
        st.push(d, in.material);
      }
    } else if (kind <= kSdfSmoothIntersect) {
      uint32_t m;
      const float r = combine(kind, in.p[0], st.d1, st.m1, st.d0, st.m0, m, blend_u(seed, ordinal++));
      st.reduce(r, m);
    } else if (kind == kSdfPushTranslate) {
      ps.push(x, y, z);
      x = x - in.p[0];
      y = y - in.p[1];
      z = z - in.p[2];
    } else if (kind == kSdfPushMirror) {
      ps.push(x, y, z);
      const int axis = int(in.p[0]);
      if (axis == 0) x = fabsf(x);
      else if (axis == 1) y = fabsf(y);
      else z = fabsf(z);
    } else if (kind == kSdfPop) {
      ps.pop(x, y, z);
    }
  }
  mat = st.m0;
  return st.d0;
}

// Distance of a chain primitive (sphere, box, capsule or torus: the kinds a culling grid admits).
// Lanes of one wave usually sit at different chain positions, so the kinds differ across the wave;
// sphere, box and capsule therefore share one code path: each lane selects the vector whose length
// its kind takes and the term added to that length, so a single square root serves all three and
// nothing diverges. Every lane performs exactly its own kind's binary32 operations, in order.
__device__ inline float chain_distance(const SdfInsn& in, float x, float y, float z) {
  const uint32_t kind = in.op & 0xffu;
  if (kind == kSdfTorus) {
    const float qx = sqrtf(x * x + z * z) - in.p[0];
    return sqrtf(qx * qx + y * y) - in.p[1];
  }
  const bool box = kind == kSdfBox, capsule = kind == kSdfCapsule;
  const float qx = fabsf(x) - in.p[0], qy = fabsf(y) - in.p[1], qz = fabsf(z) - in.p[2];
  const float yy = y - fminf(fmaxf(y, 0.0f), in.p[0]);
  const float ax = box ? fmaxf(qx, 0.0f) : x;
  const float ay = box ? fmaxf(qy, 0.0f) : (capsule ? yy : y);
  const float az = box ? fmaxf(qz, 0.0f) : z;
  const float l = len3(ax, ay, az);
  return box ? l + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f) : l - (capsule ? in.p[1] : in.p[0]);
}

// The culling mask of the cell a march last looked up: a ray that stays in one cell (grazing
// rays take many short steps) reads the mask once.
struct CellCache {
  uint32_t cell = 0xFFFFFFFFu;
  uint64_t mask = 0;
};

// chain_distance() at p less the instruction's fused translation, if it has one.
__device__ inline float chain_distance_at(const SdfInsn& in, float px, float py, float pz) {
  float qx = px, qy = py, qz = pz;
  if (in.op & kSdfFuseTranslate) {
    qx = px - in.t[0];
    qy = py - in.t[1];
    qz = pz - in.t[2];
  }
  return chain_distance(in, qx, qy, qz);
}

// map() of a union-chain program with a culling grid: the lane looks up the cell of its point and
// evaluates the head and the chain instructions of that cell's mask, in chain order. Dropped
// instructions leave the running distance and material unchanged bit for bit, so the result is
// map()'s. Points outside the grid evaluate every instruction.
__device__ inline float map_chain(const SdfInsn* prog, uint32_t len, const SdfGrid& g, const uint64_t* masks, float px,
                                  float py, float pz, uint32_t& mat, CellCache& cc, uint32_t seed) {
  uint64_t mask = len >= 64 ? ~0ull : ((1ull << len) - 1ull);
  const float fx = (px - g.lo[0]) * g.inv[0], fy = (py - g.lo[1]) * g.inv[1], fz = (pz - g.lo[2]) * g.inv[2];
  const float dim = float(g.dim);
  if (fx >= 0.f && fx < dim && fy >= 0.f && fy < dim && fz >= 0.f && fz < dim) {
    const uint32_t cx = uint32_t(fx), cy = uint32_t(fy), cz = uint32_t(fz);
    const uint32_t cell = (cz * g.dim + cy) * g.dim + cx;
    if (cell != cc.cell) {
      cc.mask = masks[g.mask_offset + cell];
      cc.cell = cell;
    }
    mask = cc.mask;
  }
  const SdfInsn h = prog[0];
  float qx = px, qy = py, qz = pz;
  if (h.op & kSdfFuseTranslate) {
    qx = px - h.t[0];
    qy = py - h.t[1];
    qz = pz - h.t[2];
  }
  float r = prim_distance(h.op & 0xffu, h, qx, qy, qz);
  uint32_t m = h.material;
  mask &= ~1ull;
  // Two instructions per iteration: their primitive distances do not depend on the running value,
  // so both are in flight before the two combines, which stay in chain order. A lane with one
  // instruction left evaluates it twice and keeps the first result.
  while (mask) {
    const uint32_t j0 = uint32_t(__builtin_ctzll(mask));
    mask &= mask - 1ull;
    const bool two = mask != 0ull;
    const uint32_t j1 = two ? uint32_t(__builtin_ctzll(mask)) : j0;
    mask &= mask - 1ull;
    const SdfInsn a = prog[j0], b = prog[j1];
    const float da = chain_distance_at(a, px, py, pz), db = chain_distance_at(b, px, py, pz);
    uint32_t ma, mb;
    // Chain instruction j carries binary operator j - 1.
    const float ra = combine((a.op >> 8) & 0xffu, a.k, r, m, da, a.material, ma, blend_u(seed, j0 - 1));
    const float rb = combine((b.op >> 8) & 0xffu, b.k, ra, ma, db, b.material, mb, blend_u(seed, j1 - 1));
    r = two ? rb : ra;
    m = two ? mb : ma;
  }
  mat = m;
  return r;
}

// Distance of drawable program at a point, culled when the program has a grid. With CHAIN every
// drawable of the launch has a grid, and the general interpreter (whose stacks hold 28 registers)
// is left out of the kernel.
template <bool CHAIN = false>
__device__ inline float map_any(const SdfInsn* prog, uint32_t len, const SdfGrid& g, const uint64_t* masks, float px,
                                float py, float pz, uint32_t& mat, CellCache& cc, uint32_t seed = 0) {
  if (CHAIN) return map_chain(prog, len, g, masks, px, py, pz, mat, cc, seed);
  return g.dim ? map_chain(prog, len, g, masks, px, py, pz, mat, cc, seed) : map(prog, len, px, py, pz, mat, seed);
}

__device__ inline uint32_t oct_encode(float x, float y, float z) {
  const float s = fabsf(x) + fabsf(y) + fabsf(z);
  float u = x / s, v = y / s;
  if (z < 0.f) {
    const float uu = (1.0f - fabsf(v)) * (u >= 0.f ? 1.f : -1.f);
    const float vv = (1.0f - fabsf(u)) * (v >= 0.f ? 1.f : -1.f);
    u = uu;
    v = vv;
  }
  const uint32_t a = uint32_t(__saturatef(u * 0.5f + 0.5f) * 65535.f + 0.5f);
  const uint32_t b = uint32_t(__saturatef(v * 0.5f + 0.5f) * 65535.f + 0.5f);
  return a | (b << 16);
}

// Slab clip of the world ray against a drawable's box: returns false if it misses [t0, t1].
__device__ inline bool clip(const SdfDrawable& dr, float ox, float oy, float oz, float dx, float dy, float dz, float& t0,
                            float& t1) {
  const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
  float a = (dr.box_lo[0] - ox) * ix, b = (dr.box_hi[0] - ox) * ix;
  float lo = fminf(a, b), hi = fmaxf(a, b);
  a = (dr.box_lo[1] - oy) * iy;
  b = (dr.box_hi[1] - oy) * iy;
  lo = fmaxf(lo, fminf(a, b));
  hi = fminf(hi, fmaxf(a, b));
  a = (dr.box_lo[2] - oz) * iz;
  b = (dr.box_hi[2] - oz) * iz;
  lo = fmaxf(lo, fminf(a, b));
  hi = fminf(hi, fmaxf(a, b));
  t0 = fmaxf(t0, lo);
  t1 = fminf(t1, hi * 1.0000005f);
  return t0 <= t1;
}

// march_sdf (raymarch.glsl:116-171): world hit distance or -1.
template <bool CHAIN>
__device__ inline float march(const SdfDrawable& dr, const SdfInsn* tape, const SdfGrid& grid, const uint64_t* masks,
                              float ox, float oy, float oz, float dx, float dy, float dz, float t_enter, float t_exit,
                              int steps, float eps, float& lx, float& ly, float& lz, float& len_out) {
  const float* m = dr.inv;
  lx = m[0] * ox + m[1] * oy + m[2] * oz + m[3];
  ly = m[4] * ox + m[5] * oy + m[6] * oz + m[7];
  lz = m[8] * ox + m[9] * oy + m[10] * oz + m[11];
  const float rx = m[0] * dx + m[1] * dy + m[2] * dz;
  const float ry = m[4] * dx + m[5] * dy + m[6] * dz;
  const float rz = m[8] * dx + m[9] * dy + m[10] * dz;
  const float len = len3(rx, ry, rz);
  len_out = len;
  float t = fmaxf(10.0f * eps, t_enter * len);
  int si = 0;
  float res = -1.0f;
  CellCache cc;
  for (int i = 0; i < steps && t / len < t_exit; ++i) {
    const float s = t / len;
    uint32_t mat;
    const float d = map_any<CHAIN>(tape, dr.tape_len, grid, masks, lx + rx * s, ly + ry * s, lz + rz * s, mat, cc);
    if (si == 0) si = d > 0.f ? 1 : (d < 0.f ? -1 : 0);
    if (fabsf(d) < eps) {
      res = t;
      break;
    }
    t = t + float(si) * d;
  }
  // Keep the object-space ray for the normal.
  const float tw = res < 0.f ? -1.0f : res / len;
  lx = res < 0.f ? lx : lx + rx * tw;
  ly = res < 0.f ? ly : ly + ry * tw;
  lz = res < 0.f ? lz : lz + rz * tw;
  return tw;
}

// Appends `ray` to the list of SDF hits (called by the lanes that found one; one atomic per wave).
__device__ inline void append_hit(const SdfParams& p, uint32_t ray) {
  const uint64_t lanes = __ballot(1);
  const int lane = __lane_id(), leader = __ffsll((unsigned long long)lanes) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(p.hit_count, uint32_t(__popcll(lanes)));
  base = __shfl(base, leader);
  p.hit_list[base + uint32_t(__popcll(lanes & ((1ull << lane) - 1ull)))] = ray;
}

// Second pass of a closest-hit query: every record the march marked as an SDF hit (instance bit
// 31) gets its material and normal, the tetrahedral gradient of raymarch.glsl:100-114 (offset
// 0.0005 / len) at the object-space hit point the march left in `side`, turned to world space.
// Kept out of the marching kernel so that its registers do not lower that kernel's occupancy.
// Material and normal of one ray whose record the march marked as an SDF hit.
__device__ inline void sdf_normal(const SdfParams& p, uint32_t ray) {
  uint4 h = reinterpret_cast<const uint4*>(p.out)[ray];
  if (h.z == 0xFFFFFFFFu || !(h.z & 0x80000000u)) return;
  const int best_j = int(h.z & 0x7fffffffu);
  const float4 sp = p.side[ray];
  const float bx = sp.x, by = sp.y, bz = sp.z, blen = sp.w;
  const SdfDrawable dr = p.drawables[best_j];
  const SdfInsn* tape = p.prog + dr.tape_offset;
  const SdfGrid grid = p.grids[best_j];
  float nx = 0, ny = 0, nz = 0;
  CellCache cc;
  for (int k = 0; k < 4; ++k) {
    const float ex = 0.5773f * (2.0f * float(((k + 3) >> 1) & 1) - 1.0f);
    const float ey = 0.5773f * (2.0f * float((k >> 1) & 1) - 1.0f);
    const float ez = 0.5773f * (2.0f * float(k & 1) - 1.0f);
    const float o = 0.0005f / blen;
    uint32_t mm;
    const float d = map_any(tape, dr.tape_len, grid, p.masks, bx + o * ex, by + o * ey, bz + o * ez, mm, cc) * blen;
    nx = nx + ex * d;
    ny = ny + ey * d;
    nz = nz + ez * d;
  }
  const float* wr = reinterpret_cast<const float*>(p.rays) + size_t(ray) * 8;  // mrt_ray: o, tmin, d, tmax
  uint32_t mat;
  map_any(tape, dr.tape_len, grid, p.masks, bx, by, bz, mat, cc, blend_seed(wr, wr + 4));
  const float* iv = dr.inv;  // world normal = (M^-1)^T n
  float wx = iv[0] * nx + iv[4] * ny + iv[8] * nz;
  float wy = iv[1] * nx + iv[5] * ny + iv[9] * nz;
  float wz = iv[2] * nx + iv[6] * ny + iv[10] * nz;
  const float l = len3(wx, wy, wz);
  wx = wx / l;
  wy = wy / l;
  wz = wz / l;
  h.y = mat;
  h.w = oct_encode(wx, wy, wz);
  reinterpret_cast<uint4*>(p.out)[ray] = h;
}

// With a hit list the pass walks only the listed rays (grid-stride), else every ray.
__global__ __launch_bounds__(kBlock) void sdf_normal_kernel(SdfParams p) {
  const uint32_t count = p.hit_list ? min(*p.hit_count, p.n) : (p.n_dev ? min(*p.n_dev, p.n) : p.n);
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock)
    sdf_normal(p, p.hit_list ? p.hit_list[k] : k);
}


// One thread per ray (the sequential algorithm as it stands): clip every drawable, march the
// ones the segment [tmin, best] crosses, keep the closest hit. A wave waits for its longest march;
// a persistent form that refilled finished lanes measured slower in the render.
template <bool ANY, bool CHAIN>
__device__ __attribute__((always_inline)) inline void march_per_ray(const SdfParams& p, const SdfInsn* prog) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = p.n_dev ? min(*p.n_dev, p.n) : p.n;
  if (i >= n) return;
  const float4* rp = reinterpret_cast<const float4*>(p.rays) + 2 * size_t(i);
  const float4 r0 = rp[0], r1 = rp[1];
  if (!(r1.x != 0.f || r1.y != 0.f || r1.z != 0.f)) return;
  float limit = r1.w;
  if (ANY) {
    if (reinterpret_cast<const uint8_t*>(p.out)[i]) return;
  } else {
    const uint4 h = reinterpret_cast<const uint4*>(p.out)[i];
    if (h.z != 0xFFFFFFFFu) limit = fminf(limit, __uint_as_float(h.x));
  }
  float best = limit;
  int best_j = -1;
  float bx = 0, by = 0, bz = 0, blen = 1;
  for (uint32_t j = 0; j < p.num_drawables; ++j) {
    const SdfDrawable dr = p.drawables[j];
    float t0 = r0.w, t1 = best;
    if (!clip(dr, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, t0, t1)) continue;
    float lx, ly, lz, len;
    const SdfGrid grid = p.grids[j];
    const float tw = march<CHAIN>(dr, prog + dr.tape_offset, grid, p.masks, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, t0, t1,
                           p.steps, p.eps, lx, ly, lz, len);
    if (tw >= 0.f && tw < best && tw > r0.w) {
      best = tw;
      best_j = int(j);
      bx = lx;
      by = ly;
      bz = lz;
      blen = len;
      if (ANY) break;
    }
  }
  if (best_j < 0) return;
  if (ANY) {
    reinterpret_cast<uint8_t*>(p.out)[i] = 1;
  } else {
    reinterpret_cast<uint4*>(p.out)[i] = make_uint4(__float_as_uint(best), 0u, 0x80000000u | uint32_t(best_j), 0u);
    p.side[i] = make_float4(bx, by, bz, blen);
    if (p.hit_list) append_hit(p, i);
  }
}

// The compiled programs of all drawables are copied to LDS when they fit (kProgCache
// instructions), so the per-lane instruction fetches of the culled chains are LDS reads rather
// than L1/L2 hits.
constexpr uint32_t kProgCache = 256;

// Waves per SIMD the marching kernels are compiled for. The march is latency bound (dependent
// square-root and division chains, mask and program fetches), so occupancy pays even where the
// register budget of 8 waves spills a few values to scratch: render of the 4M-triangle + SDF
// scene 0.130 s at 4 waves, 0.124 s at 8; with the chain-only kernel 0.125 s at 6, 0.123 s at 8.
#ifndef MRT_SDF_RAY_WAVES
#define MRT_SDF_RAY_WAVES 8
#endif
#ifndef MRT_SDF_CHAIN_WAVES
#define MRT_SDF_CHAIN_WAVES 8
#endif
template <bool ANY, bool CHAIN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    CHAIN ? MRT_SDF_CHAIN_WAVES : MRT_SDF_RAY_WAVES, 8))) void sdf_kernel(SdfParams p) {
  __shared__ uint4 s_prog[kProgCache * (sizeof(SdfInsn) / 16)];
  static_assert(sizeof(SdfInsn) % 16 == 0, "instruction records are whole 16-byte words");
  // The grid covers the queue's capacity; a block past its live rays leaves before copying the
  // program (the whole block at once, so the barrier below is never split).
  if (blockIdx.x * kBlock >= (p.n_dev ? min(*p.n_dev, p.n) : p.n)) return;
  const bool cached = p.prog_len <= kProgCache;
  if (cached) {
    const uint4* src = reinterpret_cast<const uint4*>(p.prog);
    const uint32_t words = p.prog_len * uint32_t(sizeof(SdfInsn) / 16);
    for (uint32_t k = threadIdx.x; k < words; k += kBlock) s_prog[k] = src[k];
  }
  __syncthreads();
  // Two code paths, so that each knows where its program lives: LDS reads for the cached copy
  // (a pointer that may point to either would make every fetch a flat load, which takes the
  // latency of the vector memory path).
  if (cached)
    march_per_ray<ANY, CHAIN>(p, reinterpret_cast<const SdfInsn*>(s_prog));
  else
    march_per_ray<ANY, CHAIN>(p, p.prog);
}

}  // namespace sdfk

namespace {
template <bool ANY, bool CHAIN>
void launch_march(const SdfParams& p, hipStream_t stream) {
  const int grid = std::max(int((p.n + sdfk::kBlock - 1) / sdfk::kBlock), 1);
  hipLaunchKernelGGL((sdfk::sdf_kernel<ANY, CHAIN>), dim3(grid), dim3(sdfk::kBlock), 0, stream, p);
}
template <bool ANY>
void launch_march_any(const SdfParams& p, hipStream_t stream) {
  if (p.chain_only)
    launch_march<ANY, true>(p, stream);
  else
    launch_march<ANY, false>(p, stream);
}
}  // namespace

int launch_sdf(const SdfParams& p, bool any, int num_cus, hipStream_t stream) {
  if (any) {
    launch_march_any<true>(p, stream);
  } else {
    launch_march_any<false>(p, stream);
    unsigned nb = unsigned((p.n + sdfk::kBlock - 1) / sdfk::kBlock);
    if (p.hit_list) nb = std::min(nb, unsigned(std::max(num_cus, 1)) * 16u);  // grid-stride over the hits
    hipLaunchKernelGGL(sdfk::sdf_normal_kernel, dim3(nb ? nb : 1), dim3(sdfk::kBlock), 0, stream, p);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
