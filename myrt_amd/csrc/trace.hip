// Closest-hit and any-hit ray traversal for gfx950.
//
// Replaces the reference's stackless binary traversal (src/pt/bvh.impl.glsl:8-72), its
// Moeller-Trumbore test (src/pt/intersect.glsl:3-28) and the two-level visitors
// (src/pt/pt_trace.comp:229-329) with:
//   * 4-wide BVH nodes of one 64-byte line each, quantised child boxes, tested with a robust slab
//     test (far distances scaled by 1 + 2*gamma(3), Ize 2013) so no box that contains a hit is
//     culled;
//   * the watertight ray/triangle test of Woop, Benthin and Wald (JCGT 2013) with the FP64
//     fallback for zero edge functions, so the hit set does not depend on traversal order;
//   * closest-hit ties broken by the lowest (instance, triangle) pair;
//   * persistent waves that refill idle lanes when fewer than refill_below of them are busy. A
//     wave takes its rays from a reserve of its own, which it refills with one atomic on the
//     global ray counter for at least `grab` rays (128 by default, MRT_TRACE_GRAB overrides);
//   * lanes at a leaf that wait until leaf_wait/64 of the wave's busy lanes are at one, so the
//     triangle and instance-entry blocks run for many lanes at a time;
//   * an instantiation for frames of one instance that enters it at refill and walks BLAS nodes
//     only;
//   * a per-lane ring stack in LDS (16 entries by default) that spills to a per-thread slice of
//     global memory when a path needs more.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.hpp"
#include "trace.hpp"

namespace mrt {
namespace dev {

constexpr int kBlock = 256;
constexpr float kRobust = 1.000001f;  // 1 + 2^-20 (8 ulp): covers the relative rounding of the slab test
constexpr float kSlack = 1.0f / 1048576.0f;  // 2^-20 of |b|: covers the absolute rounding of b
constexpr float kSlackFloor = 7.8886091e-31f;  // 2^-100: covers a plane term that underflows
constexpr float kInvMax = 18446744073709551616.0f;  // 2^64: |1/d| is clamped here for box tests

// 1/d for box tests, from the hardware reciprocal (1 ulp), clamped to [-2^64, 2^64]: a zero
// component gives +-2^64 instead of +-inf (a NaN component gives -2^64), so every plane distance
// below stays finite. Its relative error is part of what kRobust covers.
__device__ inline float box_inv(float d) { return fminf(fmaxf(__builtin_amdgcn_rcpf(d), -kInvMax), kInvMax); }

struct RayPre {  // per-space ray data for the watertight test
  float ox, oy, oz;   // origin permuted to (kx, ky, kz)
  float sx, sy, sz;   // shear constants
  int kx, ky, kz;
};

__device__ inline float sel3(int k, float x, float y, float z) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ inline void make_pre(float ox, float oy, float oz, float dx, float dy, float dz, RayPre& r) {
  const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  const float dkz = sel3(kz, dx, dy, dz);
  if (dkz < 0.f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  // One division per ray and space; the shear is the same for every triangle of the ray,
  // which is all the watertight property needs.
  const float inv = 1.0f / dkz;
  r.sx = sel3(kx, dx, dy, dz) * inv;
  r.sy = sel3(ky, dx, dy, dz) * inv;
  r.sz = inv;
  r.ox = sel3(kx, ox, oy, oz);
  r.oy = sel3(ky, ox, oy, oz);
  r.oz = sel3(kz, ox, oy, oz);
}

// Watertight test. Returns true when the ray's line crosses the triangle (either side);
// t = T / det is then the hit distance and (b1, b2) the weights of v1 and v2. The shear of each
// vertex is one fused multiply-add per coordinate: a function of the vertex and the ray alone, so
// triangles that share an edge see the same sheared endpoints and the edge functions stay
// consistent (watertight).
__device__ inline bool watertight(const RayPre& r, float4 a, float4 b, float4 c, float& t, float& b1, float& b2) {
  const float Axk = sel3(r.kx, a.x, a.y, a.z) - r.ox, Ayk = sel3(r.ky, a.x, a.y, a.z) - r.oy,
              Azk = sel3(r.kz, a.x, a.y, a.z) - r.oz;
  const float Bxk = sel3(r.kx, b.x, b.y, b.z) - r.ox, Byk = sel3(r.ky, b.x, b.y, b.z) - r.oy,
              Bzk = sel3(r.kz, b.x, b.y, b.z) - r.oz;
  const float Cxk = sel3(r.kx, c.x, c.y, c.z) - r.ox, Cyk = sel3(r.ky, c.x, c.y, c.z) - r.oy,
              Czk = sel3(r.kz, c.x, c.y, c.z) - r.oz;
  const float Ax = fmaf(-r.sx, Azk, Axk), Ay = fmaf(-r.sy, Azk, Ayk);
  const float Bx = fmaf(-r.sx, Bzk, Bxk), By = fmaf(-r.sy, Bzk, Byk);
  const float Cx = fmaf(-r.sx, Czk, Cxk), Cy = fmaf(-r.sy, Czk, Cyk);
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
    V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
    W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  const float det = U + V + W;
  if (det == 0.f) return false;
  const float Az = r.sz * Azk, Bz = r.sz * Bzk, Cz = r.sz * Czk;
  const float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
  t = T / det;
  const float rcp = 1.0f / det;
  b1 = V * rcp;
  b2 = W * rcp;
  return true;
}

__device__ inline uint32_t pack_bary(float b1, float b2) {
  const uint32_t u = uint32_t(__saturatef(b1) * 65535.f + 0.5f);
  const uint32_t v = uint32_t(__saturatef(b2) * 65535.f + 0.5f);
  return u | (v << 16);
}

__device__ inline float byte_of(uint32_t w, int c) { return float((w >> (8 * c)) & 0xffu); }

// SINGLE: a frame of one instance (TraceParams::single_instance set). Each ray enters it at
// refill and never returns to world space, so this instantiation keeps no world ray past the
// refill, walks the BLAS nodes only and has no instance-leaf or RETURN handling in its loop.
template <bool ANY, bool STATS, int WAVES, bool SINGLE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void
trace_kernel(TraceParams p) {
  extern __shared__ uint32_t lds_stack[];  // [lds_depth][kBlock]
  const uint32_t kStackLds = p.lds_depth;
  const uint32_t tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * kBlock + tid;
  const int lane = __lane_id();

  int ray_id = -1;
  bool queue_empty = false;
  uint32_t res_next = 0, res_end = 0;  // rays reserved by this wave (wave-uniform): [res_next, res_end)
  // A device-side count (a wavefront queue) never exceeds the launch's capacity.
  const uint32_t n_rays = p.n_dev ? min(*p.n_dev, p.n) : p.n;
  // The persistent grid is sized for full queues. Waves past the first ceil(n_rays / 64) leave at
  // once: the waves before them take every ray from the counter, and a short queue (the last
  // bounces of a render) then costs no counter atomic per resident wave.
  if ((gtid & ~63u) >= n_rays) return;

  float wox = 0, woy = 0, woz = 0, wdx = 0, wdy = 0, wdz = 0, wix = 0, wiy = 0, wiz = 0;
  float tmin = 0, tmax = 0;
  float ox = 0, oy = 0, oz = 0, ix = 0, iy = 0, iz = 0;
  RayPre pre{};
  float t_best = 0, b1_best = 0, b2_best = 0;
  uint32_t inst_best = 0, prim_best = 0, cur_inst = 0;
  uint32_t entry = 0, sp = 0;
  int level = 0;
  bool found = false;
  uint32_t n_nodes = 0, n_tris = 0, n_inst = 0;

  const auto push = [&](uint32_t e) {
    const uint32_t slot = sp & (kStackLds - 1);
    if (sp >= kStackLds) {
      const uint32_t deep = sp - kStackLds;
      if (deep < p.spill_depth) {
        p.spill[size_t(deep) * p.total_threads + gtid] = lds_stack[slot * kBlock + tid];
      } else {
        atomicOr(p.flags, 1u);  // overflow: the entry is dropped and the call reports an error
      }
    }
    lds_stack[slot * kBlock + tid] = e;
    ++sp;
  };
  const auto pop = [&]() -> uint32_t {
    --sp;
    const uint32_t slot = sp & (kStackLds - 1);
    const uint32_t e = lds_stack[slot * kBlock + tid];
    if (sp >= kStackLds) {
      const uint32_t deep = sp - kStackLds;
      if (deep < p.spill_depth) lds_stack[slot * kBlock + tid] = p.spill[size_t(deep) * p.total_threads + gtid];
    }
    return e;
  };

  // Moves the ray into instance `inst`'s object space; returns the instance's BLAS root.
  const auto enter = [&](uint32_t inst) -> uint32_t {
    const float4* ip = reinterpret_cast<const float4*>(p.instances + inst);
    const float4 m0 = ip[0], m1 = ip[1], m2 = ip[2];
    const uint4 meta = reinterpret_cast<const uint4*>(ip)[3];
    ox = fmaf(m0.x, wox, fmaf(m0.y, woy, fmaf(m0.z, woz, m0.w)));
    oy = fmaf(m1.x, wox, fmaf(m1.y, woy, fmaf(m1.z, woz, m1.w)));
    oz = fmaf(m2.x, wox, fmaf(m2.y, woy, fmaf(m2.z, woz, m2.w)));
    const float dx = fmaf(m0.x, wdx, fmaf(m0.y, wdy, m0.z * wdz));
    const float dy = fmaf(m1.x, wdx, fmaf(m1.y, wdy, m1.z * wdz));
    const float dz = fmaf(m2.x, wdx, fmaf(m2.y, wdy, m2.z * wdz));
    ix = box_inv(dx); iy = box_inv(dy); iz = box_inv(dz);
    make_pre(ox, oy, oz, dx, dy, dz, pre);
    cur_inst = meta.z;
    level = 1;
    if (STATS) ++n_inst;
    return meta.x;
  };

  for (;;) {
    // ---- refill idle lanes from the global ray counter ----
    const bool want = ray_id < 0 && !queue_empty;
    const uint64_t want_mask = __ballot(want);
    if (want_mask) {
      const int leader = __ffsll((unsigned long long)want_mask) - 1;
      const uint32_t cnt = uint32_t(__popcll(want_mask));
      // Lanes of rank < have take the reserve; the rest take a fresh reservation of at least
      // p.grab rays, whose remainder becomes the new reserve. Atomics on one address are
      // serialised (about 11 ns each), and one per refill (about 30 rays) kept the ray counter
      // near that limit.
      const uint32_t have = res_end - res_next;
      const uint32_t rank = uint32_t(__popcll(want_mask & ((1ull << lane) - 1ull)));
      uint32_t fresh = 0;
      const uint32_t need = cnt > have ? cnt - have : 0u;
      // Within one resident grid of the end (as far as this wave's last reservation tells), take
      // only what is needed, so that no wave sits on reserved rays while others run dry.
      const bool near_end = res_end + p.total_threads >= n_rays;
      const uint32_t grab = (need > p.grab || near_end) ? need : p.grab;
      if (need) {
        if (lane == leader) fresh = atomicAdd(p.counter, grab);
        fresh = __shfl(fresh, leader);
      }
      // base + rank is this lane's ray in both cases.
      const uint32_t base = rank < have ? res_next : fresh - have;
      if (need) {
        res_next = fresh + need;
        res_end = fresh + grab;
      } else {
        res_next += cnt;
      }
      if (want) {
        const uint32_t r = base + rank;
        if (r < n_rays) {
          const uint32_t rid = r;
          ray_id = int(rid);
          const float4* rp = reinterpret_cast<const float4*>(p.rays) + 2 * size_t(rid);
          const float4 r0 = rp[0], r1 = rp[1];
          wox = r0.x; woy = r0.y; woz = r0.z; tmin = r0.w;
          wdx = r1.x; wdy = r1.y; wdz = r1.z; tmax = r1.w;
          wix = box_inv(wdx); wiy = box_inv(wdy); wiz = box_inv(wdz);
          ox = wox; oy = woy; oz = woz; ix = wix; iy = wiy; iz = wiz;
          t_best = tmax;
          inst_best = prim_best = 0xFFFFFFFFu;
          b1_best = b2_best = 0.f;
          found = false;
          level = 0;
          sp = 0;
          entry = 0;  // TLAS root
          if (STATS) n_nodes = n_tris = n_inst = 0;
          if (!p.has_tlas || !(wdx != 0.f || wdy != 0.f || wdz != 0.f)) {
            entry = kStackReturn;  // finishes at once
          } else if (SINGLE) {
            // A frame of one instance: enter it here, with the other fresh rays of the wave,
            // instead of through the TLAS root; no RETURN marker, the ray ends with its BLAS.
            entry = enter(p.single_instance);
          }
        }
      }
    }
    if (__ballot(want && ray_id < 0)) queue_empty = true;
    if (__ballot(ray_id >= 0) == 0) break;  // every lane idle: the queue is drained

    // ---- traverse until too many lanes are idle ----
    // One step on an entry that is not a triangle leaf: an inner node (test the four children,
    // push the far ones, continue with the nearest), an instance leaf of the TLAS (move the ray
    // into object space and continue at the BLAS root) or a RETURN marker (back to world space).
    // Returns kStepPop (nothing to continue with: pop), kStepNext (continue at `next`) or kStepDone.
    // A status value rather than flags written through references, which the compiler kept in a
    // scratch slot in the single-instance instantiation.
    constexpr int kStepPop = 0, kStepNext = 1, kStepDone = 2;
    const auto node_step = [&](uint32_t e, uint32_t& next) -> int {
      int status = kStepPop;
      if (e == kStackReturn) {
        if (SINGLE || level == 0) {
          status = kStepDone;  // no scene or degenerate ray
        } else {
          level = 0;
          ox = wox; oy = woy; oz = woz; ix = wix; iy = wiy; iz = wiz;
        }
      } else if (!SINGLE && (e & 0x80000000u)) {
        // Enter an instance: the remaining instances of this leaf stay on the stack. A leaf of no
        // items (an unused slot whose box was accepted) is skipped.
        const uint32_t cnt = (e >> 24) & 0x7fu;
        const uint32_t start = e & 0xffffffu;
        if (cnt) {
          const uint32_t inst = p.tlas_items[start];
          if (cnt > 1) push(0x80000000u | ((cnt - 1) << 24) | (start + 1));
          push(kStackReturn);
          next = enter(inst);  // BLAS root node
          status = kStepNext;
        }
      } else {
        // ---- inner node: test the four quantised child boxes ----
        const WideNode* nodes = (SINGLE || level) ? p.blas_nodes : p.tlas_nodes;
        const uint4* np = reinterpret_cast<const uint4*>(nodes + e);
        const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3];
        if (STATS) ++n_nodes;
        const float orx = __uint_as_float(q0.x), ory = __uint_as_float(q0.y), orz = __uint_as_float(q0.z);
        // The per-axis exponents e (signed bytes, -126..127: 2^e is a normal power of two).
        const int ex = int(int8_t(q0.w & 0xffu)), ey = int(int8_t((q0.w >> 8) & 0xffu)),
                  ez = int(int8_t((q0.w >> 16) & 0xffu));
        // A child plane at origin + q * scale is crossed at t = q * a + b, with a = scale * i
        // (exact: scale is a power of two) and b = (origin - o) * i. Computing b rounds twice,
        // an absolute error of at most 2.2 ulp of |b|; the near planes use b - S and the far
        // planes b + S with S = 2^-20 |b| + 2^-100, so no box that the exact slab test accepts
        // is culled. The remaining relative rounding is covered by kRobust on the far side. The
        // direction signs pick which quantised plane is near on each axis.
        // a = 2^e * i: v_ldexp_f32 scales exactly as the product with 2^e would (one rounding,
        // the same result), in one instruction per axis.
        const float ax = __builtin_amdgcn_ldexpf(ix, ex), ay = __builtin_amdgcn_ldexpf(iy, ey),
                    az = __builtin_amdgcn_ldexpf(iz, ez);
        const float bx = (orx - ox) * ix, by = (ory - oy) * iy, bz = (orz - oz) * iz;
        const float sxx = fmaf(fabsf(bx), kSlack, kSlackFloor), syy = fmaf(fabsf(by), kSlack, kSlackFloor),
                    szz = fmaf(fabsf(bz), kSlack, kSlackFloor);
        const float bnx = bx - sxx, bfx = bx + sxx, bny = by - syy, bfy = by + syy, bnz = bz - szz, bfz = bz + szz;
        const bool nx = ix < 0.f, ny = iy < 0.f, nz = iz < 0.f;
        // Low codes of x, y, z in q1.x, q1.y, q1.z; high codes in q1.w, q2.x, q2.y.
        const uint32_t wnx = nx ? q1.w : q1.x, wfx = nx ? q1.x : q1.w;
        const uint32_t wny = ny ? q2.x : q1.y, wfy = ny ? q1.y : q2.x;
        const uint32_t wnz = nz ? q2.y : q1.z, wfz = nz ? q1.z : q2.y;
        float key[4];
        uint32_t ent[4] = {q3.x, q3.y, q3.z, q3.w};  // the children's stack entries, stored as such
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float tn = fmaxf(fmaxf(fmaf(byte_of(wnx, c), ax, bnx), fmaf(byte_of(wny, c), ay, bny)),
                                 fmaxf(fmaf(byte_of(wnz, c), az, bnz), tmin));
          const float tf0 = fminf(fminf(fmaf(byte_of(wfx, c), ax, bfx), fmaf(byte_of(wfy, c), ay, bfy)),
                                  fminf(fmaf(byte_of(wfz, c), az, bfz), t_best));
          // Widened by 2^-20 of its magnitude, so also a far plane behind the origin (a negative
          // t_min) moves outwards; for tf0 >= 0 this equals tf0 * kRobust bit for bit.
          const float tf = fmaf(fabsf(tf0), kRobust - 1.0f, tf0);
          const bool hit = tn <= tf;  // unused slots hold inverted boxes and harmless entries
          key[c] = hit ? tn : __int_as_float(0x7f800000);
        }
        // Sort the four (distance, entry) pairs ascending; misses carry +inf and sort last.
#define MRT_CSWAP(i, j)                                  \
  if (key[j] < key[i]) {                                 \
    const float tk = key[i]; key[i] = key[j]; key[j] = tk; \
    const uint32_t te = ent[i]; ent[i] = ent[j]; ent[j] = te; \
  }
        MRT_CSWAP(0, 1) MRT_CSWAP(2, 3) MRT_CSWAP(0, 2) MRT_CSWAP(1, 3) MRT_CSWAP(1, 2)
#undef MRT_CSWAP
        const float inf = __int_as_float(0x7f800000);
        uint32_t nhit = 0;  // plane distances are finite, so +inf marks exactly the missed children
#pragma unroll
        for (int c = 0; c < 4; ++c) nhit += (key[c] < inf) ? 1u : 0u;
        if (nhit > 3) push(ent[3]);
        if (nhit > 2) push(ent[2]);
        if (nhit > 1) push(ent[1]);
        if (nhit > 0) {
          next = ent[0];
          status = kStepNext;
        }
      }
      return status;
    };
    // The triangles of a BLAS leaf; returns true when an any-hit query is answered.
    const auto leaf_step = [&](uint32_t e) -> bool {
      const uint32_t cnt = (e >> 24) & 0x7fu;
      const uint32_t start = e & 0xffffffu;
      const float4* tp = reinterpret_cast<const float4*>(p.tris) + 3 * size_t(start);
      for (uint32_t k = 0; k < cnt; ++k) {
        const float4 a = tp[3 * k], b = tp[3 * k + 1], c = tp[3 * k + 2];
        if (STATS) ++n_tris;
        float t, b1, b2;
        if (!watertight(pre, a, b, c, t, b1, b2)) continue;
        if (!(t > tmin && t < tmax)) continue;
        const uint32_t prim = __float_as_uint(a.w);
        if (ANY) {
          found = true;
          return true;
        }
        if (t < t_best || (t == t_best && (cur_inst < inst_best || (cur_inst == inst_best && prim < prim_best)))) {
          t_best = t;
          inst_best = cur_inst;
          prim_best = prim;
          b1_best = b1;
          b2_best = b2;
          found = true;
        }
      }
      return false;
    };
    const auto finish = [&]() {
      if (ANY) {
        reinterpret_cast<uint8_t*>(p.out)[ray_id] = found ? 1 : 0;
      } else {
        uint4 h;
        h.x = found ? __float_as_uint(t_best) : 0x7f800000u;
        h.y = found ? prim_best : 0xFFFFFFFFu;
        h.z = found ? inst_best : 0xFFFFFFFFu;
        h.w = found ? pack_bary(b1_best, b2_best) : 0u;
        reinterpret_cast<uint4*>(p.out)[ray_id] = h;
      }
      if (STATS) {
        atomicAdd(p.stats + 0, 1ull);
        atomicAdd(p.stats + 1, (unsigned long long)n_nodes);
        atomicAdd(p.stats + 2, (unsigned long long)n_tris);
        atomicAdd(p.stats + 3, (unsigned long long)n_inst);
      }
      ray_id = -1;
    };
    const auto is_tri_leaf = [&](uint32_t e) {
      return (SINGLE || level == 1) && e != kStackReturn && (e & 0x80000000u);
    };

    // Every lane processes its current entry each iteration, whatever its kind, except that a
    // lane at a triangle leaf, or at an instance leaf of the TLAS, waits while fewer than
    // leaf_wait/64 of the busy lanes are at an entry of that kind and some lane has an inner node
    // to test: the wave then runs the triangle block (~150 VALU instructions for two triangles) and
    // the instance-entry block (~75) every few iterations instead of in almost every one, which is
    // what a single lane's need made it do. A ray's sequence of entries does not change.
    {
      for (;;) {
        const bool busy = ray_id >= 0;
        const bool at_leaf = busy && is_tri_leaf(entry);
        // A frame of one instance enters it at refill, so it has no instance leaves to wait at.
        const bool multi = !SINGLE;
        const bool at_inst = multi && busy && level == 0 && entry != kStackReturn && (entry & 0x80000000u);
        bool run = busy;
        if (p.leaf_wait) {
          const uint32_t n_busy = uint32_t(__popcll(__ballot(busy)));
          const uint32_t n_leaf = uint32_t(__popcll(__ballot(at_leaf)));
          const uint32_t n_inst = multi ? uint32_t(__popcll(__ballot(at_inst))) : 0u;
          bool leaf_ok = n_leaf * 64u >= p.leaf_wait * n_busy, inst_ok = n_inst * 64u >= p.leaf_wait * n_busy;
          if (n_leaf + n_inst == n_busy && !leaf_ok && !inst_ok) leaf_ok = inst_ok = true;  // no node lane
          run = busy && (at_leaf ? leaf_ok : (at_inst ? inst_ok : true));
        }
        if (run) {
          uint32_t next = 0;
          const int status = at_leaf ? (leaf_step(entry) ? kStepDone : kStepPop) : node_step(entry, next);
          bool done = status == kStepDone, have_next = status == kStepNext;
          if (!done && !have_next) {
            if (sp == 0) {
              done = true;
            } else {
              next = pop();
              have_next = true;
            }
          }
          if (done) {
            finish();
          } else {
            entry = next;
          }
        }
        const uint64_t active = __ballot(ray_id >= 0);
        if (active == 0) break;
        if (!queue_empty && __popcll(active) < p.refill_below) break;
      }
    }
  }
}

}  // namespace dev

namespace {
// The instantiation for (any, stats, waves 6 or 7, single instance).
template <bool A, bool S, int W>
const void* pick_kernel(bool single) {
  return single ? reinterpret_cast<const void*>(dev::trace_kernel<A, S, W, true>)
                : reinterpret_cast<const void*>(dev::trace_kernel<A, S, W, false>);
}
template <bool A, bool S>
const void* pick_kernel(uint32_t waves, bool single) {
  return waves == 7 ? pick_kernel<A, S, 7>(single) : pick_kernel<A, S, 6>(single);
}
const void* pick_kernel(bool any, bool stats, uint32_t waves, bool single) {
  if (any) return stats ? pick_kernel<true, true>(waves, single) : pick_kernel<true, false>(waves, single);
  return stats ? pick_kernel<false, true>(waves, single) : pick_kernel<false, false>(waves, single);
}
}  // namespace

int launch_trace(const TraceParams& p, bool any, bool stats, int grid, hipStream_t stream) {
  const size_t lds = size_t(p.lds_depth) * dev::kBlock * 4;
  const void* k = pick_kernel(any, stats, p.waves, p.single_instance != 0xFFFFFFFFu);
  TraceParams arg = p;
  void* args[] = {&arg};
  return hipLaunchKernel(k, dim3(grid), dim3(dev::kBlock), args, lds, stream) == hipSuccess ? 0 : -1;
}

int trace_blocks_per_cu(bool any, bool stats, uint32_t waves, uint32_t lds_depth, bool single) {
  int blocks = 0;
  const size_t lds = size_t(lds_depth) * dev::kBlock * 4;
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, pick_kernel(any, stats, waves, single), dev::kBlock, lds);
  if (e != hipSuccess || blocks < 1) blocks = 1;
  return blocks;
}

}  // namespace mrt
