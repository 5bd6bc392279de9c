// Oracle (measurement infrastructure only): a CPU replica of the product traversal kernel
// (myrt_amd/csrc/trace.hip, trace_kernel) walking the product's own device-format arrays, as
// mrt_scene_copy_wide returns them, to count per ray the 64-B wide nodes, 48-B triangle records
// and 64-B instance records the kernel fetches. The roofline's bytes per ray
// (32 + 16 + 64 N_node + 48 N_tri + 64 N_inst) take these counts.
//
// Each ray's walk is the kernel's: the one-FMA slab test with the same slack terms and far-plane
// widening, the 5-comparator child sort (misses sort last), nearest child next and the others
// pushed far to near, instance entry through M^-1 with the watertight shear, closest hits by
// (t, instance, triangle) and any hits ending at the first hit, and the single-instance frame
// entered at ray start. A lane's walk does not depend on its wave, so the per-ray counts are the
// kernel's. One difference is documented, not hidden: the kernel takes 1/d from the hardware
// reciprocal (1 ulp), this replica from the correctly rounded division, so a box test on the edge
// of a ray's slab can differ and the totals agree to a small fraction of a per cent.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "ray_math.hpp"

namespace {
using namespace orc_math;

struct WideNode {
  float origin[3];
  int8_t exp[3];
  uint8_t nchildren;
  uint8_t qlo[3][4];
  uint8_t qhi[3][4];
  uint32_t pad[2];
  uint32_t entry[4];  // an inner child's node index, or 0x80000000 | count << 24 | first item
};
static_assert(sizeof(WideNode) == 64, "wide node record");
struct TriRecord {
  float v0[3];
  uint32_t prim;
  float v1[3];
  uint32_t pad1;
  float v2[3];
  uint32_t pad2;
};
static_assert(sizeof(TriRecord) == 48, "triangle record");
struct InstanceRecord {
  float inv[12];
  uint32_t blas_root, geom, inst, pad;
};
static_assert(sizeof(InstanceRecord) == 64, "instance record");

constexpr uint32_t kReturn = 0xFFFFFFFFu;
constexpr float kRobust = 1.000001f;
constexpr float kSlack = 1.0f / 1048576.0f;
constexpr float kSlackFloor = 7.8886091e-31f;
constexpr float kInvMax = 18446744073709551616.0f;

inline float box_inv(float d) { return std::fmin(std::fmax(1.0f / d, -kInvMax), kInvMax); }
inline float exp2i(int e) {
  const uint32_t bits = uint32_t(e + 127) << 23;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

struct Arrays {
  const WideNode* blas;
  const TriRecord* tris;
  const WideNode* tlas;
  const uint32_t* items;
  const InstanceRecord* inst;
  bool has_tlas;
  uint32_t single_instance;
};

struct Counts {
  uint32_t nodes = 0, tris = 0, inst = 0;
};

// One ray, as one lane of trace_kernel runs it.
Counts walk(const Arrays& A, const orc_ray& r, bool any, orc_hit* hit_out) {
  Counts c;
  const float wo[3] = {r.o[0], r.o[1], r.o[2]}, wd[3] = {r.d[0], r.d[1], r.d[2]};
  const float tmin = r.tmin, tmax = r.tmax;
  const float wi[3] = {box_inv(wd[0]), box_inv(wd[1]), box_inv(wd[2])};
  float o[3] = {wo[0], wo[1], wo[2]}, iv[3] = {wi[0], wi[1], wi[2]};
  Pre pre{};
  float t_best = tmax, b1_best = 0, b2_best = 0;
  uint32_t inst_best = 0xFFFFFFFFu, prim_best = 0xFFFFFFFFu, cur_inst = 0;
  int level = 0;
  bool found = false;
  std::vector<uint32_t> stack;
  stack.reserve(64);

  const auto enter = [&](uint32_t k) -> uint32_t {
    const InstanceRecord& ir = A.inst[k];
    float d[3];
    to_object(ir.inv, wo, wd, o, d);
    for (int a = 0; a < 3; ++a) iv[a] = box_inv(d[a]);
    make_pre(o, d, pre);
    cur_inst = ir.inst;
    level = 1;
    ++c.inst;
    return ir.blas_root;
  };

  uint32_t entry = 0;
  if (!A.has_tlas || !(wd[0] != 0.f || wd[1] != 0.f || wd[2] != 0.f))
    entry = kReturn;
  else if (A.single_instance != 0xFFFFFFFFu)
    entry = enter(A.single_instance);

  for (;;) {
    bool have_next = false, done = false;
    uint32_t next = 0;
    const bool tri_leaf = level == 1 && entry != kReturn && (entry & 0x80000000u);
    if (tri_leaf) {
      const uint32_t cnt = (entry >> 24) & 0x7fu, start = entry & 0xffffffu;
      for (uint32_t k = 0; k < cnt && !done; ++k) {
        const TriRecord& tr = A.tris[start + k];
        ++c.tris;
        float t, b1, b2;
        if (!watertight(pre, tr.v0, tr.v1, tr.v2, t, b1, b2)) continue;
        if (!(t > tmin && t < tmax)) continue;
        if (any) {
          found = true;
          done = true;
          break;
        }
        const uint32_t prim = tr.prim;
        if (t < t_best || (t == t_best && (cur_inst < inst_best || (cur_inst == inst_best && prim < prim_best)))) {
          t_best = t;
          inst_best = cur_inst;
          prim_best = prim;
          b1_best = b1;
          b2_best = b2;
          found = true;
        }
      }
    } else if (entry == kReturn) {
      if (level == 0) {
        done = true;
      } else {
        level = 0;
        for (int a = 0; a < 3; ++a) {
          o[a] = wo[a];
          iv[a] = wi[a];
        }
      }
    } else if (entry & 0x80000000u) {
      const uint32_t cnt = (entry >> 24) & 0x7fu, start = entry & 0xffffffu;
      if (cnt) {  // a leaf of no items is an unused slot
        const uint32_t k = A.items[start];
        if (cnt > 1) stack.push_back(0x80000000u | ((cnt - 1) << 24) | (start + 1));
        stack.push_back(kReturn);
        next = enter(k);
        have_next = true;
      }
    } else {
      const WideNode& n = (level ? A.blas : A.tlas)[entry];
      ++c.nodes;
      const float sc[3] = {exp2i(n.exp[0]), exp2i(n.exp[1]), exp2i(n.exp[2])};
      float a_[3], bn[3], bf[3];
      for (int a = 0; a < 3; ++a) {
        a_[a] = sc[a] * iv[a];
        const float b = (n.origin[a] - o[a]) * iv[a];
        const float s = std::fmaf(std::fabs(b), kSlack, kSlackFloor);
        bn[a] = b - s;
        bf[a] = b + s;
      }
      float key[4];
      uint32_t ent[4];
      for (int ch = 0; ch < 4; ++ch) {
        float tn3[3], tf3[3];
        for (int a = 0; a < 3; ++a) {
          const bool neg = iv[a] < 0.f;
          const float qn = float(neg ? n.qhi[a][ch] : n.qlo[a][ch]);
          const float qf = float(neg ? n.qlo[a][ch] : n.qhi[a][ch]);
          tn3[a] = std::fmaf(qn, a_[a], bn[a]);
          tf3[a] = std::fmaf(qf, a_[a], bf[a]);
        }
        const float tn = std::fmax(std::fmax(tn3[0], tn3[1]), std::fmax(tn3[2], tmin));
        const float tf0 = std::fmin(std::fmin(tf3[0], tf3[1]), std::fmin(tf3[2], t_best));
        const float tf = std::fmaf(std::fabs(tf0), kRobust - 1.0f, tf0);
        const bool hit = tn <= tf;  // unused slots: inverted boxes, harmless entries
        ent[ch] = n.entry[ch];
        key[ch] = hit ? tn : INFINITY;
      }
      const auto cswap = [&](int i, int j) {
        if (key[j] < key[i]) {
          std::swap(key[i], key[j]);
          std::swap(ent[i], ent[j]);
        }
      };
      cswap(0, 1);
      cswap(2, 3);
      cswap(0, 2);
      cswap(1, 3);
      cswap(1, 2);
      uint32_t nhit = 0;
      for (int ch = 0; ch < 4; ++ch) nhit += key[ch] < INFINITY ? 1u : 0u;
      if (nhit > 3) stack.push_back(ent[3]);
      if (nhit > 2) stack.push_back(ent[2]);
      if (nhit > 1) stack.push_back(ent[1]);
      if (nhit > 0) {
        next = ent[0];
        have_next = true;
      }
    }
    if (!done && !have_next) {
      if (stack.empty()) {
        done = true;
      } else {
        next = stack.back();
        stack.pop_back();
        have_next = true;
      }
    }
    if (done) break;
    entry = next;
  }
  if (hit_out) {
    hit_out->t = found ? (any ? 0.f : t_best) : INFINITY;
    hit_out->prim = found && !any ? prim_best : 0xFFFFFFFFu;
    hit_out->inst = found && !any ? inst_best : (found ? 0u : 0xFFFFFFFFu);
    hit_out->b1 = b1_best;
    hit_out->b2 = b2_best;
  }
  return c;
}

}  // namespace

extern "C" int orc_fetch_counts(const void* blas_nodes, const void* tris, const void* tlas_nodes,
                                const uint32_t* tlas_items, const void* instances, int has_tlas,
                                uint32_t single_instance, const orc_ray* rays, size_t n, int any_hit,
                                uint32_t* counts, orc_hit* hits, int threads) {
  const Arrays A{static_cast<const WideNode*>(blas_nodes), static_cast<const TriRecord*>(tris),
                 static_cast<const WideNode*>(tlas_nodes), tlas_items, static_cast<const InstanceRecord*>(instances),
                 has_tlas != 0, single_instance};
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (size_t i = size_t(t); i < n; i += size_t(threads)) {
        const Counts c = walk(A, rays[i], any_hit != 0, hits ? hits + i : nullptr);
        counts[3 * i] = c.nodes;
        counts[3 * i + 1] = c.tris;
        counts[3 * i + 2] = c.inst;
      }
    });
  for (auto& th : pool) th.join();
  return 0;
}
