// Oracle (test infrastructure only): SDF tape evaluation and sphere tracing in binary32, operation
// for operation as the product's sdf.hip, following src/pt/raymarch.glsl:95-171 and the primitive /
// operator formulas of src/pathtracer/sdf.hpp:399-633. Compiled with -ffp-contract=off.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

inline float len3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline float clamp01(float v) { return fmin_(fmax_(v, 0.0f), 1.0f); }
inline float mixf(float a, float b, float h) { return a * (1.0f - h) + b * h; }

float menger(float x, float y, float z) {
  for (int n = 0; n < 4; ++n) {
    x = std::fabs(x);
    y = std::fabs(y);
    z = std::fabs(z);
    if (x < y) std::swap(x, y);
    if (x < z) std::swap(x, z);
    if (y < z) std::swap(y, z);
    z = z - 1.0f / 3.0f;
    z = -std::fabs(z);
    z = z + 1.0f / 3.0f;
    x = x * 3.0f;
    y = y * 3.0f;
    z = z * 3.0f;
    x = x - 2.0f;
    y = y - 2.0f;
  }
  const float dx = std::fabs(x) - 1.0f, dy = std::fabs(y) - 1.0f, dz = std::fabs(z) - 1.0f;
  const float inner = fmin_(fmax_(dx, fmax_(dy, dz)), 0.0f);
  return (inner + len3(fmax_(dx, 0.0f), fmax_(dy, 0.0f), fmax_(dz, 0.0f))) * (1.0f / 81.0f);
}

// Material of a smooth operator (src/pt/raymarch.glsl:43-56: `_mxm(m0, m1, t)` takes m1 when
// random_next() < t, and src/pathtracer/sdf.hpp:462-522 call it with t = 1 - h): the right
// operand's material when u < 1 - h. u is this ray's uniform number for the operator: MurmurHash3's
// 32-bit finaliser (Appleby) over the world ray's bits gives the ray's seed, and the finaliser of
// seed ^ (ordinal * 0x9E3779B9 + 0x7F4A7C15), top 24 bits, the number of the ordinal-th binary
// operator of the tape.
uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
uint32_t bits_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
uint32_t blend_seed(const float o[3], const float dd[3]) {
  uint32_t h = 0x2545F491u;
  for (int a = 0; a < 3; ++a) h = fmix32(h ^ bits_of(o[a]));
  for (int a = 0; a < 3; ++a) h = fmix32(h ^ bits_of(dd[a]));
  return h;
}
float blend_u(uint32_t seed, uint32_t ordinal) {
  return float(fmix32(seed ^ (ordinal * 0x9E3779B9u + 0x7F4A7C15u)) >> 8) * (1.0f / 16777216.0f);
}

float map(const orc_sdf_node* t, uint32_t len, float x, float y, float z, uint32_t* mat_out, uint32_t seed = 0) {
  uint32_t ordinal = 0;
  std::vector<float> d;
  std::vector<uint32_t> m;
  std::vector<float> ps;
  d.reserve(8);
  m.reserve(8);
  for (uint32_t i = 0; i < len; ++i) {
    const orc_sdf_node& nd = t[i];
    switch (nd.op) {
      case 1:
        d.push_back(len3(x, y, z) - nd.p[0]);
        m.push_back(nd.material);
        break;
      case 2: {
        const float qx = std::fabs(x) - nd.p[0], qy = std::fabs(y) - nd.p[1], qz = std::fabs(z) - nd.p[2];
        d.push_back(len3(fmax_(qx, 0.0f), fmax_(qy, 0.0f), fmax_(qz, 0.0f)) + fmin_(fmax_(qx, fmax_(qy, qz)), 0.0f));
        m.push_back(nd.material);
        break;
      }
      case 3: {
        const float qx = std::sqrt(x * x + z * z) - nd.p[0];
        d.push_back(std::sqrt(qx * qx + y * y) - nd.p[1]);
        m.push_back(nd.material);
        break;
      }
      case 4: {
        const float yy = y - fmin_(fmax_(y, 0.0f), nd.p[0]);
        d.push_back(len3(x, yy, z) - nd.p[1]);
        m.push_back(nd.material);
        break;
      }
      case 5:
        d.push_back(menger(x, y, z));
        m.push_back(nd.material);
        break;
      default:
        if (nd.op >= 16 && nd.op <= 21) {
          const float b = d.back(), a = d[d.size() - 2];
          const uint32_t mb = m.back(), ma = m[m.size() - 2];
          d.pop_back();
          m.pop_back();
          float r;
          uint32_t mr;
          ++ordinal;
          if (nd.op <= 18) {
            r = nd.op == 16 ? fmin_(a, b) : (nd.op == 17 ? fmax_(-a, b) : fmax_(a, b));
            mr = r == b ? mb : ma;
          } else {
            const float k = nd.p[0];
            float h;
            if (nd.op == 19) {
              h = clamp01(0.5f + 0.5f * (b - a) / k);
              r = mixf(b, a, h) - k * h * (1.0f - h);
            } else if (nd.op == 20) {
              h = clamp01(0.5f - 0.5f * (b + a) / k);
              r = mixf(b, -a, h) + k * h * (1.0f - h);
            } else {
              h = clamp01(0.5f - 0.5f * (b - a) / k);
              r = mixf(b, a, h) + k * h * (1.0f - h);
            }
            mr = blend_u(seed, ordinal - 1) < 1.0f - h ? mb : ma;
          }
          d.back() = r;
          m.back() = mr;
        } else if (nd.op == 32) {
          ps.push_back(x);
          ps.push_back(y);
          ps.push_back(z);
          x = x - nd.p[0];
          y = y - nd.p[1];
          z = z - nd.p[2];
        } else if (nd.op == 33) {
          ps.push_back(x);
          ps.push_back(y);
          ps.push_back(z);
          const int axis = int(nd.p[0]);
          if (axis == 0)
            x = std::fabs(x);
          else if (axis == 1)
            y = std::fabs(y);
          else
            z = std::fabs(z);
        } else if (nd.op == 34) {
          z = ps.back();
          ps.pop_back();
          y = ps.back();
          ps.pop_back();
          x = ps.back();
          ps.pop_back();
        }
    }
  }
  if (mat_out) *mat_out = m.empty() ? 0u : m.back();
  return d.empty() ? INFINITY : d.back();
}

bool clip(const orc_sdf_drawable& dr, const float o[3], const float dd[3], float& t0, float& t1) {
  float lo = -INFINITY, hi = INFINITY;
  for (int a = 0; a < 3; ++a) {
    const float inv = 1.0f / dd[a];
    const float ta = (dr.box_lo[a] - o[a]) * inv, tb = (dr.box_hi[a] - o[a]) * inv;
    if (a == 0) {
      lo = fmin_(ta, tb);
      hi = fmax_(ta, tb);
    } else {
      lo = fmax_(lo, fmin_(ta, tb));
      hi = fmin_(hi, fmax_(ta, tb));
    }
  }
  t0 = fmax_(t0, lo);
  t1 = fmin_(t1, hi * 1.0000005f);
  return t0 <= t1;
}

}  // namespace

extern "C" {

float orc_sdf_map(const orc_sdf_node* tape, uint32_t len, const float p[3], uint32_t* mat) {
  return map(tape, len, p[0], p[1], p[2], mat);
}

float orc_sdf_map_seeded(const orc_sdf_node* tape, uint32_t len, const float p[3], uint32_t seed, uint32_t* mat) {
  return map(tape, len, p[0], p[1], p[2], mat, seed);
}

uint32_t orc_sdf_blend_seed(const float o[3], const float d[3]) { return blend_seed(o, d); }

void orc_sdf_trace(const orc_sdf_node* tape, const orc_sdf_drawable* drs, uint32_t n_dr, const orc_ray* rays,
                   size_t n, int any_hit, int steps, float eps, orc_hit* hits, float* normals, int threads) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int th = 0; th < threads; ++th) {
    pool.emplace_back([=] {
      for (size_t i = size_t(th); i < n; i += size_t(threads)) {
        const orc_ray& r = rays[i];
        orc_hit& h = hits[i];
        if (!(r.d[0] != 0.f || r.d[1] != 0.f || r.d[2] != 0.f)) continue;
        float limit = r.tmax;
        if (any_hit) {
          if (std::isfinite(h.t)) continue;  // already occluded
        } else if (h.inst != 0xFFFFFFFFu) {
          limit = fmin_(limit, h.t);
        }
        float best = limit;
        int best_j = -1;
        float bp[3] = {0, 0, 0}, blen = 1;
        for (uint32_t j = 0; j < n_dr; ++j) {
          const orc_sdf_drawable& dr = drs[j];
          float t0 = r.tmin, t1 = best;
          if (!clip(dr, r.o, r.d, t0, t1)) continue;
          const float* m = dr.inv;
          float l[3], rd[3];
          for (int a = 0; a < 3; ++a) {
            l[a] = m[4 * a] * r.o[0] + m[4 * a + 1] * r.o[1] + m[4 * a + 2] * r.o[2] + m[4 * a + 3];
            rd[a] = m[4 * a] * r.d[0] + m[4 * a + 1] * r.d[1] + m[4 * a + 2] * r.d[2];
          }
          const float len = len3(rd[0], rd[1], rd[2]);
          float t = fmax_(10.0f * eps, t0 * len);
          int si = 0;
          float res = -1.0f;
          for (int it = 0; it < steps && t / len < t1; ++it) {
            const float s = t / len;
            const float dist = map(tape + dr.tape_offset, dr.tape_len, l[0] + rd[0] * s, l[1] + rd[1] * s,
                                   l[2] + rd[2] * s, nullptr);
            if (si == 0) si = dist > 0.f ? 1 : (dist < 0.f ? -1 : 0);
            if (std::fabs(dist) < eps) {
              res = t;
              break;
            }
            t = t + float(si) * dist;
          }
          if (res < 0.f) continue;
          const float tw = res / len;
          if (tw < best && tw > r.tmin) {
            best = tw;
            best_j = int(j);
            for (int a = 0; a < 3; ++a) bp[a] = l[a] + rd[a] * tw;
            blen = len;
            if (any_hit) break;
          }
        }
        if (best_j < 0) continue;
        if (any_hit) {
          h.t = 1.0f;
          continue;
        }
        const orc_sdf_drawable& dr = drs[best_j];
        float nrm[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k) {
          const float e[3] = {0.5773f * (2.0f * float(((k + 3) >> 1) & 1) - 1.0f),
                              0.5773f * (2.0f * float((k >> 1) & 1) - 1.0f), 0.5773f * (2.0f * float(k & 1) - 1.0f)};
          const float o = 0.0005f / blen;
          const float d = map(tape + dr.tape_offset, dr.tape_len, bp[0] + o * e[0], bp[1] + o * e[1],
                              bp[2] + o * e[2], nullptr) *
                          blen;
          for (int a = 0; a < 3; ++a) nrm[a] = nrm[a] + e[a] * d;
        }
        uint32_t mat = 0;
        map(tape + dr.tape_offset, dr.tape_len, bp[0], bp[1], bp[2], &mat, blend_seed(r.o, r.d));
        const float* iv = dr.inv;
        float w[3];
        for (int a = 0; a < 3; ++a) w[a] = iv[a] * nrm[0] + iv[4 + a] * nrm[1] + iv[8 + a] * nrm[2];
        const float ln = len3(w[0], w[1], w[2]);
        h.t = best;
        h.prim = mat;
        h.inst = 0x80000000u | uint32_t(best_j);
        h.b1 = h.b2 = 0.f;
        if (normals) {
          normals[3 * i + 0] = w[0] / ln;
          normals[3 * i + 1] = w[1] / ln;
          normals[3 * i + 2] = w[2] / ln;
        }
      }
    });
  }
  for (auto& t : pool) t.join();
}

}  // extern "C"
