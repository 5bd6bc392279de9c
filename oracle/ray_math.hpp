// Oracle (test infrastructure only): the ray arithmetic the product kernels use, in the same
// binary32 operation order (compiled with -ffp-contract=off): the watertight ray/triangle test
// (Woop, Benthin and Wald, JCGT 2013) with its binary64 fallback, and the object-space ray of an
// instance. Shared by the oracle's queries (trace_ref.cpp) and the fetch-count replica
// (fetch_count.cpp).
#pragma once
#include <cmath>
#include <utility>

namespace orc_math {

// ---- watertight test, same arithmetic order as the product kernel --------------------------
struct Pre {
  float o[3];  // permuted origin
  float sx, sy, sz;
  int kx, ky, kz;
};

inline void make_pre(const float o[3], const float d[3], Pre& p) {
  const float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
  int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (d[kz] < 0.f) std::swap(kx, ky);
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  const float inv = 1.0f / d[kz];
  p.sx = d[kx] * inv;
  p.sy = d[ky] * inv;
  p.sz = inv;
  p.o[0] = o[kx];
  p.o[1] = o[ky];
  p.o[2] = o[kz];
}

inline bool watertight(const Pre& p, const float* a, const float* b, const float* c, float& t, float& b1, float& b2) {
  const float Axk = a[p.kx] - p.o[0], Ayk = a[p.ky] - p.o[1], Azk = a[p.kz] - p.o[2];
  const float Bxk = b[p.kx] - p.o[0], Byk = b[p.ky] - p.o[1], Bzk = b[p.kz] - p.o[2];
  const float Cxk = c[p.kx] - p.o[0], Cyk = c[p.ky] - p.o[1], Czk = c[p.kz] - p.o[2];
  // The shear is one fused multiply-add per vertex coordinate (a function of the vertex and the
  // ray alone, which keeps shared edges consistent).
  const float Ax = std::fmaf(-p.sx, Azk, Axk), Ay = std::fmaf(-p.sy, Azk, Ayk);
  const float Bx = std::fmaf(-p.sx, Bzk, Bxk), By = std::fmaf(-p.sy, Bzk, Byk);
  const float Cx = std::fmaf(-p.sx, Czk, Cxk), Cy = std::fmaf(-p.sy, Czk, Cyk);
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
  const float Az = p.sz * Azk, Bz = p.sz * Bzk, Cz = p.sz * Czk;
  const float T = std::fmaf(U, Az, std::fmaf(V, Bz, W * Cz));
  t = T / det;
  const float rcp = 1.0f / det;
  b1 = V * rcp;
  b2 = W * rcp;
  return true;
}

// Object-space ray of an instance: rows 0-2 of M^-1 applied to (o, 1) and (d, 0) as nested fused
// multiply-adds, innermost term last (the product kernel's order).
inline void to_object(const float* inv, const float o[3], const float d[3], float oo[3], float od[3]) {
  for (int r = 0; r < 3; ++r) {
    const float* m = inv + 4 * r;
    oo[r] = std::fmaf(m[0], o[0], std::fmaf(m[1], o[1], std::fmaf(m[2], o[2], m[3])));
    od[r] = std::fmaf(m[0], d[0], std::fmaf(m[1], d[1], m[2] * d[2]));
  }
}

}  // namespace orc_math
