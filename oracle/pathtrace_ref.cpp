// Oracle (test infrastructure only): FP64 restatement of the wavefront path tracer's estimator.
//
// Same camera model, material lobes, light sampling, MIS weights, Russian roulette and sampler
// sequence as the product kernels, evaluated in binary64 and accumulated in binary64.
// Intersections use the oracle's watertight binary32 query (the hit set the kernels must match
// bit for bit); barycentrics are rounded to the 16-bit grid of the product's hit record.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

struct D3 {
  double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 normalize(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }
inline D3 ld(const float* p) { return {p[0], p[1], p[2]}; }
inline bool finite3(D3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
const double kPi = 3.14159265358979323846;

// lowbias32 chained over (pixel, sample, dimension): the product's sampler, bit for bit.
inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
inline double rnd(uint32_t pixel, uint32_t sample, uint32_t dim) {
  const uint32_t h = mix32(mix32(mix32(pixel) + sample) + dim * 0x9E3779B9u);
  return double(h >> 8) * (1.0 / 16777216.0);
}
inline uint32_t dim_of(uint32_t bounce, uint32_t k) { return 2u + 8u * bounce + k; }

inline D3 rgba8(uint32_t c) {
  return {double(c & 0xffu) / 255.0, double((c >> 8) & 0xffu) / 255.0, double((c >> 16) & 0xffu) / 255.0};
}

inline double ggx_d(double nh, double a2) {
  const double den = nh * nh * (a2 - 1.0) + 1.0;
  return a2 / (kPi * den * den);
}
inline double ggx_g1(double nx, double a2) { return 2.0 * nx / (nx + std::sqrt(a2 + (1.0 - a2) * nx * nx)); }

struct Lobes {
  D3 albedo, spec_col;
  double a2, metallic, ws, wt, wd;
};

void eval_bsdf(const Lobes& L, D3 n, D3 wo, D3 wi, D3& f, double& pdf) {
  const double nv = dot(n, wo), nl = dot(n, wi);
  f = {0, 0, 0};
  pdf = 0;
  if (nv <= 0 || nl <= 0) return;
  const D3 h = normalize(wo + wi);
  const double nh = std::max(dot(n, h), 0.0), vh = std::max(dot(wo, h), 1e-8);
  const double D = ggx_d(nh, L.a2);
  const double G = ggx_g1(nv, L.a2) * ggx_g1(nl, L.a2);
  const double fs = D * G / (4.0 * nv * nl);
  const double fd = (1.0 - L.metallic) / kPi;
  f = L.spec_col * (L.ws * fs) + L.albedo * (L.wd * fd);
  pdf = L.ws * (D * nh / (4.0 * vh)) + L.wd * (nl / kPi);
}

void onb(D3 n, D3& t, D3& b) {
  const double s = n.z >= 0 ? 1.0 : -1.0;
  const double a = -1.0 / (s + n.z);
  const double c = n.x * n.y * a;
  t = {1.0 + s * n.x * n.x * a, s * c, -s * n.x};
  b = {c, s + n.y * n.y * a, -n.y};
}

// Lobe weights and parameters of a material at a surface seen from wo (shading normal ns): Schlick
// Fresnel of n.v with f0 = mix(((ior-1)/(ior+1))^2, 1, metallic) (pbr.glsl:29-34) weighs the GGX
// reflection (alpha = roughness^2, floored at 1e-3), the smooth-dielectric transmission and the
// Lambert lobe; the weights are the selection probabilities and sum to 1.
Lobes make_lobes(const orc_material& mat, D3 albedo, D3 ns, D3 wo) {
  Lobes L;
  L.albedo = albedo;
  L.metallic = mat.metallic;
  const double sm = std::sqrt(std::max(double(mat.metallic), 0.0));
  L.spec_col = D3{1 + (L.albedo.x - 1) * sm, 1 + (L.albedo.y - 1) * sm, 1 + (L.albedo.z - 1) * sm};
  const double alpha = std::max(double(mat.roughness) * mat.roughness, 1e-3);
  L.a2 = alpha * alpha;
  const double f0d = (mat.ior - 1.0) / (mat.ior + 1.0);
  const double f0 = f0d * f0d + (1 - f0d * f0d) * mat.metallic;
  const double cv = std::min(std::max(dot(ns, wo), 0.0), 1.0);
  const double om = 1 - cv;
  const double Fv = f0 + (1 - f0) * (om * om * om * om * om);
  L.ws = Fv;
  L.wt = (1 - Fv) * mat.transmission;
  L.wd = (1 - Fv) * (1 - mat.transmission);
  return L;
}

// One BSDF sample from (ul, u1, u2): ul picks the lobe by weight. Transmission is a smooth
// dielectric (Snell with eta = 1/ior entering, ior leaving; total internal reflection as a
// mirror) and scales the throughput by the albedo; GGX reflection samples the half vector with
// cos(theta_h) = sqrt((1-u1)/(1+(a2-1)u1)) and reflects wo about it; Lambert samples cosine-
// weighted. Returns 0 when the path ends (direction below the geometric surface or pdf 0),
// 1 for a reflected direction, 2 for a refracted one; `weight` multiplies the throughput and pdf
// is the solid-angle density of the non-delta lobes (0 for delta lobes).
int sample_bsdf(const Lobes& L, double ior, bool entering, D3 ns, D3 ng, D3 wo, double ul, double u1, double u2,
                D3& wi, D3& weight, double& pdf, bool& delta) {
  if (ul >= L.ws && ul < L.ws + L.wt) {
    const double eta = entering ? 1.0 / ior : ior;
    const double ci = std::min(std::max(dot(wo, ns), 0.0), 1.0);
    const double s2 = eta * eta * (1 - ci * ci);
    int side = 1;
    if (s2 > 1.0) {
      wi = ns * (2 * ci) - wo;
    } else {
      wi = normalize(-wo * eta + ns * (eta * ci - std::sqrt(1 - s2)));
      side = 2;
    }
    weight = L.albedo;
    pdf = 0;
    delta = true;
    return side;
  }
  D3 t, bt;
  onb(ns, t, bt);
  if (ul < L.ws) {
    const double cos_t = std::sqrt((1 - u1) / (1 + (L.a2 - 1) * u1));
    const double sin_t = std::sqrt(std::max(0.0, 1 - cos_t * cos_t));
    const double phi = 2 * kPi * u2;
    const D3 hh = t * (sin_t * std::cos(phi)) + bt * (sin_t * std::sin(phi)) + ns * cos_t;
    wi = hh * (2 * dot(wo, hh)) - wo;
  } else {
    const double rr = std::sqrt(u1), phi = 2 * kPi * u2;
    wi = t * (rr * std::cos(phi)) + bt * (rr * std::sin(phi)) + ns * std::sqrt(std::max(0.0, 1 - u1));
  }
  D3 f;
  eval_bsdf(L, ns, wo, wi, f, pdf);
  delta = false;
  if (dot(ng, wi) <= 0 || pdf <= 0) return 0;
  weight = f * (dot(ns, wi) / pdf);
  return 1;
}

D3 sky(D3 d, double scale) {
  const double t = (d.y + 1.0) * 0.5;
  return D3{0.3 + 0.7 * t, 0.4 + 0.58 * t, 0.5 + 0.46 * t} * scale;
}

// Bilinear rgb of a w x h rgba image at continuous texel coordinates (fx, fy) (texel centres at
// integers), neighbours given by the caller's wrap or clamp.
D3 bilinear(const float* img, int w, double fx, double fy, int x0, int y0, int x1, int y1) {
  const double tx = fx - std::floor(fx), ty = fy - std::floor(fy);
  const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
  const double ws[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
  D3 out{0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    const float* t = img + 4 * (size_t(ys[k]) * w + xs[k]);
    out = out + D3{t[0], t[1], t[2]} * ws[k];
  }
  return out;
}

// Albedo texture: bilinear with wrap at (u, v), texel centres at (i + 0.5) / size.
D3 texture_lookup(const orc_pt_scene* S, uint32_t id, double u, double v) {
  const uint32_t* ti = S->tex_info + 4 * (id - 1);
  const int w = int(ti[1]), h = int(ti[2]);
  double fx = u * w - 0.5, fy = v * h - 0.5;
  fx -= std::floor(fx / w) * w;
  fy -= std::floor(fy / h) * h;
  const int x0 = std::min(int(fx), w - 1), y0 = std::min(int(fy), h - 1);
  return bilinear(S->texels + 4 * size_t(ti[0]), w, fx, fy, x0, y0, (x0 + 1) % w, (y0 + 1) % h);
}

// Cube map per the OpenGL face table, bilinear within the face, clamped to its edges.
D3 cube_lookup(const orc_pt_scene* S, D3 d) {
  const double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  int face;
  double sc, tc, ma;
  if (ax >= ay && ax >= az) {
    face = d.x >= 0 ? 0 : 1;
    sc = d.x >= 0 ? -d.z : d.z;
    tc = -d.y;
    ma = ax;
  } else if (ay >= az) {
    face = d.y >= 0 ? 2 : 3;
    sc = d.x;
    tc = d.y >= 0 ? d.z : -d.z;
    ma = ay;
  } else {
    face = d.z >= 0 ? 4 : 5;
    sc = d.z >= 0 ? d.x : -d.x;
    tc = -d.y;
    ma = az;
  }
  const int n = S->cube_size;
  const double fx = std::min(std::max(0.5 * (sc / ma + 1) * n - 0.5, 0.0), double(n - 1));
  const double fy = std::min(std::max(0.5 * (tc / ma + 1) * n - 0.5, 0.0), double(n - 1));
  const int x0 = int(fx), y0 = int(fy);
  return bilinear(S->cubemap + 4 * size_t(face) * n * n, n, fx, fy, x0, y0, std::min(x0 + 1, n - 1),
                  std::min(y0 + 1, n - 1));
}

D3 environment(const orc_pt_scene* S, D3 d) {
  if (S->cubemap) return cube_lookup(S, d) * S->env_scale;
  return sky(d, S->env_scale);
}

bool mat_inv4(const double m[16], double out[16]) {
  // Gauss-Jordan on the column-major matrix (as a[r][c] = m[c*4 + r]).
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][c + 4] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (a[piv][c] == 0.0) return false;
    for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
    const double inv = 1.0 / a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[c * 4 + r] = a[r][c + 4];
  return true;
}

constexpr uint32_t kLensDim = 0x30000000u;

// Bilinear bokeh mask lookup, texel centres at (i + 0.5) / size, clamped to the edge texels.
D3 bokeh_sample(const orc_pt_scene* S, double u, double v) {
  const int w = S->bokeh_w, h = S->bokeh_h;
  const double fx = std::min(std::max(u * w - 0.5, 0.0), double(w - 1));
  const double fy = std::min(std::max(v * h - 0.5, 0.0), double(h - 1));
  const int x0 = int(fx), y0 = int(fy), x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
  const double tx = fx - x0, ty = fy - y0;
  D3 out{0, 0, 0};
  const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
  const double ws[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
  for (int k = 0; k < 4; ++k) {
    const float* t = S->bokeh + 4 * (size_t(ys[k]) * w + xs[k]);
    out = out + D3{t[0], t[1], t[2]} * ws[k];
  }
  return out;
}

struct Ctx {
  const orc_pt_scene* s;
  orc_accel a;
  double cam[16];
  D3 eye;
  int w, h, max_depth;
};

void trace_path(const Ctx& C, uint32_t pixel, uint32_t sample, double* out, uint64_t* nonfinite) {
  const orc_pt_scene* S = C.s;
  const orc_scene* G = S->geo;
  const uint32_t x = pixel % uint32_t(C.w), y = pixel / uint32_t(C.w);
  const double nx = (double(x) + rnd(pixel, sample, 0)) / double(C.w) * 2.0 - 1.0;
  const double ny = (double(y) + rnd(pixel, sample, 1)) / double(C.h) * 2.0 - 1.0;
  const double* m = C.cam;
  const double qx = m[0] * nx + m[4] * ny + m[8] + m[12];
  const double qy = m[1] * nx + m[5] * ny + m[9] + m[13];
  const double qz = m[2] * nx + m[6] * ny + m[10] + m[14];
  const double qw = m[3] * nx + m[7] * ny + m[11] + m[15];
  D3 ro = C.eye;
  D3 rd = qw != 0 ? normalize(D3{qx / qw, qy / qw, qz / qw} - C.eye) : normalize(D3{qx, qy, qz});
  D3 thr{1, 1, 1}, sum{0, 0, 0};
  if (S->dof) {
    // Thin lens: offset uniform in the disk of radius 0.49 (redrawn where the mask is black, up
    // to 8 times), focal point fixed at `focus` along the pinhole ray, mask colour as the tint.
    double ox = 0, oy = 0;
    bool ok = false;
    D3 tint{1, 1, 1};
    for (uint32_t k = 0; k < 8 && !ok; ++k) {
      const double a = rnd(pixel, sample, kLensDim + 2 * k) * 6.2831853;
      const double rr = std::sqrt(rnd(pixel, sample, kLensDim + 2 * k + 1)) * 0.49;
      ox = rr * std::sin(a);
      oy = rr * std::cos(a);
      if (!S->bokeh) {
        ok = true;
      } else {
        tint = bokeh_sample(S, 0.5 + ox, 0.5 + oy);
        ok = std::max(std::max(tint.x, tint.y), tint.z) >= 1e-5;
      }
    }
    if (!ok) tint = D3{0, 0, 0};
    const D3 center = ro + rd * double(S->focus);
    const D3 local = normalize(D3{ox * S->lens[0] * 0.5, oy * S->lens[1] * 0.5, double(S->focus)});
    D3 u = cross(rd, D3{1, 0, 0});
    if (dot(u, u) == 0.0) u = cross(rd, D3{0, 1, 0});
    u = normalize(u);
    const D3 v = cross(rd, u);
    const D3 nd = normalize(u * local.x + v * local.y + rd * local.z);
    ro = center - nd * double(S->focus);
    rd = nd;
    thr = tint;
  }
  bool delta = true;
  double last_pdf = 0;
  for (int bounce = 0; bounce < C.max_depth; ++bounce) {
    orc_ray r;
    for (int k = 0; k < 3; ++k) {
      r.o[k] = float((&ro.x)[k]);
      r.d[k] = float((&rd.x)[k]);
    }
    r.tmin = 0.f;
    r.tmax = INFINITY;
    orc_hit hit;
    orc_accel_trace(C.a, &r, 1, 0, &hit, 1);
    float sdf_n[3] = {0, 0, 0};
    if (S->n_sdf)
      orc_sdf_trace(static_cast<const orc_sdf_node*>(S->sdf_tape), static_cast<const orc_sdf_drawable*>(S->sdf_drawables),
                    S->n_sdf, &r, 1, 0, S->sdf_steps, S->sdf_eps, &hit, sdf_n, 1);
    bool found = hit.inst != 0xFFFFFFFFu;
    const double t_hit = found ? double(hit.t) : INFINITY;
    // The product works on the binary32 ray it stored; continue from the same values.
    ro = ld(r.o);
    rd = ld(r.d);
    // Non-finite geometric normal: the hit is a miss (pt_trace.comp:186-189); non-finite shading
    // normal: the geometric one. The product's rules, in binary64.
    D3 ng{0, 0, 1}, ns = ng;
    uint32_t mi = 0;
    double hu = 0, hv = 0;
    bool has_uv = false;
    if (found) {
    const uint32_t inst = hit.inst, prim = hit.prim;
    if (inst & 0x80000000u) {
      ng = normalize(D3{sdf_n[0], sdf_n[1], sdf_n[2]});
      ns = ng;
      mi = prim;
    } else {
      const uint32_t g = G->inst_geom[inst];
      const uint32_t* tri = G->idx + G->geo_idx_off[g] + 3 * prim;
      const float* P = G->pos + 3 * size_t(G->geo_vtx_off[g]);
      const D3 p0 = ld(P + 3 * tri[0]), p1 = ld(P + 3 * tri[1]), p2 = ld(P + 3 * tri[2]);
      const double b1 = std::round(std::fmin(std::fmax(double(hit.b1), 0.0), 1.0) * 65535.0) / 65535.0;
      const double b2 = std::round(std::fmin(std::fmax(double(hit.b2), 0.0), 1.0) * 65535.0) / 65535.0;
      const double b0 = 1.0 - b1 - b2;
      if (S->uv) {
        const float* T = S->uv + 2 * size_t(G->geo_vtx_off[g]);
        hu = T[2 * tri[0]] * b0 + T[2 * tri[1]] * b1 + T[2 * tri[2]] * b2;
        hv = T[2 * tri[0] + 1] * b0 + T[2 * tri[1] + 1] * b1 + T[2 * tri[2] + 1] * b2;
        has_uv = true;
      }
      const D3 ngo = cross(p1 - p0, p2 - p0);
      D3 nso = ngo;
      if (S->geo_has_nrm[g]) {
        const float* N = S->nrm + 3 * size_t(G->geo_vtx_off[g]);
        nso = ld(N + 3 * tri[0]) * b0 + ld(N + 3 * tri[1]) * b1 + ld(N + 3 * tri[2]) * b2;
      }
      const float* iv = G->inst_inv + 12 * inst;
      auto xf = [&](D3 n) {
        return normalize(D3{iv[0] * n.x + iv[4] * n.y + iv[8] * n.z, iv[1] * n.x + iv[5] * n.y + iv[9] * n.z,
                            iv[2] * n.x + iv[6] * n.y + iv[10] * n.z});
      };
      ng = xf(ngo);
      ns = xf(nso);
      mi = S->inst_material[inst];
    }
    if (!finite3(ng)) {
      found = false;
      ++*nonfinite;
    } else if (!finite3(ns)) {
      ns = ng;
      ++*nonfinite;
    }
    }
    int light_hit = -1;
    double t_light = t_hit, cos_light = 0;
    for (uint32_t l = 0; l < S->n_lights; ++l) {
      const float* q = S->lights + 12 * l;
      const D3 c = ld(q), u = ld(q + 3), v = ld(q + 6);
      const D3 nl = cross(u, v);
      const double area = std::sqrt(dot(nl, nl));
      const D3 n1 = nl * (1.0 / area);
      const double den = dot(n1, rd);
      if (den >= 0) continue;
      const double t = dot(n1, c - ro) / den;
      if (!(t > 1e-6 && t < t_light)) continue;
      const D3 pq = ro + rd * t - c;
      const double a = dot(pq, u) / dot(u, u), b = dot(pq, v) / dot(v, v);
      if (a < 0 || a > 1 || b < 0 || b > 1) continue;
      light_hit = int(l);
      t_light = t;
      cos_light = -den;
    }
    if (light_hit >= 0) {
      const float* q = S->lights + 12 * light_hit;
      double w = 1.0;
      if (!delta) {
        const D3 nl = cross(ld(q + 3), ld(q + 6));
        const double area = std::sqrt(dot(nl, nl));
        const double pdf_l = t_light * t_light / (cos_light * area * double(S->n_lights));
        w = last_pdf * last_pdf / (last_pdf * last_pdf + pdf_l * pdf_l);
      }
      const D3 c = thr * ld(q + 9) * w;
      if (finite3(c)) sum = sum + c; else ++*nonfinite;
      break;
    }
    if (!found) {
      const D3 c = thr * environment(S, rd);
      if (finite3(c)) sum = sum + c; else ++*nonfinite;
      break;
    }
    const D3 p = ro + rd * t_hit;
    const D3 wo = -rd;
    const bool entering = dot(rd, ng) < 0;
    if (!entering) {
      ng = -ng;
      ns = -ns;
    }
    if (dot(ns, ng) <= 0) ns = ng;
    const double eps = 1e-4 * (1.0 + std::max(std::max(std::fabs(p.x), std::fabs(p.y)), std::fabs(p.z)));
    if (mi >= S->n_materials) mi = 0;
    const orc_material& mat = S->materials[mi];
    D3 albedo = rgba8(mat.albedo_rgba8);
    if (has_uv && mat.texture >= 1 && mat.texture <= S->n_textures) {
      const D3 t = texture_lookup(S, mat.texture, hu, 1.0 - hv);
      albedo = D3{std::pow(std::max(t.x, 0.0), 2.2), std::pow(std::max(t.y, 0.0), 2.2), std::pow(std::max(t.z, 0.0), 2.2)};
    }
    const Lobes L = make_lobes(mat, albedo, ns, wo);

    if (S->n_lights > 0 && (L.ws + L.wd) > 0) {
      uint32_t l = uint32_t(rnd(pixel, sample, dim_of(bounce, 6)) * double(S->n_lights));
      if (l >= S->n_lights) l = S->n_lights - 1;
      const float* q = S->lights + 12 * l;
      const D3 u = ld(q + 3), v = ld(q + 6);
      const double r1 = rnd(pixel, sample, dim_of(bounce, 4)), r2 = rnd(pixel, sample, dim_of(bounce, 5));
      const D3 xl = ld(q) + u * r1 + v * r2;
      const D3 nl = cross(u, v);
      const double area = std::sqrt(dot(nl, nl));
      D3 wi = xl - p;
      const double d2 = dot(wi, wi);
      wi = wi * (1.0 / std::sqrt(d2));
      const double cl = -dot(nl, wi) / area;
      if (cl > 0 && dot(ng, wi) > 0) {
        D3 f;
        double pdf_b;
        eval_bsdf(L, ns, wo, wi, f, pdf_b);
        const double pdf_l = d2 / (cl * area * double(S->n_lights));
        const double w = pdf_l * pdf_l / (pdf_l * pdf_l + pdf_b * pdf_b);
        const D3 c = thr * f * ld(q + 9) * (dot(ns, wi) * w / pdf_l);
        if (!finite3(c)) {
          ++*nonfinite;
        } else if (c.x > 0 || c.y > 0 || c.z > 0) {
          const D3 so = p + ng * eps;
          orc_ray sr;
          for (int k = 0; k < 3; ++k) {
            sr.o[k] = float((&so.x)[k]);
            sr.d[k] = float((&xl.x)[k] - (&so.x)[k]);
          }
          sr.tmin = 0.f;
          sr.tmax = 1.0f - 1e-3f;
          orc_hit sh;
          orc_accel_trace(C.a, &sr, 1, 1, &sh, 1);
          if (S->n_sdf && !std::isfinite(sh.t))
            orc_sdf_trace(static_cast<const orc_sdf_node*>(S->sdf_tape),
                          static_cast<const orc_sdf_drawable*>(S->sdf_drawables), S->n_sdf, &sr, 1, 1, S->sdf_steps,
                          S->sdf_eps, &sh, nullptr, 1);
          if (!std::isfinite(sh.t)) sum = sum + c;
        }
      }
    }

    const double ul = rnd(pixel, sample, dim_of(bounce, 0));
    const double u1 = rnd(pixel, sample, dim_of(bounce, 1)), u2 = rnd(pixel, sample, dim_of(bounce, 2));
    D3 wi, weight;
    double pdf = 0;
    bool was_delta = false;
    const int side = sample_bsdf(L, mat.ior, entering, ns, ng, wo, ul, u1, u2, wi, weight, pdf, was_delta);
    bool ok = side != 0;
    const D3 origin = side == 2 ? p - ng * eps : p + ng * eps;
    if (ok) {
      thr = thr * weight;
      delta = was_delta;
      if (!was_delta) last_pdf = pdf;
    }
    if (ok && !finite3(thr)) {
      ok = false;
      ++*nonfinite;
    }
    if (ok && bounce >= 3) {
      const double q = std::min(0.95, std::max(std::max(thr.x, thr.y), thr.z));
      if (rnd(pixel, sample, dim_of(bounce, 3)) >= q)
        ok = false;
      else
        thr = thr * (1.0 / q);
    }
    if (!(ok && bounce + 1 < C.max_depth && (thr.x > 0 || thr.y > 0 || thr.z > 0))) break;
    ro = origin;
    rd = wi;
  }
  out[0] += sum.x;
  out[1] += sum.y;
  out[2] += sum.z;
}

}  // namespace

extern "C" int orc_render(const orc_pt_scene* s, orc_accel a, const float view[16], const float proj[16], int w,
                          int h, int spp, int sample_base, int max_depth, double* out_rgb, int threads) {
  Ctx C;
  C.s = s;
  C.a = a;
  C.w = w;
  C.h = h;
  C.max_depth = max_depth;
  double v[16], p[16], vi[16], pi[16];
  for (int i = 0; i < 16; ++i) {
    v[i] = view[i];
    p[i] = proj[i];
  }
  if (!mat_inv4(v, vi) || !mat_inv4(p, pi)) return -1;
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double acc = 0;
      for (int k = 0; k < 4; ++k) acc += vi[k * 4 + r] * pi[c * 4 + k];
      C.cam[c * 4 + r] = acc;
    }
  C.eye = D3{vi[12], vi[13], vi[14]};
  const size_t npix = size_t(w) * h;
  std::fill(out_rgb, out_rgb + 3 * npix, 0.0);
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<uint64_t> bad(size_t(threads), 0);
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (size_t px = size_t(t); px < npix; px += size_t(threads))
        for (int sidx = 0; sidx < spp; ++sidx)
          trace_path(C, uint32_t(px), uint32_t(sample_base + sidx), out_rgb + 3 * px, &bad[size_t(t)]);
    });
  }
  for (auto& th : pool) th.join();
  uint64_t total = 0;
  for (uint64_t b : bad) total += b;
  return int(std::min<uint64_t>(total, 0x7fffffff));
}

// The material model on its own, for tests that pin it against the mathematics (sampling
// densities, reciprocity, albedo, Snell's law). n is the shading normal on wo's side; albedo comes
// from the material record. Batched over directions (eval) or uniform triples (sample).
extern "C" int orc_bsdf_eval(const orc_material* m, const double n[3], const double wo[3], const double* wi, size_t count,
                             double* f, double* pdf) {
  const D3 N{n[0], n[1], n[2]}, WO{wo[0], wo[1], wo[2]};
  const Lobes L = make_lobes(*m, rgba8(m->albedo_rgba8), N, WO);
  for (size_t k = 0; k < count; ++k) {
    D3 F;
    eval_bsdf(L, N, WO, D3{wi[3 * k], wi[3 * k + 1], wi[3 * k + 2]}, F, pdf[k]);
    f[3 * k] = F.x;
    f[3 * k + 1] = F.y;
    f[3 * k + 2] = F.z;
  }
  return 0;
}

extern "C" int orc_bsdf_sample(const orc_material* m, const double n[3], const double wo[3], int entering,
                               const double* u, size_t count, double* wi, double* weight, double* pdf, int32_t* side,
                               int32_t* is_delta) {
  const D3 N{n[0], n[1], n[2]}, WO{wo[0], wo[1], wo[2]};
  const Lobes L = make_lobes(*m, rgba8(m->albedo_rgba8), N, WO);
  for (size_t k = 0; k < count; ++k) {
    D3 WI{0, 0, 0}, W{0, 0, 0};
    bool delta = false;
    pdf[k] = 0;
    side[k] = sample_bsdf(L, m->ior, entering != 0, N, N, WO, u[3 * k], u[3 * k + 1], u[3 * k + 2], WI, W, pdf[k],
                          delta);
    wi[3 * k] = WI.x;
    wi[3 * k + 1] = WI.y;
    wi[3 * k + 2] = WI.z;
    weight[3 * k] = W.x;
    weight[3 * k + 1] = W.y;
    weight[3 * k + 2] = W.z;
    is_delta[k] = delta ? 1 : 0;
  }
  return 0;
}
