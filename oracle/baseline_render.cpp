// Oracle (measurement infrastructure only): the CPU baseline of the reference's plain renderer.
//
// One frame of src/pathtracer/sequential_pathtracer.cpp:45-64 (generate; per bounce trace, color,
// filter; a final trace and color that writes every live path) restated on the host, pass by
// pass over arrays of paths, each pass split into jobs that a clone of myrt::thread_pool runs
// (the same pool as orc_reference_time). The passes follow:
//   generate  src/pt/pt_generate.comp:50-84 (pinhole: disk jitter of radius 0.7071, no divide by w)
//   trace     src/pt/pt_trace.comp:134-225 (the two hard-coded sphere lights, a random point on one
//             of them, the reference's stackless traversal from orc_reference_trace for the
//             closest hit and the shadow ray, interpolated normals, pt_trace.comp:138-154)
//   color     src/pt/pt_color.comp:113-202 with pbr_resample / pbr_eval (src/pt/pbr.glsl:29-232),
//             src/pt/ggx.glsl:1-43, src/pt/material_new.glsl:17-48 and src/pt/transform.glsl:3-15
//   filter    src/pt/pt_ray_filter.comp:49-63 (survivors are compacted)
// and the random numbers of src/pt/random.glsl:1-84 over a 1024-entry random texture
// (sequential_pathtracer.cpp:276-307). It exists to time the reference's work on the GPU box's
// host cores; the product's estimator and its parity live in pathtrace_ref.cpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <future>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

struct V3 {
  float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline bool bad(V3 a) { return !std::isfinite(a.x) || !std::isfinite(a.y) || !std::isfinite(a.z); }
inline V3 mix(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
inline V3 clamp3(V3 a, float lo, float hi) {
  return {std::fmin(std::fmax(a.x, lo), hi), std::fmin(std::fmax(a.y, lo), hi), std::fmin(std::fmax(a.z, lo), hi)};
}
// GLSL faceforward(N, I, Nref): N if dot(Nref, I) < 0, else -N.
inline V3 faceforward(V3 n, V3 i, V3 nref) { return dot(nref, i) < 0.f ? n : -n; }
inline V3 reflect(V3 i, V3 n) { return i - n * (2.0f * dot(n, i)); }
inline V3 refract(V3 i, V3 n, float eta) {
  const float d = dot(n, i);
  const float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.f) return {0, 0, 0};
  return i * eta - n * (eta * d + std::sqrt(k));
}
inline uint32_t pack_unorm8(V3 c) {
  auto q = [](float v) { return uint32_t(std::lround(std::fmin(std::fmax(v, 0.f), 1.f) * 255.f)); };
  return q(c.x) | (q(c.y) << 8) | (q(c.z) << 16);
}
inline V3 unpack_unorm8(uint32_t p) {
  return {float(p & 255u) / 255.f, float((p >> 8) & 255u) / 255.f, float((p >> 16) & 255u) / 255.f};
}

constexpr float kPi = 3.1415926535897f;

// ---- random.glsl:1-84 ------------------------------------------------------------------------
struct Rng {
  uint32_t state = 0;
  int prime = 0;
};
inline uint32_t wang_hash(uint32_t a) {
  a = (a ^ 61u) ^ (a >> 16);
  a = a + (a << 3);
  a = a ^ (a >> 4);
  a = a * 0x27d4eb2du;
  a = a ^ (a >> 15);
  return a;
}
inline float random_value(Rng& r, int32_t seed) {
  r.state += wang_hash(uint32_t(seed));
  r.state ^= r.state << 13;
  r.state ^= r.state >> 17;
  r.state ^= r.state << 5;
  return std::fmin(std::fmax(float(r.state) * (1.f / float(0xFFFFFFFFu)), 0.f), 1.f);
}
inline int32_t wrap_int(double v) { return int32_t(uint32_t(int64_t(v))); }  // GLSL int() of a large float, wrapped
inline int32_t random_int(Rng& r, int32_t seed) { return wrap_int(double(random_value(r, seed)) * double(0x7fffffff)); }
inline void random_init(Rng& r, const std::vector<float>& tex, int px, int py, int32_t start) {
  r.state = uint32_t(start);
  r.prime = 0;
  const uint32_t base = uint32_t(random_int(r, px ^ 0xba77fa)) * 69019u + uint32_t(random_int(r, py ^ 0xcca6df)) * 96013u +
                        uint32_t(random_int(r, start));
  const int32_t sb = int32_t(base), n = int32_t(tex.size());
  const int32_t slot = ((sb % n) + n) % n;
  random_value(r, int32_t(uint32_t(sb) + uint32_t(wrap_int(double(0xffaf86) * tex[size_t(slot)]))));
}
inline float random_next(Rng& r) {
  r.prime = (r.prime + 1) % 6;
  return random_value(r, r.prime);
}

// ---- pbr.glsl / ggx.glsl / transform.glsl ------------------------------------------------------
struct Mat {
  V3 color;
  float ior, roughness, metallic, transmission;
};
enum { kReflect = 0, kTransmit = 1, kDiffuse = 2 };
struct PbrState {
  int state = -1;
  float resample_factor = 1.0f;
  bool after_sample = false;
};

inline V3 to_world(V3 v, V3 n) {
  V3 perp{1, 0, 0};
  V3 u = normalize(cross(n, perp));
  for (int i = 1; bad(u) && i < 3; ++i) {
    perp = i == 1 ? V3{0, 1, 0} : V3{0, 0, 1};
    u = normalize(cross(n, perp));
  }
  return normalize(u * v.x + cross(n, u) * v.y + n * v.z);
}
inline float fresnel(float cos_theta, float ior_in, float ior_out, float metallic) {
  float f0 = std::fabs((ior_in - ior_out) / (ior_in + ior_out));
  f0 = f0 * f0 * (1.0f - metallic) + metallic;
  return f0 + (1 - f0) * std::pow(1 - cos_theta, 5.0f);
}
inline float ggx_distribution(V3 n, V3 h, float alpha) {
  const float ndotm = dot(n, h), a2 = alpha * alpha;
  const float den = ndotm * ndotm * (a2 - 1) + 1;
  return a2 / (kPi * den * den);
}
inline V3 ggx_sample(float rx, float ry, float roughness) {
  const float phi = ry * 2.f * kPi;
  const float rough2 = roughness * roughness;
  const float theta = std::atan(std::sqrt(rough2 * rx / (1 - rx)));
  return {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
}
inline V3 nreflect(V3 towards_view, V3 normal, V3 m) {
  V3 ray = reflect(-towards_view, m);
  const float d = dot(ray, normal);
  if (d < 0) ray = ray - normal * (2 * d);
  return normalize(ray);
}
inline float smith(V3 n, V3 v, V3 l, float k) {
  const float ndv = std::fabs(dot(n, v)), ndl = std::fabs(dot(n, l));
  return ndl * (1 / (ndv * (1.0f - k) + k)) * (1 / (ndl * (1.0f - k) + k));
}
inline void ggx_eval(const Mat& m, const PbrState& ps, V3 v, V3 l, V3 n, float ior_f, float ior_b, V3& refl, float& pdf) {
  const float roughness = m.roughness * m.roughness;
  const V3 h = normalize(l + v);
  const float D = (ps.after_sample && roughness < 0.01f) ? 1 : ggx_distribution(n, h, roughness);
  const float G = smith(n, v, l, roughness);
  const float F = fresnel(dot(v, h), ior_f, ior_b, m.metallic);
  const float den = 4 * std::fabs(dot(v, n)) * std::fabs(dot(l, n));
  const float cos_theta = dot(h, n);
  const V3 col = mix(V3{1, 1, 1}, m.color, std::sqrt(m.metallic));
  float mden = G / den;
  mden = std::isnan(mden) ? 1 : mden;
  pdf = F * D * cos_theta * mden;
  refl = col * pdf;
}
inline void transmit_eval(const Mat& m, const PbrState& ps, V3 v, V3 l, V3 n, float ior_f, float ior_b, V3& refl,
                          float& pdf) {
  const float roughness = m.roughness * m.roughness;
  V3 h = normalize(-(v * ior_f + l * ior_b));
  const float s = dot(h, n);
  h = h * float((s > 0) - (s < 0));
  const float D = (ps.after_sample && roughness < 0.01f) ? 1 : ggx_distribution(n, h, roughness);
  const float G = smith(n, v, l, roughness);
  const float idoth = dot(v, h), odoth = dot(l, h);
  const float prefac = (std::fabs(idoth) * std::fabs(odoth)) / (std::fabs(dot(v, n)) * std::fabs(dot(l, n)));
  const float den = ior_f * idoth + ior_b * odoth;
  const float F = fresnel(dot(v, h), ior_f, ior_b, m.metallic);
  const float ior2_by_den2 = ior_b * ior_b / (den * den);
  const float ft = prefac * (1 - F) * G * D * ior2_by_den2;
  pdf = dot(h, n) * ft;
  refl = m.color * pdf;
}
// material_new.glsl:17-48 open / evaluate / bounce over the pbr.glsl functions.
struct Shader {
  Mat mat;
  PbrState ps;
  bool incoming = false;
  float ior1 = 1, ior2 = 1;

  void sides(V3 normal, V3 towards_viewer) {
    incoming = dot(normal, -towards_viewer) < 0;
    ior1 = incoming ? 1.0f : mat.ior;
    ior2 = incoming ? mat.ior : 1.0f;
  }
  V3 bounce(Rng& rng_for_state, float rx, float ry, V3 towards_viewer, V3 normal) {
    sides(normal, towards_viewer);
    const V3 n = incoming ? normal : -normal;
    // pbr.glsl:164-201: the lobe choice uses its own two values, drawn from the shared state.
    const float frand = random_value(rng_for_state, int32_t(uint32_t(int32_t(19022 * rx)) * 2909293u));
    const float trand = random_value(rng_for_state, int32_t(uint32_t(int32_t(23452 * ry)) * 8994355u));
    V3 m = to_world(ggx_sample(rx, ry, mat.roughness * mat.roughness), n);
    const float F = fresnel(dot(towards_viewer, m), ior1, ior2, mat.metallic);
    ps.after_sample = true;
    const V3 refracted = refract(-towards_viewer, m, ior1 / ior2);
    const bool tir = refracted.x == 0.f && refracted.y == 0.f && refracted.z == 0.f;
    if ((mat.transmission > 0 && tir) || F >= frand) {
      ps.state = kReflect;
      ps.resample_factor = F;
      return nreflect(towards_viewer, n, m);
    }
    if (mat.transmission >= trand) {
      ps.resample_factor = (1 - F) * mat.transmission;
      ps.state = kTransmit;
      return refracted;
    }
    ps.resample_factor = (1 - F) * (1 - mat.transmission);
    ps.state = kDiffuse;
    m = to_world(ggx_sample(rx, ry, 1.0f), n);
    return nreflect(towards_viewer, n, m);
  }
  void evaluate(V3 normal, V3 towards_light, V3 towards_viewer, V3& refl, float& pdf) {
    sides(normal, towards_viewer);
    const V3 n = incoming ? normal : -normal;
    V3 rt, rg, rd;
    float pt, pg, pd;
    transmit_eval(mat, ps, towards_viewer, towards_light, n, ior1, ior2, rt, pt);
    ggx_eval(mat, ps, towards_viewer, towards_light, n, ior1, ior2, rg, pg);
    Mat diffuse = mat;
    diffuse.roughness = 1.0f;
    diffuse.metallic = 1.0f;
    ggx_eval(diffuse, ps, towards_viewer, towards_light, n, ior1, ior2, rd, pd);
    refl = {0, 0, 0};
    pdf = 1;
    switch (ps.state) {
      case kTransmit:
        refl = rt;
        pdf = pt;
        break;
      case kDiffuse:
        refl = rd;
        pdf = pd;
        break;
      case kReflect:
        refl = rg;
        pdf = pg;
        break;
    }
    ps.after_sample = false;
    pdf *= ps.resample_factor;
    refl = refl * ps.resample_factor;
  }
};

// ---- the frame -------------------------------------------------------------------------------
struct Light {
  V3 position;
  float radius;
  V3 color;
};
// pt_trace.comp:141-146
const Light kLights[2] = {{{8, 10, 8}, 1.5f, {6, 10.8f, 18}}, {{-8, 6, 10}, 1.5f, {24, 14.4f, 8}}};

float intersect_sphere(V3 r0, V3 rd, V3 s0, float sr) {  // intersect.glsl:30-39
  const float a = dot(rd, rd);
  const V3 d = r0 - s0;
  const float b = 2.0f * dot(rd, d);
  const float c = dot(d, d) - sr * sr;
  if (b * b - 4.0f * a * c < 0.0f) return -1.0f;
  return (-b - std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
}

struct Path {  // generate_output_t (pt_color.comp:12-24)
  V3 origin;
  float random_value;
  V3 direction;
  uint32_t pixel;
  V3 reflectance;
  float pdf;
  V3 color;
  uint32_t id;
};
struct TraceOut {  // trace_output_t (pt_color.comp:26-39)
  V3 point;
  int intersects;
  V3 normal;
  uint32_t material;
  uint32_t light_color;
  V3 light_direction;
  float light_strength;
};

class ThreadPool {  // the FIFO pool of baseline.cpp, restated here so this file stands alone
 public:
  explicit ThreadPool(unsigned n) {
    for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~ThreadPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  std::future<void> run_async(std::function<void()> f) {
    auto task = std::make_shared<std::packaged_task<void()>>(std::move(f));
    std::future<void> fut = task->get_future();
    {
      std::lock_guard<std::mutex> lk(m_);
      jobs_.emplace_back([task] { (*task)(); });
    }
    cv_.notify_one();
    return fut;
  }

 private:
  void loop() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (stop_) return;
      auto job = std::move(jobs_.front());
      jobs_.pop_front();
      lk.unlock();
      job();
      lk.lock();
    }
  }
  std::vector<std::thread> workers_;
  std::deque<std::function<void()>> jobs_;
  std::mutex m_;
  std::condition_variable cv_;
  bool stop_ = false;
};

struct Frame {
  const orc_pt_scene* s;
  orc_accel a;
  int w, h, rows;
  float inv_view[16], inv_proj[16];  // column-major
  std::vector<float> random_texture;
  std::vector<float> fwd;  // per instance: rows 0-2 of M
  std::mt19937 rng{1};
  std::vector<double> sums;  // per pixel rgb
  uint64_t rays = 0;

  int32_t draw() { return int32_t(std::uniform_int_distribution<int32_t>(0, 0x7fffffff)(rng)); }
};

bool invert4(const float* m, float* out) {  // column-major 4x4 inverse (cofactors, double)
  double a[16], inv[16];
  for (int i = 0; i < 16; ++i) a[i] = m[i];
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  if (det == 0.0 || !std::isfinite(det)) return false;
  for (int i = 0; i < 16; ++i) out[i] = float(inv[i] / det);
  return true;
}

// Forward rows of an instance from the inverse rows: the inverse of the 3x4 affine map.
void forward_rows(const float* inv, float* out) {
  float m[16] = {inv[0], inv[4], inv[8], 0, inv[1], inv[5], inv[9], 0, inv[2], inv[6], inv[10], 0, inv[3], inv[7], inv[11], 1};
  float f[16];
  invert4(m, f);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) out[4 * r + c] = f[4 * c + r];
}

void accumulate(Frame& F, const Path& p, V3 c) {
  if (bad(c)) return;  // pt_color.comp:84-98
  double* s = &F.sums[3 * size_t(p.pixel)];
  s[0] += c.x;
  s[1] += c.y;
  s[2] += c.z;
}

// Runs fn over [0, n) in jobs of `chunk` on the pool and waits.
template <typename Fn>
void jobs(ThreadPool& pool, size_t n, size_t chunk, Fn fn) {
  std::vector<std::future<void>> futs;
  for (size_t b = 0; b < n; b += chunk) {
    const size_t e = std::min(n, b + chunk);
    futs.push_back(pool.run_async([=, &fn] { fn(b, e); }));
  }
  for (auto& f : futs) f.get();
}

// One path per pixel (1 sample), so no two jobs write one pixel.
void render_frame(Frame& F, ThreadPool& pool, int max_depth, size_t chunk) {
  const orc_pt_scene* s = F.s;
  const orc_scene* g = s->geo;
  std::vector<Path> paths(size_t(F.w) * F.rows);
  // generate (pt_generate.comp:50-84)
  const int32_t seed = F.draw();
  jobs(pool, paths.size(), chunk, [&](size_t b, size_t e) {
    for (size_t id = b; id < e; ++id) {
      const int px = int(id % size_t(F.w)), py = int(id / size_t(F.w));
      Rng r;
      random_init(r, F.random_texture, px, py, seed);
      const float ang = random_next(r) * 2 * kPi, rad = random_next(r);
      const float ox = std::sqrt(rad) * 0.7071067812f * std::sin(ang), oy = std::sqrt(rad) * 0.7071067812f * std::cos(ang);
      const float ndc[4] = {((ox + float(px)) / float(F.w)) * 2 - 1, ((oy + float(py)) / float(F.h)) * 2 - 1, 1, 1};
      float t[4], d[4];
      for (int i = 0; i < 4; ++i) t[i] = F.inv_proj[i] * ndc[0] + F.inv_proj[4 + i] * ndc[1] + F.inv_proj[8 + i] * ndc[2] + F.inv_proj[12 + i] * ndc[3];
      for (int i = 0; i < 4; ++i) d[i] = F.inv_view[i] * t[0] + F.inv_view[4 + i] * t[1] + F.inv_view[8 + i] * t[2] + F.inv_view[12 + i] * t[3];
      Path& p = paths[id];
      p.direction = normalize(V3{d[0], d[1], d[2]});
      p.origin = {F.inv_view[12], F.inv_view[13], F.inv_view[14]};
      p.random_value = random_next(r);
      p.reflectance = {1, 1, 1};
      p.color = {0, 0, 0};
      p.pdf = 1.0f;
      p.pixel = uint32_t(py * F.w + px);
      p.id = uint32_t(id);
    }
  });
  std::vector<TraceOut> out;
  std::vector<orc_ray> rays;
  std::vector<orc_hit> hits;
  std::vector<V3> light_point;
  std::vector<uint32_t> light_index;
  for (int bounce = 0; bounce <= max_depth; ++bounce) {
    const bool last = bounce == max_depth;
    const size_t n = paths.size();
    if (!n) break;
    out.assign(n, TraceOut{});
    rays.resize(n);
    hits.resize(n);
    light_point.resize(n);
    light_index.resize(n);
    // trace (pt_trace.comp:134-225): the light sample, the closest hit and the lights
    const int32_t trace_seed = F.draw();
    jobs(pool, n, chunk, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        const Path& p = paths[i];
        Rng r;
        random_init(r, F.random_texture, int(p.pixel % uint32_t(F.w)), int(p.pixel / uint32_t(F.w)), trace_seed);
        const int li = std::min(std::max(int(2 * random_next(r)), 0), 1);
        const float a = random_next(r), bb = random_next(r), c = random_next(r);
        const V3 dir = normalize(V3{a, bb, c});
        light_index[i] = uint32_t(li);
        light_point[i] = kLights[li].position + dir * (kLights[li].radius * std::sqrt(random_next(r)));
        rays[i] = orc_ray{{p.origin.x, p.origin.y, p.origin.z}, 0.f, {p.direction.x, p.direction.y, p.direction.z}, INFINITY};
      }
      orc_reference_trace(F.a, rays.data() + b, e - b, 0, hits.data() + b);
    });
    // hit records, light hits and shadow rays toward the sampled light point
    std::vector<uint8_t> shadow_wanted(n, 0);
    std::vector<orc_ray> shadow(n);
    std::vector<orc_hit> occl(n);
    jobs(pool, n, chunk, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        const Path& p = paths[i];
        const orc_hit& h = hits[i];
        bool hit = h.inst != 0xFFFFFFFFu;
        float t = hit ? h.t : INFINITY;
        int light_hit = -1;
        for (int l = 0; l < 2; ++l) {
          const float tl = intersect_sphere(p.origin, p.direction, kLights[l].position, kLights[l].radius);
          if (tl > 0 && (!hit || tl < t)) {
            hit = true;
            t = tl;
            light_hit = l;
          }
        }
        TraceOut& o = out[i];
        if (!hit) continue;
        o.intersects = 1;
        if (light_hit < 0) {
          const uint32_t inst = h.inst, geo = g->inst_geom[inst];
          const uint32_t* tri = g->idx + g->geo_idx_off[geo] + 3 * size_t(h.prim);
          const size_t vb = g->geo_vtx_off[geo];
          const float* P[3] = {g->pos + 3 * (vb + tri[0]), g->pos + 3 * (vb + tri[1]), g->pos + 3 * (vb + tri[2])};
          const float* N[3] = {s->nrm + 3 * (vb + tri[0]), s->nrm + 3 * (vb + tri[1]), s->nrm + 3 * (vb + tri[2])};
          const float w0 = 1.0f - h.b1 - h.b2;
          V3 lp, ln;
          lp.x = h.b1 * P[1][0] + h.b2 * P[2][0] + w0 * P[0][0];
          lp.y = h.b1 * P[1][1] + h.b2 * P[2][1] + w0 * P[0][1];
          lp.z = h.b1 * P[1][2] + h.b2 * P[2][2] + w0 * P[0][2];
          ln.x = h.b1 * N[1][0] + h.b2 * N[2][0] + w0 * N[0][0];
          ln.y = h.b1 * N[1][1] + h.b2 * N[2][1] + w0 * N[0][1];
          ln.z = h.b1 * N[1][2] + h.b2 * N[2][2] + w0 * N[0][2];
          const float* M = &F.fwd[12 * size_t(inst)];
          const float* I = g->inst_inv + 12 * size_t(inst);
          o.point = {M[0] * lp.x + M[1] * lp.y + M[2] * lp.z + M[3], M[4] * lp.x + M[5] * lp.y + M[6] * lp.z + M[7],
                     M[8] * lp.x + M[9] * lp.y + M[10] * lp.z + M[11]};
          o.normal = normalize(V3{I[0] * ln.x + I[4] * ln.y + I[8] * ln.z, I[1] * ln.x + I[5] * ln.y + I[9] * ln.z,
                                  I[2] * ln.x + I[6] * ln.y + I[10] * ln.z});
          o.material = s->inst_material[inst];
          if (bad(o.normal)) o.intersects = 0;
        }
        const V3 to_light = light_point[i] - o.point;
        const V3 dir_light = normalize(to_light);
        const float dist_light = length(to_light);
        const V3 lnorm = faceforward(o.normal, p.direction, o.normal);
        if (light_hit >= 0) {
          o.intersects = 2;
          const V3 col = kLights[light_hit].color;
          o.light_strength = length(col);
          o.light_color = pack_unorm8(col / o.light_strength);
          o.light_direction = dir_light;
        } else if (o.intersects && dot(dir_light, lnorm) > 0) {
          const V3 so = o.point + lnorm * 0.01f;
          shadow[i] = orc_ray{{so.x, so.y, so.z}, 0.f, {dir_light.x, dir_light.y, dir_light.z}, dist_light};
          shadow_wanted[i] = 1;
        }
      }
    });
    // any_hit toward the light, then the NEE record (pt_trace.comp:202-211)
    std::vector<uint32_t> sidx;
    for (size_t i = 0; i < n; ++i)
      if (shadow_wanted[i]) sidx.push_back(uint32_t(i));
    std::vector<orc_ray> srays(sidx.size());
    for (size_t k = 0; k < sidx.size(); ++k) srays[k] = shadow[sidx[k]];
    std::vector<orc_hit> shits(sidx.size());
    jobs(pool, srays.size(), chunk, [&](size_t b, size_t e) { orc_reference_trace(F.a, srays.data() + b, e - b, 1, shits.data() + b); });
    for (size_t k = 0; k < sidx.size(); ++k) {
      if (std::isfinite(shits[k].t)) continue;
      TraceOut& o = out[sidx[k]];
      const V3 col = kLights[light_index[sidx[k]]].color;
      o.light_strength = length(col);
      o.light_color = pack_unorm8(col / o.light_strength);
      o.light_strength *= 2.0f / length(light_point[sidx[k]] - o.point);
      o.light_direction = normalize(light_point[sidx[k]] - o.point);
    }
    F.rays += n + srays.size();
    // color (pt_color.comp:113-202), then filter (pt_ray_filter.comp:49-63)
    const int32_t color_seed = F.draw();
    std::vector<uint8_t> alive(n, 0);
    jobs(pool, n, chunk, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        Path& p = paths[i];
        const TraceOut& tr = out[i];
        auto write = [&](V3 c) { accumulate(F, p, c); };
        if (tr.intersects == 0) {
          const V3 sky = mix(V3{0.3f, 0.4f, 0.5f}, V3{1.0f, 0.98f, 0.96f}, (p.direction.y + 1) / 2);
          write(p.color + clamp3(p.reflectance * sky / p.pdf, 0, 16));
          continue;
        }
        if (tr.intersects == 2) {
          write(unpack_unorm8(tr.light_color) * tr.light_strength * p.reflectance / p.pdf + p.color);
          continue;
        }
        Rng r;
        random_init(r, F.random_texture, int(p.pixel % uint32_t(F.w)), int(p.pixel / uint32_t(F.w)), color_seed);
        random_value(r, wrap_int(double(int32_t(56128966.0 * p.random_value)) + 20945293.0 * double(p.id)));
        random_value(r, wrap_int(double(int32_t(12233432.0 * p.random_value)) + 74867839.0 * double(p.id)));
        const orc_material& om = s->materials[std::min(tr.material, s->n_materials - 1)];
        Shader sh;
        sh.mat = Mat{unpack_unorm8(om.albedo_rgba8), om.ior, om.roughness, om.metallic, om.transmission};
        const float rx = random_next(r), ry = random_next(r);
        const V3 gen = sh.bounce(r, rx, ry, -p.direction, tr.normal);
        V3 refl;
        float pdf;
        sh.evaluate(tr.normal, gen, -p.direction, refl, pdf);
        if (random_next(r) > p.pdf) {  // Russian roulette with p = path pdf (pt_color.comp:145-154)
          write(p.color + p.reflectance / p.pdf);
          continue;
        }
        if (last || p.pdf < 1e-5f || pdf < 1e-6f || std::isnan(pdf) || std::isinf(pdf) || bad(refl) || bad(gen)) {
          if (bad(refl) || !std::isfinite(p.pdf) || p.pdf < 1e-5f)
            write(p.color);
          else
            write(p.color + p.reflectance / p.pdf);
          continue;
        }
        if (tr.light_strength > 0) {
          V3 lr{1, 1, 1};
          float lp = 1;
          sh.evaluate(tr.normal, tr.light_direction, -p.direction, lr, lp);
          if (!bad(lr)) p.color = p.color + unpack_unorm8(tr.light_color) * tr.light_strength * clamp3(lr, 0, 1);
        }
        p.direction = gen;
        p.origin = tr.point + faceforward(tr.normal, -p.direction, tr.normal) * 1e-4f;
        p.pdf *= pdf;
        p.reflectance = p.reflectance * refl;
        p.random_value = random_value(r, wrap_int(double(wrap_int(2341335623.0 * p.random_value)) + 357883444.0 * double(p.id)));
        alive[i] = 1;
      }
    });
    std::vector<Path> next;
    next.reserve(n);
    for (size_t i = 0; i < n; ++i)
      if (alive[i]) next.push_back(paths[i]);
    paths.swap(next);
  }
}

}  // namespace

extern "C" double orc_reference_render_time(const orc_pt_scene* s, orc_accel a, const float view_colmajor[16],
                                            const float proj_colmajor[16], int width, int height, int rows,
                                            int max_depth, int threads, int reps, int* threads_used,
                                            double* rays_per_s, double* checksum) {
  if (!s || !a || width <= 0 || height <= 0) return 0.0;
  if (rows <= 0 || rows > height) rows = height;
  const unsigned nt = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
  if (reps < 1) reps = 1;
  Frame F;
  F.s = s;
  F.a = a;
  F.w = width;
  F.h = height;
  F.rows = rows;
  if (!invert4(view_colmajor, F.inv_view) || !invert4(proj_colmajor, F.inv_proj)) return 0.0;
  std::mt19937 tex_rng(0);
  std::uniform_real_distribution<float> u01(0.f, 1.f);
  F.random_texture.resize(1024);
  for (float& v : F.random_texture) v = u01(tex_rng);
  F.fwd.resize(12 * size_t(s->geo->n_inst));
  for (uint32_t i = 0; i < s->geo->n_inst; ++i) forward_rows(s->geo->inst_inv + 12 * size_t(i), &F.fwd[12 * size_t(i)]);
  F.sums.assign(3 * size_t(width) * rows, 0.0);
  ThreadPool pool(nt);
  if (threads_used) *threads_used = int(nt);
  const size_t chunk = size_t(64) * size_t(width);  // 64-row jobs
  render_frame(F, pool, max_depth, chunk);  // warm-up
  std::vector<double> secs, rps;
  for (int r = 0; r < reps; ++r) {
    F.rays = 0;
    const auto t0 = std::chrono::steady_clock::now();
    render_frame(F, pool, max_depth, chunk);
    const auto t1 = std::chrono::steady_clock::now();
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    secs.push_back(sec);
    rps.push_back(double(F.rays) / sec);
  }
  std::sort(secs.begin(), secs.end());
  std::sort(rps.begin(), rps.end());
  const double med = secs[secs.size() / 2];
  if (rays_per_s) *rays_per_s = rps[rps.size() / 2];
  if (checksum) {
    double c = 0;
    for (double v : F.sums) c += v;
    *checksum = c / double(reps + 1);
  }
  return med > 0 ? double(width) * double(rows) / med : 0.0;
}
