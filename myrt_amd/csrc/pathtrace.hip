// Wavefront path-tracing passes for gfx950: generate, shade and resolve. Extension and shadow
// rays go through the traversal kernels of trace.hip between these passes.
//
// Replaces src/pt/pt_generate.comp:50-84, pt_color.comp:113-202, pt_ray_filter.comp:49-63 and the
// light/shadow part of pt_trace.comp:134-226:
//   * the sampler is counter-based: u = hash(pixel, sample, dimension), so a CPU oracle can
//     replay exactly the same random numbers;
//   * materials keep the reference's parameters {albedo, ior, roughness, metallic, transmission}
//     and its Fresnel-weighted lobe choice (src/pt/pbr.glsl:164-201), with physically based
//     lobes: GGX reflection (alpha = roughness^2, Smith G2), Lambert diffuse and a smooth
//     dielectric transmission lobe;
//   * quad area lights with next-event estimation and power-heuristic MIS replace the two
//     hard-coded sphere lights;
//   * survivors and shadow rays are appended to dense queues with ballots and one atomic per
//     block and queue instead of one atomic per ray.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pathtrace.hpp"

namespace mrt {
namespace ptk {

constexpr int kBlock = 256;
constexpr int kShadeBlock = 512;  // shade_kernel: 8 waves share one queue atomic
constexpr uint32_t kDelta = 1u << 8;
constexpr uint32_t kUnbiased = 1u << 9;  // a training path that never ends in the cache
constexpr float kPi = 3.14159265358979f;

struct V3 {
  float x, y, z;
};
__device__ inline V3 mk(float a, float b, float c) { return {a, b, c}; }
__device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline V3 normalize(V3 a) {
  const float l = sqrtf(dot(a, a));
  return a * (1.0f / l);
}
__device__ inline V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ inline bool finite3(V3 a) {
  return __builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z);
}

// Adds the wave's count of lanes with `bad` set to the 64-bit counter at counts + 10 (non-finite
// values rejected). Every lane of the wave must call it.
__device__ inline void count_nonfinite(uint32_t* counts, bool bad) {
  const uint64_t m = __ballot(bad);
  if (m && __lane_id() == uint32_t(__ffsll((unsigned long long)m) - 1))
    atomicAdd(reinterpret_cast<unsigned long long*>(counts + 10), (unsigned long long)__popcll(m));
}

// Sampler: lowbias32 (C. Wellons' integer hash) chained over (pixel, sample, dimension).
__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ inline float rnd(uint32_t pixel, uint32_t sample, uint32_t dim) {
  const uint32_t h = mix32(mix32(mix32(pixel) + sample) + dim * 0x9E3779B9u);
  return float(h >> 8) * (1.0f / 16777216.0f);
}
// Dimensions: 0-1 pixel jitter; bounce b uses 2 + 8b + {0 lobe, 1-2 direction, 3 roulette,
// 4-5 light point, 6 light choice}.
__device__ inline uint32_t dim_of(uint32_t bounce, uint32_t k) { return 2u + 8u * bounce + k; }
constexpr uint32_t kTrainDim = 0x40000000u;  // dimension that selects training paths
constexpr uint32_t kUnbiasedDim = 0x40000001u;  // dimension that selects unbiased training paths
constexpr uint32_t kLensDim = 0x30000000u;   // lens samples: kLensDim + 2 * try + {0, 1}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Appends the threads with pred[q] set to queue q (Q queues at once) with one atomic per block and
// queue. A queue counter is a single address and atomics on one address are serialised (about
// 11 ns each here), so one atomic per wave and queue set the pace of a pass over tens of millions
// of paths. Slots follow thread order within the block. Every thread of the block must call it.
template <int BLOCK, int Q>
__device__ inline void block_append(uint32_t* const (&counter)[Q], const bool (&pred)[Q], uint32_t (&slot)[Q]) {
  constexpr int kWaves = BLOCK / 64;
  __shared__ uint32_t cnt[Q][kWaves];
  __shared__ uint32_t base[Q];
  const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
  uint64_t mask[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    mask[q] = __ballot(pred[q]);
    if (lane == 0) cnt[q][wave] = uint32_t(__popcll(mask[q]));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if (threadIdx.x == uint32_t(q)) {
      uint32_t tot = 0;
      for (int w = 0; w < kWaves; ++w) tot += cnt[q][w];
      base[q] = tot ? atomicAdd(counter[q], tot) : 0u;
    }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    uint32_t s = base[q];
    for (uint32_t w = 0; w < wave; ++w) s += cnt[q][w];
    slot[q] = s + uint32_t(__popcll(mask[q] & below));
  }
}

// Adds v[0..N) to dst[0..N) for the lanes with `on` set, one atomic per float and per run of
// neighbouring lanes with the same key: each run is summed in registers first (a suffix scan that
// stops at the run's last lane), and the run's first lane adds the total. Paths are laid out
// pixel-major, so the samples of one pixel sit in neighbouring lanes. Lanes without a contribution
// carry the key 0xFFFFFFFF and add nothing. Every lane of the wave must call it.
template <int N>
__device__ inline void wave_run_add(float* dst, uint32_t key, float (&v)[N], bool on) {
  const int lane = int(lane_id());
  if (!on) key = 0xFFFFFFFFu;
  const uint32_t next_key = __shfl_down(key, 1u);
  const uint32_t prev_key = __shfl_up(key, 1u);
  const uint64_t ends = __ballot(lane == 63 || next_key != key);  // last lane of each run
  const int run_end = lane + __builtin_ctzll(ends >> lane);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (!on) v[k] = 0.f;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const bool take = lane + int(off) <= run_end;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const float o = __shfl_down(v[k], off);
      if (take) v[k] += o;
    }
  }
  if (on && (lane == 0 || prev_key != key)) {
#pragma unroll
    for (int k = 0; k < N; ++k) unsafeAtomicAdd(dst + k, v[k]);
  }
}

// add_pixel for a whole wave (uniform control flow): lanes of one pixel add once.
__device__ inline void wave_add_pixel(float4* accum, uint32_t pixel, V3 c, bool on) {
  float v[3] = {c.x, c.y, c.z};
  wave_run_add<3>(reinterpret_cast<float*>(accum + pixel), pixel, v, on);
}

// Adds a path contribution to its training header (one writer per path and pass).
__device__ inline void add_train(const PtBuffers& b, uint32_t slot, V3 c) {
  if (slot == 0xFFFFFFFFu) return;
  TrainPath& tp = b.cache.tpaths[slot];
  tp.T[0] += c.x;
  tp.T[1] += c.y;
  tp.T[2] += c.z;
}

__device__ inline V3 unpack_rgba8(uint32_t c) {
  return mk(float(c & 0xffu) * (1.0f / 255.0f), float((c >> 8) & 0xffu) * (1.0f / 255.0f),
            float((c >> 16) & 0xffu) * (1.0f / 255.0f));
}

// Bilinear lookup with wrap (repeat) of an albedo texture at (u, v), texel centres at
// (i + 0.5) / size.
__device__ inline V3 texture_lookup(const PtScene& sc, const TexInfo& t, float u, float v) {
  float fx = u * float(t.w) - 0.5f, fy = v * float(t.h) - 0.5f;
  fx = fx - floorf(fx / float(t.w)) * float(t.w);
  fy = fy - floorf(fy / float(t.h)) * float(t.h);
  uint32_t x0 = min(uint32_t(fx), t.w - 1), y0 = min(uint32_t(fy), t.h - 1);
  const float tx = fminf(fmaxf(fx - float(x0), 0.f), 1.f), ty = fminf(fmaxf(fy - float(y0), 0.f), 1.f);
  const uint32_t x1 = x0 + 1 == t.w ? 0 : x0 + 1, y1 = y0 + 1 == t.h ? 0 : y0 + 1;
  const float4* T = sc.texels + t.offset;
  const float4 a = T[y0 * t.w + x0], b = T[y0 * t.w + x1], c = T[y1 * t.w + x0], e = T[y1 * t.w + x1];
  const float w00 = (1.f - tx) * (1.f - ty), w10 = tx * (1.f - ty), w01 = (1.f - tx) * ty, w11 = tx * ty;
  return mk(a.x * w00 + b.x * w10 + c.x * w01 + e.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + e.y * w11,
            a.z * w00 + b.z * w10 + c.z * w01 + e.z * w11);
}

// Cube-map lookup (face selection and face coordinates as in the OpenGL specification's cube
// map table), bilinear within the face, clamped to its edge texels.
__device__ inline V3 cubemap_lookup(const PtScene& sc, V3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  uint32_t face;
  float scc, tcc, ma;
  if (ax >= ay && ax >= az) {
    face = d.x >= 0.f ? 0u : 1u;
    scc = d.x >= 0.f ? -d.z : d.z;
    tcc = -d.y;
    ma = ax;
  } else if (ay >= az) {
    face = d.y >= 0.f ? 2u : 3u;
    scc = d.x;
    tcc = d.y >= 0.f ? d.z : -d.z;
    ma = ay;
  } else {
    face = d.z >= 0.f ? 4u : 5u;
    scc = d.z >= 0.f ? d.x : -d.x;
    tcc = -d.y;
    ma = az;
  }
  const float n = float(sc.cube_size);
  const float u = 0.5f * (scc / ma + 1.0f), v = 0.5f * (tcc / ma + 1.0f);
  const float fx = fminf(fmaxf(u * n - 0.5f, 0.f), n - 1.f), fy = fminf(fmaxf(v * n - 0.5f, 0.f), n - 1.f);
  const uint32_t x0 = uint32_t(fx), y0 = uint32_t(fy);
  const uint32_t x1 = min(x0 + 1, sc.cube_size - 1), y1 = min(y0 + 1, sc.cube_size - 1);
  const float tx = fx - float(x0), ty = fy - float(y0);
  const float4* F = sc.cubemap + size_t(face) * sc.cube_size * sc.cube_size;
  const float4 a = F[y0 * sc.cube_size + x0], b = F[y0 * sc.cube_size + x1];
  const float4 c = F[y1 * sc.cube_size + x0], e = F[y1 * sc.cube_size + x1];
  const float w00 = (1.f - tx) * (1.f - ty), w10 = tx * (1.f - ty), w01 = (1.f - tx) * ty, w11 = tx * ty;
  return mk(a.x * w00 + b.x * w10 + c.x * w01 + e.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + e.y * w11,
            a.z * w00 + b.z * w10 + c.z * w01 + e.z * w11);
}

// Sky gradient of src/pt/pt_color.comp:198, scaled.
__device__ inline V3 sky_gradient(V3 d, float scale) {
  const float t = (d.y + 1.0f) * 0.5f;
  return mk(0.3f + (1.0f - 0.3f) * t, 0.4f + (0.98f - 0.4f) * t, 0.5f + (0.96f - 0.5f) * t) * scale;
}

// Environment of src/pt/pt_color.comp:191-199: the cube map if one is set, else the sky
// gradient; scaled.
__device__ inline V3 environment(const PtScene& sc, V3 d) {
  if (sc.cubemap) return cubemap_lookup(sc, d) * sc.env_scale;
  return sky_gradient(d, sc.env_scale);
}

__device__ inline float ggx_d(float nh, float a2) {
  const float den = nh * nh * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * den * den);
}
__device__ inline float ggx_g1(float nx, float a2) { return 2.0f * nx / (nx + sqrtf(a2 + (1.0f - a2) * nx * nx)); }

struct Lobes {
  V3 albedo, spec_col;
  float a2, metallic;
  float ws, wt, wd;  // lobe weights = selection probabilities (sum to 1)
};

// Non-delta part of the BSDF and its sampling density: ws * GGX + wd * Lambert.
__device__ inline void eval_bsdf(const Lobes& L, V3 n, V3 wo, V3 wi, V3& f, float& pdf) {
  const float nv = dot(n, wo), nl = dot(n, wi);
  f = mk(0, 0, 0);
  pdf = 0.f;
  if (nv <= 0.f || nl <= 0.f) return;
  const V3 h = normalize(wo + wi);
  const float nh = fmaxf(dot(n, h), 0.f), vh = fmaxf(dot(wo, h), 1e-8f);
  const float D = ggx_d(nh, L.a2);
  const float G = ggx_g1(nv, L.a2) * ggx_g1(nl, L.a2);
  const float fs = D * G / (4.0f * nv * nl);
  const float fd = (1.0f - L.metallic) / kPi;
  f = L.spec_col * (L.ws * fs) + L.albedo * (L.wd * fd);
  pdf = L.ws * (D * nh / (4.0f * vh)) + L.wd * (nl / kPi);
}

// Octahedral normal of an SDF hit record (16 + 16 bits).
__device__ inline V3 oct_decode(uint32_t w) {
  float u = float(w & 0xffffu) * (2.0f / 65535.0f) - 1.0f;
  float v = float(w >> 16) * (2.0f / 65535.0f) - 1.0f;
  const float z = 1.0f - fabsf(u) - fabsf(v);
  if (z < 0.f) {
    const float uu = (1.0f - fabsf(v)) * (u >= 0.f ? 1.f : -1.f);
    const float vv = (1.0f - fabsf(u)) * (v >= 0.f ? 1.f : -1.f);
    u = uu;
    v = vv;
  }
  return normalize(mk(u, v, z));
}

// Orthonormal basis around n (Duff et al. 2017).
__device__ inline void onb(V3 n, V3& t, V3& b) {
  const float s = n.z >= 0.f ? 1.0f : -1.0f;
  const float a = -1.0f / (s + n.z);
  const float c = n.x * n.y * a;
  t = mk(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
  b = mk(c, s + n.y * n.y * a, -n.y);
}

// Bilinear lookup of the bokeh mask at (u, v) in [0, 1]^2, texel centres at (i + 0.5) / size,
// coordinates clamped to the edge texels.
__device__ inline V3 bokeh_sample(const Camera& cam, float u, float v) {
  const float fx = fminf(fmaxf(u * float(cam.bokeh_w) - 0.5f, 0.f), float(cam.bokeh_w - 1));
  const float fy = fminf(fmaxf(v * float(cam.bokeh_h) - 0.5f, 0.f), float(cam.bokeh_h - 1));
  const uint32_t x0 = uint32_t(fx), y0 = uint32_t(fy);
  const uint32_t x1 = min(x0 + 1, cam.bokeh_w - 1), y1 = min(y0 + 1, cam.bokeh_h - 1);
  const float tx = fx - float(x0), ty = fy - float(y0);
  const float4 a = cam.bokeh[y0 * cam.bokeh_w + x0], b = cam.bokeh[y0 * cam.bokeh_w + x1];
  const float4 c = cam.bokeh[y1 * cam.bokeh_w + x0], e = cam.bokeh[y1 * cam.bokeh_w + x1];
  const float w00 = (1.f - tx) * (1.f - ty), w10 = tx * (1.f - ty), w01 = (1.f - tx) * ty, w11 = tx * ty;
  return mk(a.x * w00 + b.x * w10 + c.x * w01 + e.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + e.y * w11,
            a.z * w00 + b.z * w10 + c.z * w01 + e.z * w11);
}

__global__ __launch_bounds__(kBlock) void generate_kernel(Camera cam, PtBuffers b, uint32_t cur, uint32_t n_paths,
                                                          uint32_t sample_base) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) b.counts[cur] = n_paths;
  if (blockIdx.x * kBlock >= n_paths) return;  // whole block idle
  const bool live = i < n_paths;
  const uint32_t ii = live ? i : n_paths - 1;  // idle lanes mirror a live one but write nothing
  // Pixel-major: the batch's samples of one pixel are neighbouring paths (n_paths is a whole number
  // of samples per pixel), so a wave's primary rays leave a few pixels and its pixel adds combine.
  const uint32_t npix = cam.width * cam.height;
  const uint32_t nspp = max(1u, n_paths / npix);
  const uint32_t pixel = ii / nspp;
  const uint32_t sample = sample_base + ii % nspp;
  const uint32_t x = pixel % cam.width, y = pixel / cam.width;
  const float nx = (float(x) + rnd(pixel, sample, 0)) / float(cam.width) * 2.0f - 1.0f;
  const float ny = (float(y) + rnd(pixel, sample, 1)) / float(cam.height) * 2.0f - 1.0f;
  const float* m = cam.inv_view_proj;
  const float qx = m[0] * nx + m[4] * ny + m[8] + m[12];
  const float qy = m[1] * nx + m[5] * ny + m[9] + m[13];
  const float qz = m[2] * nx + m[6] * ny + m[10] + m[14];
  const float qw = m[3] * nx + m[7] * ny + m[11] + m[15];
  V3 target = qw != 0.f ? mk(qx / qw, qy / qw, qz / qw) : mk(qx, qy, qz);
  V3 eye = ld3(cam.eye);
  V3 d = qw != 0.f ? normalize(target - eye) : normalize(target);
  V3 tint = mk(1.0f, 1.0f, 1.0f);
  if (cam.dof) {
    // A lens offset uniform in the disk of radius 0.49; with a mask, offsets where the mask is
    // black are redrawn (up to 8 tries, then the sample carries no weight) and the mask colour
    // tints the path. The focal point at `focus` along the pinhole ray stays fixed.
    float ox = 0.f, oy = 0.f;
    bool ok = false;
    for (uint32_t k = 0; k < 8 && !ok; ++k) {
      const float a = rnd(pixel, sample, kLensDim + 2 * k) * 6.2831853f;
      const float rr = sqrtf(rnd(pixel, sample, kLensDim + 2 * k + 1)) * 0.49f;
      ox = rr * sinf(a);
      oy = rr * cosf(a);
      if (!cam.bokeh) {
        ok = true;
      } else {
        tint = bokeh_sample(cam, 0.5f + ox, 0.5f + oy);
        ok = fmaxf(fmaxf(tint.x, tint.y), tint.z) >= 1e-5f;
      }
    }
    if (!ok) tint = mk(0.f, 0.f, 0.f);
    const V3 center = eye + d * cam.focus;
    const V3 local = normalize(mk(ox * cam.lens[0] * 0.5f, oy * cam.lens[1] * 0.5f, cam.focus));
    V3 u = cross(d, mk(1.f, 0.f, 0.f));
    if (dot(u, u) == 0.f) u = cross(d, mk(0.f, 1.f, 0.f));
    u = normalize(u);
    const V3 v = cross(d, u);
    const V3 nd = normalize(u * local.x + v * local.y + d * local.z);
    eye = center - nd * cam.focus;
    d = nd;
  }
  if (live) {
    float4* r = reinterpret_cast<float4*>(b.rays[cur]) + 2 * size_t(i);
    r[0] = make_float4(eye.x, eye.y, eye.z, 0.0f);
    r[1] = make_float4(d.x, d.y, d.z, __int_as_float(0x7f800000));
  }
  PathState ps;
  ps.thr[0] = tint.x;
  ps.thr[1] = tint.y;
  ps.thr[2] = tint.z;
  ps.pixel = pixel;
  ps.sample = sample;
  ps.flags = kDelta;
  ps.last_pdf = 0.f;
  ps.pad = 0xFFFFFFFFu;
  // Training paths, stratified: one path of every train_group consecutive paths (the member drawn
  // by a hash of the group's first path), whose training slot is the group index, so no queue
  // atomic is needed (a queue atomic per block took 0.3 ms for 8M paths).
  bool want_train = false;
  uint32_t slot = 0;
  if (b.cache.train) {
    const uint32_t g = b.cache.train_group;
    slot = ii / g;
    const uint32_t first = slot * g, members = min(g, n_paths - first);
    const uint32_t f_pixel = first / nspp, f_sample = sample_base + first % nspp;
    want_train = live && ii - first == mix32(mix32(mix32(f_pixel) + f_sample) + kTrainDim) % members;
    if (i == 0) b.counts[3] = (n_paths + g - 1) / g;
  }
  if (want_train && slot < b.cache.max_train_paths) {
    ps.pad = slot;
    if (rnd(pixel, sample, kUnbiasedDim) < b.cache.unbiased_fraction) ps.flags |= kUnbiased;
    TrainPath tp;
    tp.T[0] = tp.T[1] = tp.T[2] = 0.f;
    tp.nv = 0;
    b.cache.tpaths[slot] = tp;
  }
  if (live) b.paths[cur][i] = ps;
}

__global__ __launch_bounds__(kShadeBlock) void shade_kernel(PtScene sc, PtBuffers b, uint32_t cur, uint32_t max_depth) {
  const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
  const uint32_t n = b.counts[cur];
  if (i == 0) atomicAdd(reinterpret_cast<unsigned long long*>(b.counts + 8), (unsigned long long)n);  // rays traced
  if (blockIdx.x * kShadeBlock >= n) return;  // whole block idle (uniform)
  const uint32_t nxt = cur ^ 1u;
  const bool live = i < n;

  bool survive = false, shadow = false, want_query = false;
  CacheQuery cache_q{};
  V3 cache_thr = mk(0, 0, 0);
  V3 new_o = mk(0, 0, 0), new_d = mk(0, 0, 1);
  V3 sh_o = mk(0, 0, 0), sh_d = mk(0, 0, 1), sh_c = mk(0, 0, 0);
  V3 pix_c = mk(0, 0, 0);  // emitter or environment seen by the ray
  bool pix_on = false, aux_on = false;
  bool bad = false;  // a non-finite value was rejected (pt_color.comp:84-98, 156-166)
  float aux_v[8] = {};  // primary-hit features (albedo, depth, shading normal, count)
  PathState ps{};
  if (live) {
    const uint32_t src = b.perm ? b.perm[i] : i;  // shading order by material, when sorted
    ps = b.paths[cur][src];
    const float4* rp = reinterpret_cast<const float4*>(b.rays[cur]) + 2 * size_t(src);
    const float4 r0 = rp[0], r1 = rp[1];
    const V3 ro = mk(r0.x, r0.y, r0.z), rd = mk(r1.x, r1.y, r1.z);
    const uint4 hit = reinterpret_cast<const uint4*>(b.hits)[src];
    bool found = hit.z != 0xFFFFFFFFu;
    const float t_hit = found ? __uint_as_float(hit.x) : __int_as_float(0x7f800000);
    const uint32_t bounce = ps.flags & 0xffu;
    V3 thr = mk(ps.thr[0], ps.thr[1], ps.thr[2]);

    // Normals of the surface hit. A non-finite geometric normal (a degenerate triangle) makes the
    // hit a miss, as pt_trace.comp:186-189 does; a non-finite shading normal (say NaN vertex
    // normals) falls back to the geometric one.
    V3 ng = mk(0, 0, 1), ns = ng;
    uint32_t mi = 0;
    float hu = 0.f, hv = 0.f;
    bool has_uv = false;
    if (found) {
      const uint32_t inst = hit.z, prim = hit.y;
      if (inst & 0x80000000u) {
        // SDF surface: the record carries the world normal and the material.
        ng = oct_decode(hit.w);
        ns = ng;
        mi = prim;
      } else {
        const InstanceRecord& ir = sc.instances[inst];
        const GeoInfo gi = sc.geos[ir.geom];
        const uint32_t i0 = sc.idx[gi.idx_base + 3 * prim], i1 = sc.idx[gi.idx_base + 3 * prim + 1],
                       i2 = sc.idx[gi.idx_base + 3 * prim + 2];
        const float4 p0 = sc.pos[gi.vtx_base + i0], p1 = sc.pos[gi.vtx_base + i1], p2 = sc.pos[gi.vtx_base + i2];
        const float b1 = float(hit.w & 0xffffu) * (1.0f / 65535.0f), b2 = float(hit.w >> 16) * (1.0f / 65535.0f);
        const float b0 = 1.0f - b1 - b2;
        if (sc.uv) {
          const float2 t0 = sc.uv[gi.vtx_base + i0], t1 = sc.uv[gi.vtx_base + i1], t2 = sc.uv[gi.vtx_base + i2];
          hu = b0 * t0.x + b1 * t1.x + b2 * t2.x;
          hv = b0 * t0.y + b1 * t1.y + b2 * t2.y;
          has_uv = true;
        }
        const V3 e1 = mk(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z), e2 = mk(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z);
        const V3 ngo = cross(e1, e2);
        V3 nso = ngo;
        if (gi.has_normals) {
          const float4 n0 = sc.nrm[gi.vtx_base + i0], n1 = sc.nrm[gi.vtx_base + i1], n2 = sc.nrm[gi.vtx_base + i2];
          nso = mk(b0 * n0.x + b1 * n1.x + b2 * n2.x, b0 * n0.y + b1 * n1.y + b2 * n2.y,
                   b0 * n0.z + b1 * n1.z + b2 * n2.z);
        }
        const float* iv = ir.inv;  // normals transform with the transpose of M^-1
        ng = normalize(mk(iv[0] * ngo.x + iv[4] * ngo.y + iv[8] * ngo.z, iv[1] * ngo.x + iv[5] * ngo.y + iv[9] * ngo.z,
                          iv[2] * ngo.x + iv[6] * ngo.y + iv[10] * ngo.z));
        ns = normalize(mk(iv[0] * nso.x + iv[4] * nso.y + iv[8] * nso.z, iv[1] * nso.x + iv[5] * nso.y + iv[9] * nso.z,
                          iv[2] * nso.x + iv[6] * nso.y + iv[10] * nso.z));
        mi = sc.inst_material[inst];
      }
      if (!finite3(ng)) {
        found = false;
        bad = true;
      } else if (!finite3(ns)) {
        ns = ng;
        bad = true;
      }
    }

    // Emitters along the ray (lights are not in the BVH).
    int light_hit = -1;
    float t_light = t_hit;
    float cos_light = 0.f;
    for (uint32_t l = 0; l < sc.num_lights; ++l) {
      const QuadLight& q = sc.lights[l];
      const V3 u = ld3(q.u), v = ld3(q.v), c = ld3(q.corner);
      const V3 nl = cross(u, v);
      const float area = sqrtf(dot(nl, nl));
      const V3 n1 = nl * (1.0f / area);
      const float den = dot(n1, rd);
      if (den >= 0.f) continue;  // one-sided: seen from behind it neither emits nor blocks
      const float t = dot(n1, c - ro) / den;
      if (!(t > 1e-6f && t < t_light)) continue;
      const V3 pq = ro + rd * t - c;
      const float a = dot(pq, u) / dot(u, u), bb = dot(pq, v) / dot(v, v);
      if (a < 0.f || a > 1.f || bb < 0.f || bb > 1.f) continue;
      light_hit = int(l);
      t_light = t;
      cos_light = -den;
    }
    if (light_hit >= 0) {
      const QuadLight& q = sc.lights[light_hit];
      float w = 1.0f;
      if (!(ps.flags & kDelta)) {
        const V3 nl = cross(ld3(q.u), ld3(q.v));
        const float area = sqrtf(dot(nl, nl));
        const float pdf_l = t_light * t_light / (cos_light * area * float(sc.num_lights));
        w = ps.last_pdf * ps.last_pdf / (ps.last_pdf * ps.last_pdf + pdf_l * pdf_l);
      }
      const V3 c = thr * ld3(q.emission) * w;
      if (finite3(c)) {
        pix_c = c;
        pix_on = true;
        add_train(b, ps.pad, c);
      } else {
        bad = true;
      }
    } else if (!found) {
      const V3 c = thr * environment(sc, rd);
      if (finite3(c)) {
        pix_c = c;
        pix_on = true;
        add_train(b, ps.pad, c);
      } else {
        bad = true;
      }
    } else {
      // ---- surface interaction ----
      const V3 p = ro + rd * t_hit;
      const V3 wo = -rd;
      const bool entering = dot(rd, ng) < 0.f;
      if (!entering) {
        ng = -ng;
        ns = -ns;
      }
      if (dot(ns, ng) <= 0.f) ns = ng;
      const float eps = 1e-4f * (1.0f + fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z)));

      if (mi >= sc.num_materials) mi = 0;
      const Material mat = sc.materials[mi];
      Lobes L;
      L.albedo = unpack_rgba8(mat.albedo_rgba8);
      if (has_uv && mat.texture >= 1u && mat.texture <= sc.num_textures) {
        // The albedo texture replaces the colour, decoded with gamma 2.2 (material_new.glsl:29-31).
        const V3 t = texture_lookup(sc, sc.textures[mat.texture - 1], hu, 1.0f - hv);
        L.albedo = mk(powf(fmaxf(t.x, 0.f), 2.2f), powf(fmaxf(t.y, 0.f), 2.2f), powf(fmaxf(t.z, 0.f), 2.2f));
      }
      L.metallic = mat.metallic;
      const float sm = sqrtf(fmaxf(mat.metallic, 0.f));
      L.spec_col = mk(1.0f + (L.albedo.x - 1.0f) * sm, 1.0f + (L.albedo.y - 1.0f) * sm, 1.0f + (L.albedo.z - 1.0f) * sm);
      const float alpha = fmaxf(mat.roughness * mat.roughness, 1e-3f);
      L.a2 = alpha * alpha;
      const float f0d = (mat.ior - 1.0f) / (mat.ior + 1.0f);
      const float f0 = f0d * f0d + (1.0f - f0d * f0d) * mat.metallic;
      const float cv = fminf(fmaxf(dot(ns, wo), 0.f), 1.0f);
      const float om = 1.0f - cv;
      const float Fv = f0 + (1.0f - f0) * (om * om * om * om * om);
      L.ws = Fv;
      L.wt = (1.0f - Fv) * mat.transmission;
      L.wd = (1.0f - Fv) * (1.0f - mat.transmission);
      // Primary-hit features for the denoiser (sums; the host divides by the sample count).
      if (bounce == 0 && b.aux) {
        aux_on = true;
        aux_v[0] = L.albedo.x;
        aux_v[1] = L.albedo.y;
        aux_v[2] = L.albedo.z;
        aux_v[3] = t_hit;
        aux_v[4] = ns.x;
        aux_v[5] = ns.y;
        aux_v[6] = ns.z;
        aux_v[7] = 1.0f;
      }

      // ---- radiance cache: end here (non-training paths) or record the vertex (training paths) ----
      bool cached = false;
      CacheQuery cq;
      if (b.cache.terminate || ps.pad != 0xFFFFFFFFu) {
        cq.pos[0] = p.x;
        cq.pos[1] = p.y;
        cq.pos[2] = p.z;
        cq.roughness = mat.roughness;
        cq.dir[0] = wo.x;
        cq.dir[1] = wo.y;
        cq.dir[2] = wo.z;
        cq.pad0 = 0.f;
        cq.nrm[0] = ns.x;
        cq.nrm[1] = ns.y;
        cq.nrm[2] = ns.z;
        cq.pad1 = 0.f;
        cq.albedo[0] = L.albedo.x;
        cq.albedo[1] = L.albedo.y;
        cq.albedo[2] = L.albedo.z;
        cq.pad2 = 0.f;
      }
      // At vertex train_end every path but the unbiased training paths ends in the cache, on
      // dielectrics too. For a training path the cache's prediction becomes the tail radiance of
      // its earlier vertices (self-training); this vertex is not recorded, since its target would
      // be the prediction itself.
      const bool at_end = b.cache.terminate && b.cache.train_end != 0u && bounce >= b.cache.train_end &&
                          !(ps.flags & kUnbiased);
      const bool train_ends = ps.pad != 0xFFFFFFFFu && at_end;
      if (train_ends) {
        cached = true;
      } else if (ps.pad != 0xFFFFFFFFu) {
        TrainPath& tp = b.cache.tpaths[ps.pad];
        if (tp.nv < b.cache.max_vertices) {
          TrainVertex& tv = b.cache.tverts[size_t(ps.pad) * b.cache.max_vertices + tp.nv];
          tv.q = cq;
          tv.thr[0] = thr.x;
          tv.thr[1] = thr.y;
          tv.thr[2] = thr.z;
          tv.prefix[0] = tp.T[0];
          tv.prefix[1] = tp.T[1];
          tv.prefix[2] = tp.T[2];
          tp.nv = tp.nv + 1;
        }
      } else if (b.cache.terminate && bounce >= b.cache.min_bounce &&
                 (mat.roughness >= b.cache.min_roughness || bounce > b.cache.min_bounce) &&
                 mat.transmission < 0.5f) {
        cached = true;
      } else if (at_end) {
        cached = true;
      }
      if (cached) {
        cache_q = cq;
        cache_thr = thr;
        want_query = true;
      }

      // ---- next-event estimation toward one quad light ----
      if (!cached && sc.num_lights > 0 && (L.ws + L.wd) > 0.f) {
        uint32_t l = uint32_t(rnd(ps.pixel, ps.sample, dim_of(bounce, 6)) * float(sc.num_lights));
        if (l >= sc.num_lights) l = sc.num_lights - 1;
        const QuadLight& q = sc.lights[l];
        const V3 u = ld3(q.u), v = ld3(q.v);
        const float r1 = rnd(ps.pixel, ps.sample, dim_of(bounce, 4)), r2 = rnd(ps.pixel, ps.sample, dim_of(bounce, 5));
        const V3 x = ld3(q.corner) + u * r1 + v * r2;
        const V3 nl = cross(u, v);
        const float area = sqrtf(dot(nl, nl));
        V3 wi = x - p;
        const float d2 = dot(wi, wi);
        wi = wi * (1.0f / sqrtf(d2));
        const float cl = -dot(nl, wi) / area;
        if (cl > 0.f && dot(ng, wi) > 0.f) {
          V3 f;
          float pdf_b;
          eval_bsdf(L, ns, wo, wi, f, pdf_b);
          const float pdf_l = d2 / (cl * area * float(sc.num_lights));
          const float w = pdf_l * pdf_l / (pdf_l * pdf_l + pdf_b * pdf_b);
          const V3 c = thr * f * ld3(q.emission) * (dot(ns, wi) * w / pdf_l);
          if (!finite3(c)) {
            bad = true;
          } else if (c.x > 0.f || c.y > 0.f || c.z > 0.f) {
            shadow = true;
            sh_o = p + ng * eps;
            sh_d = x - sh_o;
            sh_c = c;
          }
        }
      }

      // ---- sample the next direction ----
      if (cached) goto done_shading;
      {
      const float ul = rnd(ps.pixel, ps.sample, dim_of(bounce, 0));
      const float u1 = rnd(ps.pixel, ps.sample, dim_of(bounce, 1)), u2 = rnd(ps.pixel, ps.sample, dim_of(bounce, 2));
      bool ok = true;
      V3 wi;
      V3 origin = p + ng * eps;
      if (ul >= L.ws && ul < L.ws + L.wt) {
        // Smooth dielectric transmission (delta lobe), total internal reflection as a mirror.
        const float eta = entering ? 1.0f / mat.ior : mat.ior;
        const float ci = fminf(fmaxf(dot(wo, ns), 0.f), 1.0f);
        const float s2 = eta * eta * (1.0f - ci * ci);
        if (s2 > 1.0f) {
          wi = ns * (2.0f * ci) - wo;
        } else {
          wi = normalize(-wo * eta + ns * (eta * ci - sqrtf(1.0f - s2)));
          origin = p - ng * eps;
        }
        thr = thr * L.albedo;
        ps.flags |= kDelta;
      } else {
        if (ul < L.ws) {
          const float cos_t = sqrtf((1.0f - u1) / (1.0f + (L.a2 - 1.0f) * u1));
          const float sin_t = sqrtf(fmaxf(0.f, 1.0f - cos_t * cos_t));
          const float phi = 2.0f * kPi * u2;
          V3 t, bt;
          onb(ns, t, bt);
          const V3 h = t * (sin_t * cosf(phi)) + bt * (sin_t * sinf(phi)) + ns * cos_t;
          wi = h * (2.0f * dot(wo, h)) - wo;
        } else {
          const float rr = sqrtf(u1), phi = 2.0f * kPi * u2;
          V3 t, bt;
          onb(ns, t, bt);
          wi = t * (rr * cosf(phi)) + bt * (rr * sinf(phi)) + ns * sqrtf(fmaxf(0.f, 1.0f - u1));
        }
        V3 f;
        float pdf;
        eval_bsdf(L, ns, wo, wi, f, pdf);
        if (dot(ng, wi) <= 0.f || pdf <= 0.f) {
          ok = false;
        } else {
          thr = thr * f * (dot(ns, wi) / pdf);
          ps.last_pdf = pdf;
          ps.flags &= ~kDelta;
        }
      }
      if (ok && !finite3(thr)) {
        ok = false;  // the path ends (pt_color.comp:156-166)
        bad = true;
      }
      if (ok && bounce >= 3) {
        const float q = fminf(0.95f, fmaxf(fmaxf(thr.x, thr.y), thr.z));
        if (rnd(ps.pixel, ps.sample, dim_of(bounce, 3)) >= q)
          ok = false;
        else
          thr = thr * (1.0f / q);
      }
      if (ok && bounce + 1 < max_depth && (thr.x > 0.f || thr.y > 0.f || thr.z > 0.f)) {
        survive = true;
        new_o = origin;
        new_d = wi;
        ps.thr[0] = thr.x;
        ps.thr[1] = thr.y;
        ps.thr[2] = thr.z;
        ps.flags = (ps.flags & ~0xffu) | (bounce + 1);
      }
      }
    done_shading:;
    }
  }
  wave_add_pixel(b.accum, ps.pixel, pix_c, pix_on);
  count_nonfinite(b.counts, bad);
  if (b.aux) wave_run_add<8>(reinterpret_cast<float*>(b.aux + 2 * size_t(ps.pixel)), ps.pixel, aux_v, aux_on);
  uint32_t* const queues[3] = {&b.counts[nxt], &b.counts[2], &b.counts[4]};
  const bool append[3] = {survive, shadow, want_query};
  uint32_t slots[3];
  block_append<kShadeBlock, 3>(queues, append, slots);
  const uint32_t slot = slots[0], sslot = slots[1], qslot = slots[2];
  if (survive) {
    float4* r = reinterpret_cast<float4*>(b.rays[nxt]) + 2 * size_t(slot);
    r[0] = make_float4(new_o.x, new_o.y, new_o.z, 0.0f);
    r[1] = make_float4(new_d.x, new_d.y, new_d.z, __int_as_float(0x7f800000));
    b.paths[nxt][slot] = ps;
  }
  if (shadow) {
    float4* r = reinterpret_cast<float4*>(b.shadow_rays) + 2 * size_t(sslot);
    r[0] = make_float4(sh_o.x, sh_o.y, sh_o.z, 0.0f);
    r[1] = make_float4(sh_d.x, sh_d.y, sh_d.z, 1.0f - 1e-3f);
    ShadowRecord rec;
    rec.contrib[0] = sh_c.x;
    rec.contrib[1] = sh_c.y;
    rec.contrib[2] = sh_c.z;
    rec.pixel = ps.pixel;
    rec.train_slot = ps.pad;
    rec.pad[0] = rec.pad[1] = rec.pad[2] = 0;
    b.shadow[sslot] = rec;
  }
  if (want_query && qslot < b.cache.max_queries) {
    b.cache.queries[qslot] = cache_q;
    CacheRecord cr;
    cr.thr[0] = cache_thr.x;
    cr.thr[1] = cache_thr.y;
    cr.thr[2] = cache_thr.z;
    cr.pixel = ps.pixel;
    cr.train_slot = ps.pad;
    cr.pad[0] = cr.pad[1] = cr.pad[2] = 0;
    b.cache.qrec[qslot] = cr;
  }
}

__global__ __launch_bounds__(kBlock) void resolve_kernel(PtBuffers b) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) atomicAdd(reinterpret_cast<unsigned long long*>(b.counts + 8), (unsigned long long)b.counts[2]);
  const uint32_t n = b.counts[2];
  if (blockIdx.x * kBlock >= n) return;  // whole block idle (uniform)
  const bool on = i < n && !b.occluded[i];
  ShadowRecord r{};
  if (on) r = b.shadow[i];
  const V3 c = mk(r.contrib[0], r.contrib[1], r.contrib[2]);
  wave_add_pixel(b.accum, r.pixel, c, on);
  if (on) add_train(b, r.train_slot, c);
}

// Paths that ended in the cache: add throughput times the predicted radiance.
__global__ __launch_bounds__(kBlock) void cache_resolve_kernel(PtBuffers b) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = min(b.counts[4], b.cache.max_queries);
  if (blockIdx.x * kBlock >= n) return;  // whole block idle (uniform)
  const bool on = i < n;
  CacheRecord r{};
  V3 c = mk(0, 0, 0);
  bool bad = false;
  if (on) {
    r = b.cache.qrec[i];
    const float* y = b.cache.qout + 3 * size_t(i);
    c = mk(r.thr[0] * fmaxf(y[0], 0.f), r.thr[1] * fmaxf(y[1], 0.f), r.thr[2] * fmaxf(y[2], 0.f));
    bad = !finite3(c);
  }
  wave_add_pixel(b.accum, r.pixel, c, on && !bad);
  if (on && !bad) add_train(b, r.train_slot, c);
  count_nonfinite(b.counts, bad);
}

// Training samples: suffix radiance (T - prefix) / thr of every recorded vertex.
__global__ __launch_bounds__(kBlock) void train_batch_kernel(PtBuffers b) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n_paths = min(b.counts[3], b.cache.max_train_paths);
  // Training paths were numbered in pixel order; visit them in a scattered order (j * s mod n with
  // s prime to n, a bijection) so that the first max_batch samples cover the whole image.
  const uint32_t j = i / b.cache.max_vertices, k = i % b.cache.max_vertices;
  const uint64_t stride = (n_paths % 1000003u) ? 1000003u : 999983u;
  const uint32_t path = n_paths ? uint32_t((uint64_t(j) * stride) % n_paths) : 0u;
  bool ok = false;
  CacheQuery q{};
  V3 t = mk(0, 0, 0);
  if (j < n_paths) {
    const TrainPath tp = b.cache.tpaths[path];
    if (k < tp.nv) {
      const TrainVertex v = b.cache.tverts[size_t(path) * b.cache.max_vertices + k];
      q = v.q;
      t = mk((tp.T[0] - v.prefix[0]) / fmaxf(v.thr[0], 1e-4f), (tp.T[1] - v.prefix[1]) / fmaxf(v.thr[1], 1e-4f),
             (tp.T[2] - v.prefix[2]) / fmaxf(v.thr[2], 1e-4f));
      ok = t.x == t.x && t.y == t.y && t.z == t.z;
    }
  }
  if (blockIdx.x * kBlock >= b.cache.max_train_paths * b.cache.max_vertices) return;
  uint32_t* const queues[1] = {&b.counts[5]};
  const bool append[1] = {ok};
  uint32_t slots[1];
  block_append<kBlock, 1>(queues, append, slots);
  const uint32_t slot = slots[0];
  if (ok && slot < b.cache.max_batch) {
    b.cache.batch_q[slot] = q;
    float* o = b.cache.batch_t + 3 * size_t(slot);
    o[0] = fmaxf(t.x, 0.f);
    o[1] = fmaxf(t.y, 0.f);
    o[2] = fmaxf(t.z, 0.f);
  }
}

}  // namespace ptk

namespace ptk {
// Raw denoiser features of each pixel: mean radiance, albedo, normal, hit distance (inf if no
// sample hit a surface).
__global__ __launch_bounds__(kBlock) void denoise_input_kernel(const float4* accum, const float4* aux, float inv_spp,
                                                               float* raw, size_t npix) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= npix) return;
  const float4 c = accum[i], a0 = aux[2 * i], a1 = aux[2 * i + 1];
  float* o = raw + 10 * i;
  o[0] = c.x * inv_spp;
  o[1] = c.y * inv_spp;
  o[2] = c.z * inv_spp;
  o[3] = a0.x * inv_spp;
  o[4] = a0.y * inv_spp;
  o[5] = a0.z * inv_spp;
  o[6] = a1.x * inv_spp;
  o[7] = a1.y * inv_spp;
  o[8] = a1.z * inv_spp;
  o[9] = a1.w > 0.f ? a0.w / a1.w : __int_as_float(0x7f800000);
}
}  // namespace ptk

namespace ptk {
// Mean radiance of each pixel (running sum / sample count), alpha 1.
__global__ __launch_bounds__(kBlock) void mean_kernel(const float4* accum, float inv_spp, float4* out, size_t npix) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= npix) return;
  const float4 c = accum[i];
  out[i] = make_float4(c.x * inv_spp, c.y * inv_spp, c.z * inv_spp, 1.0f);
}
}  // namespace ptk

int launch_pt_mean(const float4* accum, uint32_t spp, float4* out, size_t npix, hipStream_t stream) {
  const unsigned grid = unsigned((npix + ptk::kBlock - 1) / ptk::kBlock);
  hipLaunchKernelGGL(ptk::mean_kernel, dim3(grid ? grid : 1), dim3(ptk::kBlock), 0, stream, accum,
                     1.0f / float(spp ? spp : 1), out, npix);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_denoise_input(const float4* accum, const float4* aux, uint32_t spp, float* raw, size_t npix,
                            hipStream_t stream) {
  const unsigned grid = unsigned((npix + ptk::kBlock - 1) / ptk::kBlock);
  hipLaunchKernelGGL(ptk::denoise_input_kernel, dim3(grid ? grid : 1), dim3(ptk::kBlock), 0, stream, accum, aux,
                     1.0f / float(spp ? spp : 1), raw, npix);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_generate(const Camera& cam, const PtBuffers& b, uint32_t cur, uint32_t n_paths, uint32_t sample_base,
                       hipStream_t stream) {
  const uint32_t grid = (n_paths + ptk::kBlock - 1) / ptk::kBlock;
  hipLaunchKernelGGL(ptk::generate_kernel, dim3(grid ? grid : 1), dim3(ptk::kBlock), 0, stream, cam, b, cur, n_paths,
                     sample_base);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_shade(const PtScene& sc, const PtBuffers& b, uint32_t cur, uint32_t max_depth, hipStream_t stream) {
  const uint32_t grid = (b.capacity + ptk::kShadeBlock - 1) / ptk::kShadeBlock;
  hipLaunchKernelGGL(ptk::shade_kernel, dim3(grid), dim3(ptk::kShadeBlock), 0, stream, sc, b, cur, max_depth);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_cache_resolve(const PtBuffers& b, hipStream_t stream) {
  const uint32_t grid = (b.cache.max_queries + ptk::kBlock - 1) / ptk::kBlock;
  hipLaunchKernelGGL(ptk::cache_resolve_kernel, dim3(grid ? grid : 1), dim3(ptk::kBlock), 0, stream, b);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_train_batch(const PtBuffers& b, hipStream_t stream) {
  const uint32_t n = b.cache.max_train_paths * b.cache.max_vertices;
  const uint32_t grid = (n + ptk::kBlock - 1) / ptk::kBlock;
  hipLaunchKernelGGL(ptk::train_batch_kernel, dim3(grid ? grid : 1), dim3(ptk::kBlock), 0, stream, b);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pt_resolve(const PtBuffers& b, hipStream_t stream) {
  const uint32_t grid = (b.capacity + ptk::kBlock - 1) / ptk::kBlock;
  hipLaunchKernelGGL(ptk::resolve_kernel, dim3(grid), dim3(ptk::kBlock), 0, stream, b);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
