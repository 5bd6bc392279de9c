// Wavefront renderer host (the sequential_pathtracer of src/pathtracer/sequential_pathtracer.*)
// and the scene's material / light entry points.
//
// One render pass per sample batch: generate -> for each bounce {extend (closest hit), shade
// (NEE + BSDF sampling + compaction into the next queue), connect (any hit on shadow rays),
// resolve}. Queue sizes stay on the device between passes, like the reference's indirect
// dispatch (src/pt/pt_ray_filter.comp:58-62); the host issues launches only.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/myrt.h"
#include "cache_internal.hpp"
#include "denoiser_internal.hpp"
#include "matsort.hpp"
#include "pathtrace.hpp"
#include "scene_internal.hpp"
#include "abi_guard.hpp"

namespace mrt {
namespace {

// 4x4 column-major helpers (rnu::mat4 layout).
void mat_mul(const double a[16], const double b[16], double out[16]) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
      out[c * 4 + r] = s;
    }
}

bool mat_inv(const double m[16], double out[16]) {
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] -
           m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] -
           m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] -
           m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] -
           m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] +
           m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] -
            m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] +
           m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] -
           m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] +
            m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] -
            m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.0 || !std::isfinite(det)) return false;
  for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
  return true;
}

template <typename T>
int ensure(T*& p, size_t& cap, size_t count) {
  if (p && cap >= count) return MRT_OK;
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc renderer buffer");
  cap = count;
  return MRT_OK;
}

}  // namespace

struct Renderer {
  float view[16], proj[16];
  Camera cam{};
  int num_bounces = 10;      // sequential_pathtracer.hpp:134
  uint32_t sample_count = 0;
  uint32_t sample_base = 0;  // sampler index of the accumulation's first sample (mrt_renderer_set_sample_base)
  uint32_t width = 0, height = 0;
  size_t cap_paths = 0, cap_accum = 0;
  // The sample pipeline's queues: ray and path queues (current and next bounce), hits, shadow
  // rays and records, occlusion flags, queue counters, material-sort scratch, and the shadow
  // stream and events its passes run on.
  struct Pipe {
    void* rays[2] = {nullptr, nullptr};
    PathState* paths[2] = {nullptr, nullptr};
    void *hits = nullptr, *shadow_rays = nullptr;
    ShadowRecord* shadow = nullptr;
    uint8_t* occluded = nullptr;
    uint32_t* counts = nullptr;  // [0..1] queues, [2] shadow queue, [3..5] cache, [8..9] rays traced (u64),
                                 // [10..11] non-finite values rejected (u64)
    uint32_t *ms_keys = nullptr, *ms_perm = nullptr, *ms_counts = nullptr;
    size_t c_rays0 = 0, c_rays1 = 0, c_paths0 = 0, c_paths1 = 0, c_hits = 0, c_shr = 0, c_sh = 0, c_occ = 0, c_cnt = 0;
    size_t c_msk = 0, c_msp = 0, c_msc = 0;
    hipStream_t main = nullptr, shadow_stream = nullptr;  // main is the renderer's stream
    hipEvent_t ev_shaded = nullptr, ev_resolved = nullptr;
    ~Pipe() {
      for (void* p : {rays[0], rays[1], (void*)paths[0], (void*)paths[1], hits, shadow_rays, (void*)shadow,
                      (void*)occluded, (void*)counts, (void*)ms_keys, (void*)ms_perm, (void*)ms_counts})
        if (p) hipFree(p);
      for (hipEvent_t ev : {ev_shaded, ev_resolved})
        if (ev) hipEventDestroy(ev);
      if (shadow_stream) hipStreamDestroy(shadow_stream);
    }
    int ensure_buffers(size_t cap, bool sort) {
      int rc;
      if ((rc = ensure(reinterpret_cast<char*&>(rays[0]), c_rays0, cap * 32)) ||
          (rc = ensure(reinterpret_cast<char*&>(rays[1]), c_rays1, cap * 32)) || (rc = ensure(paths[0], c_paths0, cap)) ||
          (rc = ensure(paths[1], c_paths1, cap)) || (rc = ensure(reinterpret_cast<char*&>(hits), c_hits, cap * 16)) ||
          (rc = ensure(reinterpret_cast<char*&>(shadow_rays), c_shr, cap * 32)) || (rc = ensure(shadow, c_sh, cap)) ||
          (rc = ensure(occluded, c_occ, cap)) || (rc = ensure(counts, c_cnt, 16)))
        return rc;
      if (sort && ((rc = ensure(ms_keys, c_msk, cap)) || (rc = ensure(ms_perm, c_msp, cap)) ||
                   (rc = ensure(ms_counts, c_msc, kMatSortBins))))
        return rc;
      return MRT_OK;
    }
    int ensure_streams() {
      hipError_t e = hipSuccess;
      if (!shadow_stream) e = hipStreamCreateWithFlags(&shadow_stream, hipStreamNonBlocking);
      if (e == hipSuccess && !ev_shaded) e = hipEventCreateWithFlags(&ev_shaded, hipEventDisableTiming);
      if (e == hipSuccess && !ev_resolved) e = hipEventCreateWithFlags(&ev_resolved, hipEventDisableTiming);
      return e == hipSuccess ? MRT_OK : hip_fail(e, "creating render streams");
    }
  };
  Pipe pipe;
  float4* accum = nullptr;
  uint64_t* ctr_host = nullptr;  // pinned: rays traced and non-finite values rejected by a render
  float4* aux = nullptr;  // primary-hit feature sums, 2 per pixel
  size_t cap_aux = 0;
  float* raw = nullptr;   // denoiser input (10 floats per pixel)
  float* den = nullptr;   // denoiser output (rgba)
  size_t cap_raw = 0, cap_den = 0;
  hipStream_t stream = nullptr;
  mrt_render_stats stats{};
  // Radiance cache (optional).
  // Depth of field (sequential_pathtracer.hpp:131-139 defaults).
  bool dof = false;
  float lens[2] = {10.f, 10.f};
  float focus = 10.f;
  float4* bokeh = nullptr;
  uint32_t bokeh_w = 0, bokeh_h = 0;
  float4* cubemap = nullptr;  // set_cubemap: 6 faces of cube_size^2 rgba texels
  uint32_t cube_size = 0;
  Cache* cache = nullptr;
  int cache_mode = 0;  // 0 off, 1 train only, 2 train and end paths in the cache
  float train_fraction = 1.0f / 16.0f;
  int train_steps = 8;
  uint32_t max_batch = 1u << 18;
  float min_roughness = 0.25f;
  uint32_t min_bounce = 2;
  uint32_t train_end = 4;         // every path ends in the cache at this vertex (0: training paths never)
  float unbiased_fraction = 0.f;  // training paths exempt from train_end
  TrainPath* tpaths = nullptr;
  TrainVertex* tverts = nullptr;
  CacheQuery* queries = nullptr;
  CacheRecord* qrec = nullptr;
  float* qout = nullptr;
  CacheQuery* batch_q = nullptr;
  float* batch_t = nullptr;
  size_t c_tp = 0, c_tv = 0, c_q = 0, c_qr = 0, c_qo = 0, c_bq = 0, c_bt = 0;
  float last_loss = 0.f;
  // Shading in material order (off by default: with block-wide queue appends and pixel-major
  // paths, shading is about 4 % of the frame and the sort no longer pays for itself;
  // MRT_MATERIAL_SORT=1 or mrt_renderer_set_material_sort switches it on).
  bool material_sort = false;

  // Shadow rays of each bounce are traced and resolved on a second stream, concurrently with the
  // next bounce's closest-hit passes (MRT_SHADOW_STREAM=0: everything on one stream). Events keep
  // every data dependency: the shadow pass starts after the shade that wrote its rays, and the next
  // shade (which rewrites the shadow queue and adds to the same accumulators) waits for the resolve.
  bool shadow_overlap = true;
  // Paths in flight per batch (whole samples per pixel; at least one sample). Every launch has a
  // latency floor (the longest SDF march, the deepest traversal), so larger batches amortise it:
  // 1920x1080 at 16 spp renders 2.4x faster in one batch of 2^25 paths than in batches of 2^21.
  // The queues take about 230 B per path (7.5 GB at 2^25). MRT_BATCH_PATHS overrides.
  static constexpr size_t kCacheBatchPaths = size_t(1) << 21;
  uint32_t train_scale = 1;  // cache training of the current batch, in units of kCacheBatchPaths
  size_t max_batch_paths = size_t(1) << 25;

  Renderer() {
    if (const char* v = std::getenv("MRT_MATERIAL_SORT")) material_sort = std::strtol(v, nullptr, 10) != 0;
    if (const char* v = std::getenv("MRT_SHADOW_STREAM")) shadow_overlap = std::strtol(v, nullptr, 10) != 0;
    if (const char* v = std::getenv("MRT_BATCH_PATHS")) {
      const long long n = std::strtoll(v, nullptr, 10);
      if (n >= 1024 && n <= (1ll << 26)) max_batch_paths = size_t(n);
    }
    std::memset(view, 0, sizeof(view));
    std::memset(proj, 0, sizeof(proj));
    for (int i = 0; i < 4; ++i) view[i * 5] = proj[i * 5] = 1.f;
  }
  ~Renderer() {
    for (void* p : {(void*)accum, (void*)tpaths, (void*)tverts, (void*)queries, (void*)qrec, (void*)qout,
                    (void*)batch_q, (void*)batch_t, (void*)aux, (void*)raw, (void*)den, (void*)bokeh, (void*)cubemap})
      if (p) hipFree(p);
    if (ctr_host) hipHostFree(ctr_host);
  }

  int update_camera() {
    double v[16], pj[16], vinv[16], pinv[16], m[16];
    for (int i = 0; i < 16; ++i) {
      v[i] = view[i];
      pj[i] = proj[i];
    }
    if (!mat_inv(v, vinv) || !mat_inv(pj, pinv)) return fail(MRT_ERR_INVALID, "camera matrix is not invertible");
    mat_mul(vinv, pinv, m);
    for (int i = 0; i < 16; ++i) cam.inv_view_proj[i] = float(m[i]);
    cam.eye[0] = float(vinv[12]);
    cam.eye[1] = float(vinv[13]);
    cam.eye[2] = float(vinv[14]);
    cam.dof = dof ? 1u : 0u;
    cam.lens[0] = lens[0];
    cam.lens[1] = lens[1];
    cam.focus = focus;
    cam.bokeh = bokeh;
    cam.bokeh_w = bokeh_w;
    cam.bokeh_h = bokeh_h;
    return MRT_OK;
  }

  int render(Scene& s, uint32_t w, uint32_t h, uint32_t spp) {
    int rc = update_camera();
    if (rc != MRT_OK) return rc;
    if (w == 0 || h == 0) return fail(MRT_ERR_INVALID, "empty image");
    if (uint64_t(w) * h > (1u << 26)) return fail(MRT_ERR_LIMIT, "image larger than 64M pixels");
    // scene.prepare() at the start of every run (sequential_pathtracer.cpp:45); a change
    // restarts the progressive accumulation (:47-53).
    mrt_prepare_result pr{};
    const bool changed_before = s.geoms_dirty || s.inst_dirty || !s.pending.empty() || s.shading_dirty;
    rc = mrt_scene_prepare(reinterpret_cast<mrt_scene>(&s), &pr);
    if (rc != MRT_OK) return rc;
    const size_t npix = size_t(w) * h;
    if (w != width || h != height || changed_before) {
      width = w;
      height = h;
      sample_count = 0;
    }
    const bool use_cache = cache && cache_mode > 0;
    // Samples per batch: at most max_batch_paths paths in flight.
    const uint32_t batch = uint32_t(std::max<size_t>(1, std::min<size_t>(spp, max_batch_paths / npix)));
    const size_t cap = npix * batch;
    // The radiance cache trains after every batch; its cadence was set for batches of
    // kCacheBatchPaths paths. A larger batch trains as that many of them would have: proportionally
    // more samples and steps, each step the same size.
    const size_t cadence_spp = std::max<size_t>(1, kCacheBatchPaths / npix);
    const uint32_t batch_cap = max_batch * uint32_t((batch + cadence_spp - 1) / cadence_spp);
    if ((rc = pipe.ensure_buffers(cap, material_sort)) != MRT_OK) return rc;
    if (!accum || cap_accum < npix) {
      if ((rc = ensure(accum, cap_accum, npix)) != MRT_OK) return rc;
      if ((rc = ensure(aux, cap_aux, 2 * npix)) != MRT_OK) return rc;
      sample_count = 0;
    }
    if (sample_count == 0) {
      hipError_t e = hipMemsetAsync(accum, 0, npix * sizeof(float4), stream);
      if (e == hipSuccess) e = hipMemsetAsync(aux, 0, 2 * npix * sizeof(float4), stream);
      if (e != hipSuccess) return hip_fail(e, "clearing the accumulator");
    }
    PtCache pc{};
    if (use_cache) {
      const uint32_t depth = uint32_t(num_bounces);
      const uint32_t group = uint32_t(std::max(1.0, std::round(1.0 / double(train_fraction))));
      const size_t max_train = (cap + group - 1) / group;
      if ((rc = ensure(tpaths, c_tp, max_train)) || (rc = ensure(tverts, c_tv, max_train * depth)) ||
          (rc = ensure(queries, c_q, cap)) || (rc = ensure(qrec, c_qr, cap)) || (rc = ensure(qout, c_qo, 3 * cap)) ||
          (rc = ensure(batch_q, c_bq, batch_cap)) || (rc = ensure(batch_t, c_bt, 3 * size_t(batch_cap))))
        return rc;
      pc.terminate = cache_mode >= 2 ? 1u : 0u;
      pc.train = 1u;
      pc.train_group = group;
      pc.max_train_paths = uint32_t(max_train);
      pc.max_vertices = depth;
      pc.max_queries = uint32_t(cap);
      pc.max_batch = batch_cap;
      pc.min_roughness = min_roughness;
      pc.min_bounce = min_bounce;
      pc.train_end = train_end;
      pc.unbiased_fraction = unbiased_fraction;
      pc.tpaths = tpaths;
      pc.tverts = tverts;
      pc.queries = queries;
      pc.qrec = qrec;
      pc.qout = qout;
      pc.batch_q = batch_q;
      pc.batch_t = batch_t;
    }
    cam.width = w;
    cam.height = h;
    PtScene sc = s.shading_view();
    sc.cubemap = cubemap;
    sc.cube_size = cube_size;
    const auto t0 = std::chrono::steady_clock::now();
    pipe.main = stream;
    if (shadow_overlap && (rc = pipe.ensure_streams()) != MRT_OK) return rc;
    hipError_t e = hipMemsetAsync(pipe.counts + 8, 0, 16, stream);
    if (e != hipSuccess) return hip_fail(e, "clearing the render counters");
    for (uint32_t done = 0; done < spp;) {
      const uint32_t nb = std::min(batch, spp - done);
      if ((rc = run_batch(s, sc, pipe, nb, npix, sample_base + sample_count + done, pc, cadence_spp)) != MRT_OK)
        return rc;
      done += nb;
    }
    // Rays traced and non-finite values rejected: copied on the stream, and read_ctrl's one
    // synchronisation of the same stream covers this copy too.
    if (!ctr_host && (e = hipHostMalloc(reinterpret_cast<void**>(&ctr_host), 16, hipHostMallocDefault)) != hipSuccess) {
      ctr_host = nullptr;
      return hip_fail(e, "hipHostMalloc render counter staging");
    }
    e = hipMemcpyAsync(ctr_host, pipe.counts + 8, 16, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return hip_fail(e, "render");
    uint32_t flags = 0;
    if ((rc = s.read_ctrl(&flags, nullptr, stream)) != MRT_OK) return rc;
    const uint64_t ctr[2] = {ctr_host[0], ctr_host[1]};
    const uint64_t rays_traced = ctr[0];
    if (flags & 1u) return fail(MRT_ERR_INTERNAL, "traversal stack overflow");
    const auto t1 = std::chrono::steady_clock::now();
    sample_count += spp;
    stats.samples += uint64_t(spp) * npix;
    stats.rays += rays_traced;
    stats.nonfinite += ctr[1];
    stats.milliseconds += std::chrono::duration<double, std::milli>(t1 - t0).count();
    return MRT_OK;
  }

  // One batch of nb samples per pixel: generate, then per bounce closest hits, material sort,
  // shade, and the shadow pass (on the shadow stream when overlapped), then the cache pass. Trace
  // contexts 0 (closest) and 1 (shadow).
  int run_batch(Scene& s, const PtScene& sc, Pipe& p, uint32_t nb, size_t npix, uint32_t sample_base,
                const PtCache& pc, size_t cadence_spp) {
    int rc;
    const uint32_t n_paths = uint32_t(npix * nb);
    PtBuffers b{};
    b.rays[0] = p.rays[0];
    b.rays[1] = p.rays[1];
    b.paths[0] = p.paths[0];
    b.paths[1] = p.paths[1];
    b.hits = p.hits;
    b.shadow_rays = p.shadow_rays;
    b.shadow = p.shadow;
    b.occluded = p.occluded;
    b.counts = p.counts;
    b.accum = accum;
    b.aux = aux;
    b.cache = pc;
    b.capacity = n_paths;
    train_scale = uint32_t((nb + cadence_spp - 1) / cadence_spp);
    b.cache.max_batch = max_batch * train_scale;
    hipStream_t st = p.main;
    const bool overlap = shadow_overlap;
    hipStream_t sh_stream = overlap ? p.shadow_stream : st;
    const int ctx_closest = 0, ctx_shadow = overlap ? 1 : 0;
    uint32_t* counts = p.counts;
    uint32_t cur = 0;
    hipError_t e0 = hipMemsetAsync(counts + 3, 0, 3 * 4, st);
    if (e0 != hipSuccess) return hip_fail(e0, "clearing cache counters");
    if (launch_pt_generate(cam, b, cur, n_paths, sample_base, st)) return fail(MRT_ERR_HIP, "generate");
    bool shadow_pending = false;
    for (int bounce = 0; bounce < num_bounces; ++bounce) {
      // counts[cur] is live; clear the next queue.
      hipError_t e = hipMemsetAsync(counts + (cur ^ 1u), 0, 4, st);
      if (e != hipSuccess) return hip_fail(e, "clearing queue counters");
      if ((rc = s.trace(p.rays[cur], n_paths, p.hits, false, false, st, counts + cur, ctx_closest)) != MRT_OK) return rc;
      // Primary hits arrive grouped by screen position (paths are pixel-major), hence by surface:
      // sorting pays from the first scattered bounce on.
      b.perm = nullptr;
      if (material_sort && bounce > 0) {
        MatSortArgs ma{p.hits, counts + cur, n_paths, sc.inst_material, p.ms_keys, p.ms_counts, p.ms_perm};
        if (launch_material_sort(ma, st)) return fail(MRT_ERR_HIP, "material sort");
        b.perm = p.ms_perm;
      }
      // The previous bounce's shadow pass has to be done with the shadow queue and the
      // accumulators before this shade rewrites them.
      if (shadow_pending && (e = hipStreamWaitEvent(st, p.ev_resolved, 0)) != hipSuccess)
        return hip_fail(e, "waiting for the shadow pass");
      if ((e = hipMemsetAsync(counts + 2, 0, 4, st)) != hipSuccess) return hip_fail(e, "clearing the shadow queue");
      if (launch_pt_shade(sc, b, cur, uint32_t(num_bounces), st)) return fail(MRT_ERR_HIP, "shade");
      if (overlap) {
        if ((e = hipEventRecord(p.ev_shaded, st)) != hipSuccess || (e = hipStreamWaitEvent(sh_stream, p.ev_shaded, 0)) != hipSuccess)
          return hip_fail(e, "ordering the shadow pass");
      }
      if ((rc = s.trace(p.shadow_rays, n_paths, p.occluded, true, false, sh_stream, counts + 2, ctx_shadow)) != MRT_OK)
        return rc;
      if (launch_pt_resolve(b, sh_stream)) return fail(MRT_ERR_HIP, "resolve");
      if (overlap) {
        if ((e = hipEventRecord(p.ev_resolved, sh_stream)) != hipSuccess) return hip_fail(e, "ordering the shadow pass");
        shadow_pending = true;
      }
      cur ^= 1u;
    }
    if (shadow_pending) {
      hipError_t e = hipStreamWaitEvent(st, p.ev_resolved, 0);
      if (e != hipSuccess) return hip_fail(e, "waiting for the shadow pass");
    }
    if (b.cache.train) {
      if ((rc = cache_pass(b, st)) != MRT_OK) return rc;
    }
    return MRT_OK;
  }

  // After a pass: resolve the paths that ended in the cache with its current weights, then build
  // this pass's training batch from the training paths and take train_steps Adam steps on it.
  int cache_pass(const PtBuffers& b, hipStream_t st) {
    if (b.cache.terminate) {
      const int grid = std::max(1, std::min(cache->infer_grid, int((b.cache.max_queries + 255) / 256)));
      if (launch_cache_infer(cache->P, queries, b.cache.max_queries, b.counts + 4, qout, grid, st))
        return fail(MRT_ERR_HIP, "cache inference");
      if (launch_pt_cache_resolve(b, st)) return fail(MRT_ERR_HIP, "cache resolve");
    }
    if (launch_pt_train_batch(b, st)) return fail(MRT_ERR_HIP, "training batch");
    uint32_t n_batch = 0;
    hipError_t e = hipMemcpyAsync(&n_batch, b.counts + 5, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "reading the training batch size");
    n_batch = std::min(n_batch, b.cache.max_batch);
    if (n_batch < 64) return MRT_OK;
    // train_steps steps per kCacheBatchPaths paths rendered, on disjoint slices of the batch.
    const uint32_t steps = uint32_t(std::max(train_steps, 1)) * train_scale;
    const uint32_t per = std::max<uint32_t>(64, n_batch / steps);
    for (uint32_t off = 0; off + per <= n_batch; off += per) {
      int rc = cache->train(batch_q + off, batch_t + 3 * size_t(off), per, cache->lr, st);
      if (rc != MRT_OK) return rc;
    }
    return MRT_OK;
  }

  int read_aux(float* out) {
    if (!aux || sample_count == 0) return fail(MRT_ERR_INVALID, "nothing rendered yet");
    const size_t npix = size_t(width) * height;
    std::vector<float4> host(2 * npix);
    hipError_t e = hipMemcpy(host.data(), aux, 2 * npix * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the features");
    const float inv = 1.0f / float(sample_count);
    for (size_t i = 0; i < npix; ++i) {
      const float4 a0 = host[2 * i], a1 = host[2 * i + 1];
      float* o = out + 8 * i;
      o[0] = a0.x * inv;
      o[1] = a0.y * inv;
      o[2] = a0.z * inv;
      o[3] = a1.x * inv;
      o[4] = a1.y * inv;
      o[5] = a1.z * inv;
      o[6] = a1.w > 0.f ? a0.w / a1.w : INFINITY;
      o[7] = 0.f;
    }
    return MRT_OK;
  }

  int denoise(Denoiser& d, float* rgba) {
    if (!accum || sample_count == 0) return fail(MRT_ERR_INVALID, "nothing rendered yet");
    const size_t npix = size_t(width) * height;
    int rc;
    if ((rc = ensure(raw, cap_raw, 10 * npix)) || (rc = ensure(den, cap_den, 4 * npix))) return rc;
    if (launch_pt_denoise_input(accum, aux, sample_count, raw, npix, stream))
      return fail(MRT_ERR_HIP, "denoiser input");
    if ((rc = d.denoise(raw, int(height), int(width), den, stream)) != MRT_OK) return rc;
    hipError_t e = hipMemcpyAsync(rgba, den, npix * 16, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e == hipSuccess ? MRT_OK : hip_fail(e, "reading the denoised image");
  }

  // The mean image into a device buffer, asynchronously on `st` (null: the renderer's stream).
  int read_device(float* d_rgba, hipStream_t st) {
    if (!accum || sample_count == 0) return fail(MRT_ERR_INVALID, "nothing rendered yet");
    const size_t npix = size_t(width) * height;
    if (launch_pt_mean(accum, sample_count, reinterpret_cast<float4*>(d_rgba), npix, st ? st : stream))
      return fail(MRT_ERR_HIP, "mean image");
    return MRT_OK;
  }

  int read(float* rgba) {
    if (!accum || sample_count == 0) return fail(MRT_ERR_INVALID, "nothing rendered yet");
    const size_t npix = size_t(width) * height;
    std::vector<float4> host(npix);
    hipError_t e = hipMemcpy(host.data(), accum, npix * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the accumulator");
    const float inv = 1.0f / float(sample_count);
    for (size_t i = 0; i < npix; ++i) {
      rgba[4 * i + 0] = host[i].x * inv;
      rgba[4 * i + 1] = host[i].y * inv;
      rgba[4 * i + 2] = host[i].z * inv;
      rgba[4 * i + 3] = 1.0f;
    }
    return MRT_OK;
  }
};

}  // namespace mrt

using namespace mrt;


extern "C" {

int mrt_scene_push_texture(mrt_scene sh, const float* rgba, int w, int h, uint32_t* out) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || !rgba || w < 1 || h < 1 || size_t(w) * h > (size_t(1) << 26)) return fail(MRT_ERR_INVALID, "bad texture");
  if (s->texels.size() + size_t(w) * h >= (size_t(1) << 31)) return fail(MRT_ERR_LIMIT, "texture memory exceeds 2^31 texels");
  TexInfo t{uint32_t(s->texels.size()), uint32_t(w), uint32_t(h), 0};
  const float4* px = reinterpret_cast<const float4*>(rgba);
  s->texels.insert(s->texels.end(), px, px + size_t(w) * h);
  s->textures.push_back(t);
  s->shading_dirty = true;
  if (out) *out = uint32_t(s->textures.size());  // 0 means "no texture"
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_lights(mrt_scene sh, const mrt_quad_light* lights, size_t n) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s || (n && !lights)) return fail(MRT_ERR_INVALID, "null argument");
  static_assert(sizeof(mrt_quad_light) == sizeof(QuadLight), "light layouts differ");
  s->lights.assign(reinterpret_cast<const QuadLight*>(lights), reinterpret_cast<const QuadLight*>(lights) + n);
  s->shading_dirty = true;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_scene_set_environment(mrt_scene sh, float scale) {
  MRT_GUARD_BEGIN
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!s) return fail(MRT_ERR_INVALID, "null scene");
  s->env_scale = scale;
  s->shading_dirty = true;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_create(mrt_renderer* out) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output");
  *out = reinterpret_cast<mrt_renderer>(new Renderer);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_destroy(mrt_renderer r) {
  MRT_GUARD_BEGIN
  delete reinterpret_cast<Renderer*>(r);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_camera(mrt_renderer rh, const float view_colmajor[16], const float proj_colmajor[16]) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !view_colmajor || !proj_colmajor) return fail(MRT_ERR_INVALID, "null argument");
  std::memcpy(r->view, view_colmajor, sizeof(r->view));
  std::memcpy(r->proj, proj_colmajor, sizeof(r->proj));
  r->sample_count = 0;  // every setter invalidates the counter (sequential_pathtracer.cpp:70-144)
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_dof_enabled(mrt_renderer rh, int enabled) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r) return fail(MRT_ERR_INVALID, "null renderer");
  if (r->dof != (enabled != 0)) r->sample_count = 0;
  r->dof = enabled != 0;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_lens_size(mrt_renderer rh, float w, float h) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !std::isfinite(w) || !std::isfinite(h)) return fail(MRT_ERR_INVALID, "bad lens size");
  if (w != r->lens[0] || h != r->lens[1]) r->sample_count = 0;
  r->lens[0] = w;
  r->lens[1] = h;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_focus(mrt_renderer rh, float focus) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !(focus > 0.f) || !std::isfinite(focus)) return fail(MRT_ERR_INVALID, "focus must be positive");
  if (focus != r->focus) r->sample_count = 0;
  r->focus = focus;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_bokeh_mask(mrt_renderer rh, const float* rgba, int w, int h) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || w < 0 || h < 0 || (w > 0 && (!rgba || h == 0))) return fail(MRT_ERR_INVALID, "bad bokeh mask");
  if (r->bokeh) hipFree(r->bokeh);
  r->bokeh = nullptr;
  r->bokeh_w = r->bokeh_h = 0;
  r->sample_count = 0;
  // Like bokeh_get (bokeh.glsl:24-25), masks narrower than 10 texels mean "no mask".
  if (w < 10) return MRT_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&r->bokeh), size_t(w) * h * sizeof(float4));
  if (e == hipSuccess) e = hipMemcpy(r->bokeh, rgba, size_t(w) * h * sizeof(float4), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "uploading the bokeh mask");
  r->bokeh_w = uint32_t(w);
  r->bokeh_h = uint32_t(h);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_cubemap(mrt_renderer rh, const float* faces, int size) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || size < 0 || (size > 0 && !faces) || size > 16384) return fail(MRT_ERR_INVALID, "bad cube map");
  if (r->cubemap) hipFree(r->cubemap);
  r->cubemap = nullptr;
  r->cube_size = 0;
  r->sample_count = 0;
  // Like pt_color.comp:192, maps of at most 10 texels a side mean "use the sky gradient".
  if (size <= 10) return MRT_OK;
  const size_t n = size_t(6) * size * size;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&r->cubemap), n * sizeof(float4));
  if (e == hipSuccess) e = hipMemcpy(r->cubemap, faces, n * sizeof(float4), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "uploading the cube map");
  r->cube_size = uint32_t(size);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_num_bounces(mrt_renderer rh, int n) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || n < 1 || n > 64) return fail(MRT_ERR_INVALID, "bounces must be in [1, 64]");
  r->num_bounces = n;
  r->sample_count = 0;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_material_sort(mrt_renderer rh, int enabled) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r) return fail(MRT_ERR_INVALID, "null renderer");
  r->material_sort = enabled != 0;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_cache(mrt_renderer rh, mrt_cache c, int mode, float train_fraction, int train_steps,
                           int min_bounce) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || mode < 0 || mode > 2 || (mode > 0 && !c) || !(train_fraction > 0.f && train_fraction <= 1.f) ||
      train_steps < 1 || train_steps > 64 || min_bounce < 1 || min_bounce > 16)
    return fail(MRT_ERR_INVALID, "bad cache settings");
  r->min_bounce = uint32_t(min_bounce);
  r->cache = reinterpret_cast<Cache*>(c);
  r->cache_mode = mode;
  r->train_fraction = train_fraction;
  r->train_steps = train_steps;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_cache_training(mrt_renderer rh, int train_end, float unbiased_fraction) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || train_end < 0 || train_end > 16 || !(unbiased_fraction >= 0.f && unbiased_fraction <= 1.f))
    return fail(MRT_ERR_INVALID, "bad cache training settings");
  r->train_end = uint32_t(train_end);
  r->unbiased_fraction = unbiased_fraction;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_invalidate(mrt_renderer rh) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r) return fail(MRT_ERR_INVALID, "null renderer");
  r->sample_count = 0;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_set_sample_base(mrt_renderer rh, uint32_t base) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r) return fail(MRT_ERR_INVALID, "null renderer");
  r->sample_base = base;
  r->sample_count = 0;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_renderer_sample_count(mrt_renderer rh, uint32_t* out) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !out) return fail(MRT_ERR_INVALID, "null argument");
  *out = r->sample_count;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_render(mrt_renderer rh, mrt_scene sh, int width, int height, int spp) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  Scene* s = reinterpret_cast<Scene*>(sh);
  if (!r || !s || width <= 0 || height <= 0 || spp <= 0) return fail(MRT_ERR_INVALID, "bad render arguments");
  return r->render(*s, uint32_t(width), uint32_t(height), uint32_t(spp));
  MRT_GUARD_END
}

int mrt_renderer_read(mrt_renderer rh, float* rgba) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !rgba) return fail(MRT_ERR_INVALID, "null argument");
  return r->read(rgba);
  MRT_GUARD_END
}

int mrt_renderer_read_device(mrt_renderer rh, float* d_rgba, void* stream) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !d_rgba) return fail(MRT_ERR_INVALID, "null argument");
  return r->read_device(d_rgba, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}

int mrt_renderer_read_aux(mrt_renderer rh, float* out) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !out) return fail(MRT_ERR_INVALID, "null argument");
  return r->read_aux(out);
  MRT_GUARD_END
}

int mrt_renderer_denoise(mrt_renderer rh, mrt_denoiser dh, float* rgba) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !dh || !rgba) return fail(MRT_ERR_INVALID, "null argument");
  return r->denoise(*reinterpret_cast<Denoiser*>(dh), rgba);
  MRT_GUARD_END
}

int mrt_renderer_stats(mrt_renderer rh, mrt_render_stats* out) {
  MRT_GUARD_BEGIN
  Renderer* r = reinterpret_cast<Renderer*>(rh);
  if (!r || !out) return fail(MRT_ERR_INVALID, "null argument");
  *out = r->stats;
  return MRT_OK;
  MRT_GUARD_END
}

}  // extern "C"
