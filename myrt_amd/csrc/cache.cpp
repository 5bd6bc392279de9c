// Neural radiance cache host object and C ABI (kernels in mlp.hip).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/myrt.h"
#include "cache_internal.hpp"
#include "mlp.hpp"
#include "scene_internal.hpp"
#include "abi_guard.hpp"

using namespace mrt;

namespace {

template <typename T>
struct DevCopy {
  T* p = nullptr;
  hipError_t e = hipSuccess;
  DevCopy(const T* host, size_t n) {
    e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess && n) e = hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice);
  }
  explicit DevCopy(size_t n) { e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)); }
  ~DevCopy() {
    if (p) hipFree(p);
  }
};

}  // namespace

extern "C" {

int mrt_cache_create(mrt_cache* out, uint64_t seed) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output");
  auto* c = new Cache;
  const int rc = c->init(seed);
  if (rc != MRT_OK) {
    delete c;
    return rc;
  }
  *out = reinterpret_cast<mrt_cache>(c);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_cache_destroy(mrt_cache c) {
  MRT_GUARD_BEGIN
  delete reinterpret_cast<Cache*>(c);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_cache_set_bounds(mrt_cache ch, const float lo[3], const float hi[3]) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || !lo || !hi) return fail(MRT_ERR_INVALID, "null argument");
  for (int a = 0; a < 3; ++a) {
    if (!(hi[a] > lo[a])) return fail(MRT_ERR_INVALID, "empty cache box");
    c->P.lo[a] = lo[a];
    c->P.inv_ext[a] = 1.0f / (hi[a] - lo[a]);
  }
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_cache_set_adam(mrt_cache ch, float lr, float beta1, float beta2, float eps) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || !(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return fail(MRT_ERR_INVALID, "bad Adam parameters");
  c->lr = lr;
  c->beta1 = beta1;
  c->beta2 = beta2;
  c->eps = eps;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_cache_get_params(mrt_cache ch, float* out, size_t n) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || !out || n != size_t(kCacheParams)) return fail(MRT_ERR_INVALID, "need 20672 floats");
  return c->get(c->P.master, out);
  MRT_GUARD_END
}

int mrt_cache_set_params(mrt_cache ch, const float* in, size_t n) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || !in || n != size_t(kCacheParams)) return fail(MRT_ERR_INVALID, "need 20672 floats");
  return c->set_params(in);
  MRT_GUARD_END
}

int mrt_cache_infer_device(mrt_cache ch, const mrt_cache_query* d_q, size_t n, float* d_out, void* stream) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || (n && (!d_q || !d_out))) return fail(MRT_ERR_INVALID, "null argument");
  if (n == 0) return MRT_OK;
  if (n >= (1u << 31)) return fail(MRT_ERR_LIMIT, "too many queries");
  static_assert(sizeof(mrt_cache_query) == sizeof(CacheQuery), "query layouts differ");
  const int grid = int(std::min<size_t>(size_t(c->infer_grid), (n + 255) / 256));
  return launch_cache_infer(c->P, reinterpret_cast<const CacheQuery*>(d_q), uint32_t(n), nullptr, d_out, std::max(grid, 1),
                            static_cast<hipStream_t>(stream)) == 0
             ? MRT_OK
             : fail(MRT_ERR_HIP, "cache inference launch failed");
  MRT_GUARD_END
}

int mrt_cache_infer(mrt_cache ch, const mrt_cache_query* q, size_t n, float* out) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || (n && (!q || !out))) return fail(MRT_ERR_INVALID, "null argument");
  if (n == 0) return MRT_OK;
  DevCopy<mrt_cache_query> dq(q, n);
  DevCopy<float> dout(3 * n);
  if (dq.e != hipSuccess || dout.e != hipSuccess) return hip_fail(dq.e != hipSuccess ? dq.e : dout.e, "staging");
  int rc = mrt_cache_infer_device(ch, dq.p, n, dout.p, nullptr);
  if (rc != MRT_OK) return rc;
  hipError_t e = hipMemcpy(out, dout.p, 3 * n * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "reading cache outputs");
  MRT_GUARD_END
}

int mrt_cache_train_device(mrt_cache ch, const mrt_cache_query* d_q, const float* d_targets, size_t n, void* stream) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || (n && (!d_q || !d_targets))) return fail(MRT_ERR_INVALID, "null argument");
  return c->train(reinterpret_cast<const CacheQuery*>(d_q), d_targets, n, c->lr, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}

int mrt_cache_train(mrt_cache ch, const mrt_cache_query* q, const float* targets, size_t n, float* loss) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || (n && (!q || !targets))) return fail(MRT_ERR_INVALID, "null argument");
  if (n == 0) return MRT_OK;
  DevCopy<mrt_cache_query> dq(q, n);
  DevCopy<float> dt(targets, 3 * n);
  if (dq.e != hipSuccess || dt.e != hipSuccess) return hip_fail(dq.e != hipSuccess ? dq.e : dt.e, "staging");
  int rc = c->train(reinterpret_cast<const CacheQuery*>(dq.p), dt.p, n, c->lr, nullptr);
  if (rc == MRT_OK && loss) rc = c->read_loss(n, loss, nullptr);
  if (rc == MRT_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "cache training");
  }
  return rc;
  MRT_GUARD_END
}

int mrt_cache_gradient(mrt_cache ch, const mrt_cache_query* q, const float* targets, size_t n, float* grad,
                       float* loss) {
  MRT_GUARD_BEGIN
  Cache* c = reinterpret_cast<Cache*>(ch);
  if (!c || !grad || (n && (!q || !targets))) return fail(MRT_ERR_INVALID, "null argument");
  if (n == 0) return fail(MRT_ERR_INVALID, "empty batch");
  DevCopy<mrt_cache_query> dq(q, n);
  DevCopy<float> dt(targets, 3 * n);
  if (dq.e != hipSuccess || dt.e != hipSuccess) return hip_fail(dq.e != hipSuccess ? dq.e : dt.e, "staging");
  int rc = c->train(reinterpret_cast<const CacheQuery*>(dq.p), dt.p, n, 0.f, nullptr);
  if (rc == MRT_OK && loss) rc = c->read_loss(n, loss, nullptr);
  if (rc == MRT_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "cache gradient");
    rc = c->get(c->grad, grad);
  }
  return rc;
  MRT_GUARD_END
}

}  // extern "C"
