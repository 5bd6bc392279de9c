// Radiance cache state behind the mrt_cache handle (kernels in mlp.hip, C ABI in cache.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/myrt.h"
#include "mlp.hpp"
#include "scene_internal.hpp"

namespace mrt {

struct Cache {
  CacheParams P{};
  float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
  uint32_t step = 0;
  float *loss_tiles = nullptr, *partial = nullptr, *grad = nullptr;
  size_t cap_batch = 0;
  int infer_grid = 1024;

  ~Cache() {
    for (void* p : {(void*)P.w, (void*)P.wt, (void*)P.master, (void*)P.m, (void*)P.v, (void*)loss_tiles,
                    (void*)partial, (void*)grad})
      if (p) hipFree(p);
  }

  int init(uint64_t seed) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&P.w), kCacheParamsPadded * 2);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P.wt), kCacheParamsPadded * 2);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P.master), kCacheParamsPadded * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P.m), kCacheParamsPadded * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P.v), kCacheParamsPadded * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&grad), kCacheParamsPadded * 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc cache parameters");
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
      infer_grid = prop.multiProcessorCount * cache_infer_blocks_per_cu();
    for (int a = 0; a < 3; ++a) {
      P.lo[a] = -1.f;
      P.inv_ext[a] = 0.5f;
    }
    // He-uniform initialisation from a splitmix64 stream.
    std::vector<float> w(kCacheParams);
    uint64_t x = seed;
    size_t k = 0;
    for (int l = 0; l < kCacheLayers; ++l) {
      const int rows = l == kCacheLayers - 1 ? 3 : kCacheWidth;
      const float bound = std::sqrt(6.0f / float(kCacheWidth));
      for (int i = 0; i < rows * kCacheWidth; ++i) {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const float u = float(z >> 40) * (1.0f / 16777216.0f);
        w[k++] = (2.0f * u - 1.0f) * bound;
      }
    }
    return set_params(w.data());
  }

  static size_t padded_index(size_t packed) {
    const size_t l = packed / (kCacheWidth * kCacheWidth);
    return l < 5 ? packed : 5 * kCacheWidth * kCacheWidth + (packed - 5 * kCacheWidth * kCacheWidth);
  }

  int set_params(const float* packed) {
    std::vector<float> pad(kCacheParamsPadded, 0.f);
    for (size_t i = 0; i < size_t(kCacheParams); ++i) pad[padded_index(i)] = packed[i];
    hipError_t e = hipMemcpy(P.master, pad.data(), pad.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(P.m, 0, kCacheParamsPadded * 4);
    if (e == hipSuccess) e = hipMemset(P.v, 0, kCacheParamsPadded * 4);
    if (e != hipSuccess) return hip_fail(e, "uploading cache parameters");
    step = 0;
    if (launch_cache_sync(P, nullptr)) return fail(MRT_ERR_HIP, "cache weight conversion");
    e = hipDeviceSynchronize();
    return e == hipSuccess ? MRT_OK : hip_fail(e, "cache weight conversion");
  }

  int get(float* dev_padded, float* packed) {
    std::vector<float> pad(kCacheParamsPadded);
    hipError_t e = hipMemcpy(pad.data(), dev_padded, pad.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading cache parameters");
    for (size_t i = 0; i < size_t(kCacheParams); ++i) packed[i] = pad[padded_index(i)];
    return MRT_OK;
  }

  int ensure_batch(size_t n) {
    const size_t n_pad = (n + 31) / 32 * 32;
    if (n_pad <= cap_batch) return MRT_OK;
    if (loss_tiles) hipFree(loss_tiles);
    loss_tiles = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&loss_tiles), n_pad / 32 * 4);
    if (e == hipSuccess && !partial)
      e = hipMalloc(reinterpret_cast<void**>(&partial), size_t(cache_train_groups()) * kCacheParamsPadded * 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc cache training buffers");
    cap_batch = n_pad;
    return MRT_OK;
  }

  // One optimiser step (lr_override < 0 keeps the configured rate; 0 computes the gradient only).
  int train(const CacheQuery* d_q, const float* d_t, size_t n, float lr_used, hipStream_t s) {
    if (n == 0) return MRT_OK;
    if (n >= (1u << 26)) return fail(MRT_ERR_LIMIT, "at most 2^26 training samples per step");
    int rc = ensure_batch(n);
    if (rc != MRT_OK) return rc;
    float bc1 = 1.f, bc2 = 1.f;
    if (lr_used != 0.f) {
      ++step;
      bc1 = 1.0f - std::pow(beta1, float(step));
      bc2 = 1.0f - std::pow(beta2, float(step));
    }
    const float scale = 1.0f / float(3 * n);
    if (launch_cache_train(P, d_q, d_t, uint32_t(n), scale, loss_tiles, partial, lr_used, beta1, beta2, eps, bc1, bc2,
                           grad, s))
      return fail(MRT_ERR_HIP, "cache training launch failed");
    return MRT_OK;
  }

  int read_loss(size_t n, float* loss, hipStream_t s) {
    const size_t tiles = (n + 31) / 32;
    std::vector<float> lt(tiles);
    hipError_t e = hipMemcpyAsync(lt.data(), loss_tiles, tiles * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "reading the loss");
    double sum = 0;
    for (float v : lt) sum += v;
    *loss = float(sum / double(3 * n));
    return MRT_OK;
  }
};

}  // namespace mrt
