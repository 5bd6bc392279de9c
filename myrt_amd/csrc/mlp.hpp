// Neural radiance cache: records and launchers (mlp.hip) shared with cache.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

// One cache query, 64 B: shading point, outgoing direction, normal, roughness and albedo.
struct CacheQuery {
  float pos[3];
  float roughness;
  float dir[3];
  float pad0;
  float nrm[3];
  float pad1;
  float albedo[3];
  float pad2;
};
static_assert(sizeof(CacheQuery) == 64, "cache query is 64 bytes");

struct CacheParams {
  uint16_t* w;       // bf16 [6][64][64] W[out][in]
  uint16_t* wt;      // bf16 [6][64][64] W^T[in][out]
  float* master;     // fp32 [6][64][64]
  float* m;          // Adam first moment
  float* v;          // Adam second moment
  float lo[3];       // position box of the encoding
  float inv_ext[3];  // 1 / box extent
};

constexpr int kCacheLayers = 6;
constexpr int kCacheWidth = 64;
constexpr int kCacheParamsPadded = kCacheLayers * kCacheWidth * kCacheWidth;  // 24576
constexpr int kCacheParams = 5 * 64 * 64 + 3 * 64;                            // 20672

// Inference workgroups that fit one CU at once: a grid of that many per CU is persistent (each
// workgroup loads the weights to LDS once and loops over its tiles).
int cache_infer_blocks_per_cu();
int launch_cache_infer(const CacheParams& P, const CacheQuery* q, uint32_t n, const uint32_t* n_dev, float* out,
                       int grid, hipStream_t s);
// Workgroups of a training launch (at most kCacheMaxTrainGroups); `partial` holds that many
// kCacheParamsPadded-float weight-gradient partials.
constexpr int kCacheMaxTrainGroups = 512;
int cache_train_groups();
int launch_cache_train(const CacheParams& P, const CacheQuery* q, const float* targets, uint32_t n, float grad_scale,
                       float* loss_tiles, float* partial, float lr, float beta1, float beta2, float eps, float bc1,
                       float bc2, float* grad_out, hipStream_t s);
int launch_cache_sync(const CacheParams& P, hipStream_t s);

}  // namespace mrt
