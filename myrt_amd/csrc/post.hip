// Post-process chain after the accumulator (src/pt/pp_denoise.comp, pp_tonemap_filmic.cl,
// pp_fxaa.comp): a bilateral filter, a filmic tone curve and FXAA, one thread per pixel over
// rgba float images. Each pass follows the reference pass's formula; images are read with
// clamped coordinates (texelFetch) or bilinear filtering clamped to the edge (texture()).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/myrt.h"
#include "scene_internal.hpp"
#include "abi_guard.hpp"

namespace mrt {
namespace ppk {

constexpr int kBlock = 256;

__device__ inline float normpdf(float x, float sigma) { return 0.39894f * expf(-0.5f * x * x / (sigma * sigma)) / sigma; }

// pp_denoise.comp:24-63: 9x9 window, spatial sigma `exponent`, range sigma `strength`; the
// neighbours are clamped to [0, 1], the centre is not. The range weights are taken relative to
// the window's smallest colour distance q_min, exp(-(q - q_min) / (2 sigma_r^2)): the common
// factor cancels in the normalised sum, so the result is the reference's wherever its fp32
// weights do not all underflow, and stays defined where they do (a centre far above every
// clamped neighbour, where the reference divides 0 by 0).
__global__ __launch_bounds__(kBlock) void bilateral_kernel(const float4* in, float4* out, int w, int h, float sigma,
                                                           float bsigma) {
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= w * h) return;
  const int x = i % w, y = i / w;
  constexpr int kSize = 4;
  float kernel[2 * kSize + 1];
  for (int j = 0; j <= kSize; ++j) kernel[kSize + j] = kernel[kSize - j] = normpdf(float(j), sigma);
  const float4 c = in[i];
  auto neighbour = [&](int di, int dj) {
    const int sx = min(max(x + di, 0), w - 1), sy = min(max(y + dj, 0), h - 1);
    const float4 s = in[sy * w + sx];
    return make_float3(fminf(fmaxf(s.x, 0.f), 1.f), fminf(fmaxf(s.y, 0.f), 1.f), fminf(fmaxf(s.z, 0.f), 1.f));
  };
  auto dist2 = [&](float3 s) {
    const float dr = s.x - c.x, dg = s.y - c.y, db = s.z - c.z;
    return dr * dr + dg * dg + db * db;
  };
  float qmin = INFINITY;
  for (int di = -kSize; di <= kSize; ++di)
    for (int dj = -kSize; dj <= kSize; ++dj) qmin = fminf(qmin, dist2(neighbour(di, dj)));
  const float inv2s2 = 0.5f / (bsigma * bsigma);
  float z = 0.f, fr = 0.f, fg = 0.f, fb = 0.f;
  for (int di = -kSize; di <= kSize; ++di) {
    for (int dj = -kSize; dj <= kSize; ++dj) {
      const float3 s = neighbour(di, dj);
      const float f = expf(-(dist2(s) - qmin) * inv2s2) * kernel[kSize + dj] * kernel[kSize + di];
      z += f;
      fr += f * s.x;
      fg += f * s.y;
      fb += f * s.z;
    }
  }
  out[i] = make_float4(fr / z, fg / z, fb / z, 1.0f);
}

__device__ inline float filmic(float x) {
  const float X = fmaxf(0.0f, x - 0.004f);
  const float r = (X * (6.2f * X + 0.5f)) / (X * (6.2f * X + 1.7f) + 0.06f);
  return powf(r, 1.3f);
}

// pp_tonemap_filmic.cl:1-16.
__global__ __launch_bounds__(kBlock) void tonemap_kernel(const float4* in, float4* out, int n) {
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n) return;
  const float4 c = in[i];
  out[i] = make_float4(filmic(c.x), filmic(c.y), filmic(c.z), 1.0f);
}

// Bilinear rgb at normalised (u, v), texel centres at (i + 0.5) / size, clamped to the edge.
__device__ inline float3 tex(const float4* img, int w, int h, float u, float v) {
  const float fx = fminf(fmaxf(u * float(w) - 0.5f, 0.f), float(w - 1));
  const float fy = fminf(fmaxf(v * float(h) - 0.5f, 0.f), float(h - 1));
  const int x0 = int(fx), y0 = int(fy), x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
  const float tx = fx - float(x0), ty = fy - float(y0);
  const float4 a = img[y0 * w + x0], b = img[y0 * w + x1], c = img[y1 * w + x0], d = img[y1 * w + x1];
  const float w00 = (1.f - tx) * (1.f - ty), w10 = tx * (1.f - ty), w01 = (1.f - tx) * ty, w11 = tx * ty;
  return make_float3(a.x * w00 + b.x * w10 + c.x * w01 + d.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + d.y * w11,
                     a.z * w00 + b.z * w10 + c.z * w01 + d.z * w11);
}

__device__ inline float luma(float3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }

// pp_fxaa.comp:31-105 (FXAA with 4 corner taps); the fragment coordinate is the integer pixel
// position, as in the reference's main().
__global__ __launch_bounds__(kBlock) void fxaa_kernel(const float4* in, float4* out, int w, int h) {
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= w * h) return;
  const float px = float(i % w), py = float(i / w);
  const float ivx = 1.0f / float(w), ivy = 1.0f / float(h);
  const float3 nw = tex(in, w, h, (px - 1.f) * ivx, (py - 1.f) * ivy);
  const float3 ne = tex(in, w, h, (px + 1.f) * ivx, (py - 1.f) * ivy);
  const float3 sw = tex(in, w, h, (px - 1.f) * ivx, (py + 1.f) * ivy);
  const float3 se = tex(in, w, h, (px + 1.f) * ivx, (py + 1.f) * ivy);
  const float3 m = tex(in, w, h, px * ivx, py * ivy);
  const float lnw = luma(nw), lne = luma(ne), lsw = luma(sw), lse = luma(se), lm = luma(m);
  const float lmin = fminf(lm, fminf(fminf(lnw, lne), fminf(lsw, lse)));
  const float lmax = fmaxf(lm, fmaxf(fmaxf(lnw, lne), fmaxf(lsw, lse)));
  float dx = -((lnw + lne) - (lsw + lse));
  float dy = ((lnw + lsw) - (lne + lse));
  const float reduce = fmaxf((lnw + lne + lsw + lse) * (0.25f * (1.0f / 8.0f)), 1.0f / 128.0f);
  const float rcp = 1.0f / (fminf(fabsf(dx), fabsf(dy)) + reduce);
  dx = fminf(8.0f, fmaxf(-8.0f, dx * rcp)) * ivx;
  dy = fminf(8.0f, fmaxf(-8.0f, dy * rcp)) * ivy;
  const float u = px * ivx, v = py * ivy;
  const float3 a1 = tex(in, w, h, u + dx * (1.0f / 3.0f - 0.5f), v + dy * (1.0f / 3.0f - 0.5f));
  const float3 a2 = tex(in, w, h, u + dx * (2.0f / 3.0f - 0.5f), v + dy * (2.0f / 3.0f - 0.5f));
  const float3 ra = make_float3(0.5f * (a1.x + a2.x), 0.5f * (a1.y + a2.y), 0.5f * (a1.z + a2.z));
  const float3 b1 = tex(in, w, h, u + dx * -0.5f, v + dy * -0.5f);
  const float3 b2 = tex(in, w, h, u + dx * 0.5f, v + dy * 0.5f);
  const float3 rb = make_float3(ra.x * 0.5f + 0.25f * (b1.x + b2.x), ra.y * 0.5f + 0.25f * (b1.y + b2.y),
                                ra.z * 0.5f + 0.25f * (b1.z + b2.z));
  const float lb = luma(rb);
  const float3 o = (lb < lmin || lb > lmax) ? ra : rb;
  out[i] = make_float4(o.x, o.y, o.z, 1.0f);
}

}  // namespace ppk
}  // namespace mrt

using namespace mrt;

namespace {

int launch_post(int op, const float* in, float* out, int w, int h, float exponent, float strength, hipStream_t s) {
  const unsigned grid = unsigned((size_t(w) * h + ppk::kBlock - 1) / ppk::kBlock);
  const float4* i4 = reinterpret_cast<const float4*>(in);
  float4* o4 = reinterpret_cast<float4*>(out);
  switch (op) {
    case MRT_POST_BILATERAL:
      hipLaunchKernelGGL(ppk::bilateral_kernel, dim3(grid), dim3(ppk::kBlock), 0, s, i4, o4, w, h, exponent, strength);
      break;
    case MRT_POST_TONEMAP_FILMIC:
      hipLaunchKernelGGL(ppk::tonemap_kernel, dim3(grid), dim3(ppk::kBlock), 0, s, i4, o4, w * h);
      break;
    case MRT_POST_FXAA:
      hipLaunchKernelGGL(ppk::fxaa_kernel, dim3(grid), dim3(ppk::kBlock), 0, s, i4, o4, w, h);
      break;
    default:
      return fail(MRT_ERR_INVALID, "unknown post-process pass");
  }
  return hipGetLastError() == hipSuccess ? MRT_OK : fail(MRT_ERR_HIP, "post-process launch failed");
}

bool bad_image(int w, int h) { return w < 1 || h < 1 || size_t(w) * h > (size_t(1) << 26); }

}  // namespace

extern "C" {

int mrt_post_pass_device(int op, const float* d_in, float* d_out, int w, int h, float exponent, float strength,
                         void* stream) {
  MRT_GUARD_BEGIN
  if (!d_in || !d_out || d_in == d_out) return fail(MRT_ERR_INVALID, "need distinct input and output images");
  if (bad_image(w, h)) return fail(MRT_ERR_INVALID, "bad image size");
  if (op == MRT_POST_BILATERAL && !(exponent > 0.f && strength > 0.f)) return fail(MRT_ERR_INVALID, "sigmas must be > 0");
  return launch_post(op, d_in, d_out, w, h, exponent, strength, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}

int mrt_post_process(const float* rgba_in, float* rgba_out, int w, int h, const int* ops, int n_ops, float exponent,
                     float strength) {
  MRT_GUARD_BEGIN
  if (!rgba_in || !rgba_out || (n_ops > 0 && !ops) || n_ops < 0) return fail(MRT_ERR_INVALID, "null argument");
  if (bad_image(w, h)) return fail(MRT_ERR_INVALID, "bad image size");
  const size_t bytes = size_t(w) * h * 16;
  float *a = nullptr, *b = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&a), bytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b), bytes);
  if (e == hipSuccess) e = hipMemcpy(a, rgba_in, bytes, hipMemcpyHostToDevice);
  int rc = e == hipSuccess ? MRT_OK : hip_fail(e, "post-process buffers");
  for (int k = 0; k < n_ops && rc == MRT_OK; ++k) {
    rc = mrt_post_pass_device(ops[k], a, b, w, h, exponent, strength, nullptr);
    std::swap(a, b);
  }
  if (rc == MRT_OK) {
    e = hipMemcpy(rgba_out, a, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "reading the post-processed image");
  }
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
  MRT_GUARD_END
}

}  // extern "C"
