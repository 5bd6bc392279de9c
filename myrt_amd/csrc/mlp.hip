// Neural radiance cache: a 64-64-64-64-64-64-3 ReLU MLP (no biases) on bf16 MFMA with fp32
// accumulation, trained online with Adam.
//
// Layout (all layers padded to 64x64 so one code path serves every layer; layer 5 uses rows 0-2):
//   w      bf16 [6][64][64]  W[out][in]      A operand of the forward products
//   wt     bf16 [6][64][64]  W^T[in][out]    A operand of the backward (data) products
//   master fp32 [6][64][64], Adam m, v alike; padding rows stay zero.
// One wave handles a tile of 32 queries. Products use v_mfma_f32_32x32x16_bf16 with
// D[m][n] = sum_k A[m][k] B[k][n]: lane l supplies A[m = l%32][k0 + 8(l/32) .. +8] and
// B[k0 + 8(l/32) .. +8][n = l%32] and receives D[8(v/4) + 4(l/32) + v%4][l%32], v = 0..15.
// Forward: m = output neuron, n = query, so lane l ends with 16 neurons of query l%32, written
// back to LDS as the next layer's [query][neuron] activation tile (row stride 72 bf16 so the
// 16-byte fragment reads of 32 rows spread over the banks).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace mrt {
namespace mlpk {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kW = 64;
constexpr int kStride = 72;
constexpr int kLayers = 6;
constexpr float kPi = 3.14159265358979f;

__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
  return uint16_t(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ inline bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline float sel3(int k, float x, float y, float z) { return k == 0 ? x : (k == 1 ? y : z); }

// The 64 network inputs of one query (input encoding): position in the cache's unit box with
// sin/cos at 6 octaves (36), direction and normal with sin/cos at 2 octaves (12 + 12),
// roughness (1) and albedo (3). q0..q3 are the query's four float4 rows.
__device__ inline float encode_feature(float4 q0, float4 q1, float4 q2, float4 q3, const CacheParams& P, int f) {
  if (f < 36) {
    const int axis = f / 12, r = f % 12, k = r >> 1;
    const float x = (sel3(axis, q0.x, q0.y, q0.z) - sel3(axis, P.lo[0], P.lo[1], P.lo[2])) *
                    sel3(axis, P.inv_ext[0], P.inv_ext[1], P.inv_ext[2]);
    const float a = x * float(1 << k) * kPi;
    return (r & 1) ? cosf(a) : sinf(a);
  }
  if (f < 60) {
    const int g = f - 36;
    const float4 v = g < 12 ? q1 : q2;
    const int r = g % 12, axis = r / 4, k = (r >> 1) & 1;
    const float a = sel3(axis, v.x, v.y, v.z) * float(1 << k) * kPi;
    return (r & 1) ? cosf(a) : sinf(a);
  }
  if (f == 60) return q0.w;
  return sel3(f - 61, q3.x, q3.y, q3.z);
}

__device__ inline void encode_tile(const CacheQuery* qs, uint32_t base, uint32_t n, const CacheParams& P,
                                   uint16_t* act) {
  const int lane = __lane_id();
  const int s = lane & 31, half = lane >> 5;
  const uint32_t qi = base + uint32_t(s);
  float4 q0 = make_float4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
  if (qi < n) {
    const float4* qp = reinterpret_cast<const float4*>(qs + qi);
    q0 = qp[0];
    q1 = qp[1];
    q2 = qp[2];
    q3 = qp[3];
  }
  for (int f = half * 32; f < half * 32 + 32; ++f) {
    const float v = qi < n ? encode_feature(q0, q1, q2, q3, P, f) : 0.0f;
    act[s * kStride + f] = to_bf16(v);
  }
}

// D(64 x 32) = A(64 x 64) * act^T: A rows from `a` (row stride 64, global), act [32][kStride] in LDS.
__device__ inline void product(const uint16_t* a, const uint16_t* act, f32x16& d0, f32x16& d1) {
  const int lane = __lane_id();
  const int r = lane & 31, h = lane >> 5;
  d0 = f32x16{};
  d1 = f32x16{};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const int k = kb * 16 + 8 * h;
    const bf16x8 b = ld8(act + r * kStride + k);
    const bf16x8 a0 = ld8(a + r * kW + k);
    const bf16x8 a1 = ld8(a + (32 + r) * kW + k);
    d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, d1, 0, 0, 0);
  }
}

// Writes D (neurons x queries) back as [query][neuron] bf16, optionally through ReLU, optionally
// masked by (mask[query][neuron] > 0).
__device__ inline void store_tile(const f32x16& d0, const f32x16& d1, uint16_t* out, bool relu, const uint16_t* mask) {
  const int lane = __lane_id();
  const int s = lane & 31, h = lane >> 5;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int n0 = 8 * (v >> 2) + 4 * h + (v & 3);
    float x0 = d0[v], x1 = d1[v];
    if (relu) {
      x0 = fmaxf(x0, 0.f);
      x1 = fmaxf(x1, 0.f);
    }
    if (mask) {
      if (!(from_bf16(mask[s * kStride + n0]) > 0.f)) x0 = 0.f;
      if (!(from_bf16(mask[s * kStride + 32 + n0]) > 0.f)) x1 = 0.f;
    }
    out[s * kStride + n0] = to_bf16(x0);
    out[s * kStride + 32 + n0] = to_bf16(x1);
  }
}

__global__ __launch_bounds__(256) void infer_kernel(CacheParams P, const CacheQuery* qs, uint32_t n_host,
                                                    const uint32_t* n_dev, float* out) {
  __shared__ __attribute__((aligned(16))) uint16_t tiles[4][2][32 * kStride];
  const int wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;  // a device count is capped by the buffer size
  const uint32_t tiles_total = (n + 31) / 32;
  for (uint32_t t = blockIdx.x * 4 + wave; t < tiles_total; t += gridDim.x * 4) {
    uint16_t* a = tiles[wave][0];
    uint16_t* b = tiles[wave][1];
    encode_tile(qs, t * 32, n, P, a);
    __builtin_amdgcn_wave_barrier();
    f32x16 d0, d1;
    for (int l = 0; l < kLayers; ++l) {
      product(P.w + size_t(l) * kW * kW, a, d0, d1);
      if (l < kLayers - 1) {
        store_tile(d0, d1, b, true, nullptr);
        __builtin_amdgcn_wave_barrier();
        uint16_t* tmp = a;
        a = b;
        b = tmp;
      }
    }
    // Outputs are neurons 0-2: lanes 0-31 hold them in d0[0..2] for query lane.
    const uint32_t qi = t * 32 + uint32_t(lane & 31);
    if (lane < 32 && qi < n) {
      out[3 * size_t(qi) + 0] = d0[0];
      out[3 * size_t(qi) + 1] = d0[1];
      out[3 * size_t(qi) + 2] = d0[2];
    }
  }
}

// Forward + backward of one tile per wave; writes H_l and G_l ([neuron][query], bf16) for the
// weight-gradient reduction, and the tile's loss sum.
__global__ __launch_bounds__(64) void train_kernel(CacheParams P, const CacheQuery* qs, const float* targets,
                                                   uint32_t n, float grad_scale, uint16_t* h_out, uint16_t* g_out,
                                                   float* loss_out) {
  __shared__ __attribute__((aligned(16))) uint16_t H[kLayers][32 * kStride];
  __shared__ __attribute__((aligned(16))) uint16_t G[2][32 * kStride];
  const int lane = __lane_id();
  const int s = lane & 31, h = lane >> 5;
  const uint32_t t = blockIdx.x;
  const uint32_t qi = t * 32 + uint32_t(s);
  encode_tile(qs, t * 32, n, P, H[0]);
  __syncthreads();
  f32x16 d0, d1;
  for (int l = 0; l < kLayers; ++l) {
    product(P.w + size_t(l) * kW * kW, H[l], d0, d1);
    if (l < kLayers - 1) store_tile(d0, d1, H[l + 1], true, nullptr);
    __syncthreads();
  }
  // Loss: relative L2 (y - t)^2 / (sg(y)^2 + 0.01) per channel; dL/dy = 2 (y - t) / (y^2 + 0.01).
  for (int i = lane; i < 32 * kStride; i += 64) G[0][i] = 0;
  __syncthreads();
  float loss = 0.f;
  if (h == 0 && qi < n) {
    for (int c = 0; c < 3; ++c) {
      const float y = d0[c], tg = targets[3 * size_t(qi) + c];
      const float den = y * y + 0.01f;
      const float e = y - tg;
      loss += e * e / den;
      G[0][s * kStride + c] = to_bf16(2.0f * e / den * grad_scale);
    }
  }
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (lane == 0) loss_out[t] = loss;
  int cur = 0;
  const uint32_t ld = P.batch_stride;  // queries per row of the [neuron][query] buffers
  for (int l = kLayers - 1; l >= 0; --l) {
    // Save G_l and H_l transposed for dW_l = G_l^T H_l.
    {
      const int neuron = lane;
      uint16_t* go = g_out + (size_t(l) * kW + neuron) * ld + size_t(t) * 32;
      uint16_t* ho = h_out + (size_t(l) * kW + neuron) * ld + size_t(t) * 32;
      for (int q = 0; q < 32; ++q) {
        go[q] = G[cur][q * kStride + neuron];
        ho[q] = H[l][q * kStride + neuron];
      }
    }
    if (l > 0) {
      product(P.wt + size_t(l) * kW * kW, G[cur], d0, d1);
      store_tile(d0, d1, G[cur ^ 1], false, H[l]);
      cur ^= 1;
      __syncthreads();
    }
  }
}

// dW_l partial over a range of queries: D(64 out x 64 in) = G_l[out][q] * H_l[in][q]^T.
__global__ __launch_bounds__(64) void wgrad_kernel(const uint16_t* h_in, const uint16_t* g_in, uint32_t batch_stride,
                                                   uint32_t n_pad, uint32_t chunk, float* partial) {
  const int lane = __lane_id();
  const int r = lane & 31, hh = lane >> 5;
  const int l = blockIdx.y;
  const uint32_t c = blockIdx.x;
  const uint32_t q0 = c * chunk, q1 = min(n_pad, q0 + chunk);
  const uint16_t* Gl = g_in + size_t(l) * kW * batch_stride;
  const uint16_t* Hl = h_in + size_t(l) * kW * batch_stride;
  f32x16 acc[2][2] = {};
  for (uint32_t q = q0; q < q1; q += 16) {
    const uint32_t k = q + 8 * uint32_t(hh);
    const bf16x8 a0 = ld8(Gl + size_t(r) * batch_stride + k);
    const bf16x8 a1 = ld8(Gl + size_t(32 + r) * batch_stride + k);
    const bf16x8 b0 = ld8(Hl + size_t(r) * batch_stride + k);
    const bf16x8 b1 = ld8(Hl + size_t(32 + r) * batch_stride + k);
    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
  }
  float* out = partial + (size_t(c) * kLayers + l) * kW * kW;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = mb * 32 + 8 * (v >> 2) + 4 * hh + (v & 3);
        const int col = nb * 32 + r;
        out[row * kW + col] = acc[mb][nb][v];
      }
}

__global__ __launch_bounds__(256) void adam_kernel(CacheParams P, const float* partial, uint32_t chunks, float lr,
                                                   float beta1, float beta2, float eps, float bc1, float bc2,
                                                   float* grad_out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= uint32_t(kLayers * kW * kW)) return;
  const uint32_t l = i / (kW * kW), rem = i % (kW * kW), row = rem / kW, col = rem % kW;
  if (l == kLayers - 1 && row >= 3) return;  // padding rows of the output layer stay zero
  float g = 0.f;
  for (uint32_t c = 0; c < chunks; ++c) g += partial[size_t(c) * kLayers * kW * kW + i];
  if (grad_out) grad_out[i] = g;
  if (lr == 0.f) return;
  const float m = beta1 * P.m[i] + (1.0f - beta1) * g;
  const float v = beta2 * P.v[i] + (1.0f - beta2) * g * g;
  P.m[i] = m;
  P.v[i] = v;
  const float w = P.master[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  P.master[i] = w;
  P.w[i] = to_bf16(w);
  P.wt[size_t(l) * kW * kW + col * kW + row] = to_bf16(w);
}

__global__ __launch_bounds__(256) void sync_bf16_kernel(CacheParams P) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= uint32_t(kLayers * kW * kW)) return;
  const uint32_t l = i / (kW * kW), rem = i % (kW * kW), row = rem / kW, col = rem % kW;
  const float w = P.master[i];
  P.w[i] = to_bf16(w);
  P.wt[size_t(l) * kW * kW + col * kW + row] = to_bf16(w);
}

}  // namespace mlpk

int launch_cache_infer(const CacheParams& P, const CacheQuery* q, uint32_t n, const uint32_t* n_dev, float* out,
                       int grid, hipStream_t s) {
  hipLaunchKernelGGL(mlpk::infer_kernel, dim3(grid), dim3(256), 0, s, P, q, n, n_dev, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cache_train(const CacheParams& P, const CacheQuery* q, const float* targets, uint32_t n, float grad_scale,
                       uint16_t* h_buf, uint16_t* g_buf, float* loss_tiles, float* partial, uint32_t chunk,
                       float lr, float beta1, float beta2, float eps, float bc1, float bc2, float* grad_out,
                       hipStream_t s) {
  const uint32_t tiles = (n + 31) / 32;
  const uint32_t n_pad = tiles * 32;
  hipLaunchKernelGGL(mlpk::train_kernel, dim3(tiles), dim3(64), 0, s, P, q, targets, n, grad_scale, h_buf, g_buf,
                     loss_tiles);
  const uint32_t chunks = (n_pad + chunk - 1) / chunk;
  hipLaunchKernelGGL(mlpk::wgrad_kernel, dim3(chunks, 6), dim3(64), 0, s, h_buf, g_buf, P.batch_stride, n_pad, chunk,
                     partial);
  hipLaunchKernelGGL(mlpk::adam_kernel, dim3((6 * 64 * 64 + 255) / 256), dim3(256), 0, s, P, partial, chunks, lr,
                     beta1, beta2, eps, bc1, bc2, grad_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cache_sync(const CacheParams& P, hipStream_t s) {
  hipLaunchKernelGGL(mlpk::sync_bf16_kernel, dim3((6 * 64 * 64 + 255) / 256), dim3(256), 0, s, P);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
