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

// sin(2 pi r) and cos(2 pi r) of r revolutions: the hardware's transcendental units take their
// argument in revolutions; the fractional part keeps it in their domain. The features are rounded
// to bf16 next, far coarser than these units' error.
__device__ inline float sin_rev(float r) { return __builtin_amdgcn_sinf(__builtin_amdgcn_fractf(r)); }
__device__ inline float cos_rev(float r) { return __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(r)); }

// The 64 network inputs of one query (input encoding): position in the cache's unit box with
// sin/cos of 2^k pi x at 6 octaves (36), direction and normal with sin/cos at 2 octaves (12 + 12),
// roughness (1) and albedo (3). q0..q3 are the query's four float4 rows. 2^k pi x radians are
// 2^(k-1) x revolutions (an exact scaling).
__device__ inline float encode_feature(float4 q0, float4 q1, float4 q2, float4 q3, const CacheParams& P, int f) {
  if (f < 36) {
    const int axis = f / 12, r = f % 12, k = r >> 1;
    const float x = (sel3(axis, q0.x, q0.y, q0.z) - sel3(axis, P.lo[0], P.lo[1], P.lo[2])) *
                    sel3(axis, P.inv_ext[0], P.inv_ext[1], P.inv_ext[2]);
    const float rev = x * float(1 << k) * 0.5f;
    return (r & 1) ? cos_rev(rev) : sin_rev(rev);
  }
  if (f < 60) {
    const int g = f - 36;
    const float4 v = g < 12 ? q1 : q2;
    const int r = g % 12, axis = r / 4, k = (r >> 1) & 1;
    const float rev = sel3(axis, v.x, v.y, v.z) * float(1 << k) * 0.5f;
    return (r & 1) ? cos_rev(rev) : sin_rev(rev);
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

__device__ inline uint32_t pack_bf16(float lo, float hi) { return uint32_t(to_bf16(lo)) | (uint32_t(to_bf16(hi)) << 16); }

// Writes D (neurons x queries) back as [query][neuron] bf16, optionally through ReLU. Lane (s, h)
// holds neurons 8j + 4h + (0..3) of query s in d0[4j..4j+3] (and 32 more in d1): four consecutive
// neurons, one 8-byte LDS write. Bit v of the returned mask is (stored value > 0) of d0[v], bit
// 16 + v that of d1[v]; with use_mask, values whose bit in `mask` is clear are stored as zero
// (the ReLU derivative of the backward pass).
__device__ inline uint32_t store_tile(const f32x16& d0, const f32x16& d1, uint16_t* out, bool relu, bool use_mask,
                                      uint32_t mask) {
  const int lane = __lane_id();
  const int s = lane & 31, h = lane >> 5;
  uint32_t pos = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float x[2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = 4 * j + i;
      x[0][i] = d0[v];
      x[1][i] = d1[v];
      if (relu) {
        x[0][i] = fmaxf(x[0][i], 0.f);
        x[1][i] = fmaxf(x[1][i], 0.f);
      }
      if (use_mask) {
        if (!((mask >> v) & 1u)) x[0][i] = 0.f;
        if (!((mask >> (16 + v)) & 1u)) x[1][i] = 0.f;
      }
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const uint2 w = make_uint2(pack_bf16(x[m][0], x[m][1]), pack_bf16(x[m][2], x[m][3]));
      *reinterpret_cast<uint2*>(out + s * kStride + 32 * m + 8 * j + 4 * h) = w;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint16_t bits = uint16_t(((i < 2 ? w.x : w.y) >> (16 * (i & 1))) & 0xffffu);
        if (from_bf16(bits) > 0.f) pos |= 1u << (16 * m + 4 * j + i);
      }
    }
  }
  return pos;
}

// Tile [query][neuron] (row stride kStride) to dst[neuron][32 queries]: lane = neuron gathers its
// column and writes it as 64 contiguous bytes; the wave writes 4 KB in one run.
__device__ inline void save_columns(const uint16_t* tile, uint16_t* dst) {
  const int neuron = __lane_id();
  uint32_t w[16];
#pragma unroll
  for (int q = 0; q < 32; q += 2)
    w[q / 2] = uint32_t(tile[q * kStride + neuron]) | (uint32_t(tile[(q + 1) * kStride + neuron]) << 16);
  uint4* o = reinterpret_cast<uint4*>(dst + size_t(neuron) * 32);
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
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
        store_tile(d0, d1, b, true, false, 0u);
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

// Forward + backward of one tile per wave; writes H_l and G_l for the weight-gradient reduction
// and the tile's loss sum. Activation and gradient buffers are [layer][tile][neuron][32 queries]
// (layer stride 64 * batch_stride). The ReLU pattern of every hidden layer stays in registers
// (one bit per value a lane stores), so LDS holds only two activation and two gradient tiles.
__global__ __launch_bounds__(64) void train_kernel(CacheParams P, const CacheQuery* qs, const float* targets,
                                                   uint32_t n, float grad_scale, uint16_t* h_out, uint16_t* g_out,
                                                   float* loss_out) {
  __shared__ __attribute__((aligned(16))) uint16_t H[2][32 * kStride];
  __shared__ __attribute__((aligned(16))) uint16_t G[2][32 * kStride];
  const int lane = __lane_id();
  const int s = lane & 31, h = lane >> 5;
  const uint32_t t = blockIdx.x;
  const uint32_t qi = t * 32 + uint32_t(s);
  const size_t layer_stride = size_t(kW) * P.batch_stride;
  const size_t tile_off = size_t(t) * kW * 32;
  encode_tile(qs, t * 32, n, P, H[0]);
  __syncthreads();
  uint32_t relu_mask[kLayers];  // [l]: ReLU pattern of H_l, l = 1..5
  f32x16 d0, d1;
  int cur = 0;
  for (int l = 0; l < kLayers; ++l) {
    save_columns(H[cur], h_out + l * layer_stride + tile_off);
    product(P.w + size_t(l) * kW * kW, H[cur], d0, d1);
    if (l < kLayers - 1) {
      relu_mask[l + 1] = store_tile(d0, d1, H[cur ^ 1], true, false, 0u);
      cur ^= 1;
    }
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
  int gc = 0;
  for (int l = kLayers - 1; l >= 0; --l) {
    save_columns(G[gc], g_out + l * layer_stride + tile_off);
    if (l > 0) {
      product(P.wt + size_t(l) * kW * kW, G[gc], d0, d1);
      store_tile(d0, d1, G[gc ^ 1], false, true, relu_mask[l]);
      gc ^= 1;
      __syncthreads();
    }
  }
}

// dW_l partial over a range of queries: D(64 out x 64 in) = G_l[out][q] * H_l[in][q]^T, with G_l and
// H_l in the [tile][neuron][32 queries] layout of train_kernel.
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
    // [tile][neuron][32 queries]: the 8 queries of a fragment lie in one tile row.
    const uint32_t k = q + 8 * uint32_t(hh);
    const size_t base = size_t(k >> 5) * kW * 32 + (k & 31u);
    const bf16x8 a0 = ld8(Gl + base + size_t(r) * 32);
    const bf16x8 a1 = ld8(Gl + base + size_t(32 + r) * 32);
    const bf16x8 b0 = ld8(Hl + base + size_t(r) * 32);
    const bf16x8 b1 = ld8(Hl + base + size_t(32 + r) * 32);
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
