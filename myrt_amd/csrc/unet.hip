// Denoiser U-Net kernels for gfx950.
//
// Convolutions are implicit GEMMs on MFMA: output channels are the M dimension (32 per block),
// pixels the N dimension (32 consecutive pixels of a row per block) and input channels x taps the
// K dimension. A wave owns 64 output channels x 64 pixels (4 accumulator blocks); operands come
// straight from NHWC tensors with 16-byte loads (8 halves or 4 floats per lane). Inference runs in
// fp16 (v_mfma_f32_32x32x16_f16), training in fp32 (v_mfma_f32_32x32x2_f32), both accumulating in
// fp32. Weight gradients reduce over pixels with the fp32 MFMA (K = 2 pixels per instruction),
// whose operand layout reads 32 consecutive channels of one pixel per half wave.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "unet.hpp"

namespace mrt {
namespace unk {

typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 h8;
typedef __attribute__((__vector_size__(4 * sizeof(_Float16)))) _Float16 h4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kBlock = 256;

template <typename T>
struct Frag;

// fp16: one 32x32x16 MFMA per K step of 16 channels; lane l holds channels 8*(l/32) .. +7.
template <>
struct Frag<_Float16> {
  static constexpr int kStep = 16;
  static constexpr int kLane = 8;
  using V = h8;
  __device__ static V load(const _Float16* p) { return *reinterpret_cast<const V*>(p); }
  __device__ static void mma(const V& a, const V& b, f32x16& acc) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
  }
  __device__ static void store4(_Float16* p, float a, float b, float c, float d) {
    h4 v;
    v[0] = (_Float16)a;
    v[1] = (_Float16)b;
    v[2] = (_Float16)c;
    v[3] = (_Float16)d;
    *reinterpret_cast<h4*>(p) = v;
  }
  __device__ static void load4(const _Float16* p, float* o) {
    const h4 v = *reinterpret_cast<const h4*>(p);
    o[0] = float(v[0]);
    o[1] = float(v[1]);
    o[2] = float(v[2]);
    o[3] = float(v[3]);
  }
};

// fp32: four 32x32x2 MFMAs per K step of 8 channels; lane l holds channels 4*(l/32) .. +3 and
// MFMA j takes element j (k = 0 from the lower half wave, k = 1 from the upper).
template <>
struct Frag<float> {
  static constexpr int kStep = 8;
  static constexpr int kLane = 4;
  using V = f4;
  __device__ static V load(const float* p) { return *reinterpret_cast<const V*>(p); }
  __device__ static void mma(const V& a, const V& b, f32x16& acc) {
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[2], b[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[3], b[3], acc, 0, 0, 0);
  }
  __device__ static void store4(float* p, float a, float b, float c, float d) {
    *reinterpret_cast<f4*>(p) = f4{a, b, c, d};
  }
  __device__ static void load4(const float* p, float* o) {
    const f4 v = *reinterpret_cast<const f4*>(p);
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = v[3];
  }
};

// One wave: output channels m0 .. m0+63 (m0 .. m0+31 when cout - m0 == 32) of pixels
// x0 .. x0+63 of row y of image img. Requirements (checked by the launcher): cout a multiple of
// 32, c0 and c1 multiples of the K step, H and W even when in0 is upsampled.
template <typename T, int TAPS>
__global__ __launch_bounds__(kBlock) void conv_kernel(ConvArgs a) {
  using F = Frag<T>;
  using V = typename F::V;
  const int lane = threadIdx.x & 63;
  const int wave = int((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const int col = lane & 31, half = lane >> 5;
  const int nxb = (a.W + 63) >> 6;
  const int ncb = (a.cout + 63) >> 6;
  const int total = a.B * a.H * nxb * ncb;
  if (wave >= total) return;
  int t = wave;
  const int cb = t % ncb;
  t /= ncb;
  const int xb = t % nxb;
  t /= nxb;
  const int y = t % a.H;
  const int img = t / a.H;
  const int m0 = cb * 64;
  const bool two_m = m0 + 32 < a.cout;
  const int cin = a.c0 + a.c1;
  const int H0 = a.up0 ? (a.H >> 1) : a.H, W0 = a.up0 ? (a.W >> 1) : a.W;
  const T* in0 = static_cast<const T*>(a.in0);
  const T* in1 = static_cast<const T*>(a.in1);
  const T* w = static_cast<const T*>(a.w);

  f32x16 acc00 = {}, acc01 = {}, acc10 = {}, acc11 = {};
  for (int tap = 0; tap < TAPS; ++tap) {
    const int dy = TAPS == 9 ? tap / 3 - 1 : 0;
    const int dx = TAPS == 9 ? tap % 3 - 1 : 0;
    const int sy = y + dy;
    const int px0 = xb * 64 + col, px1 = px0 + 32;
    const int sx0 = px0 + dx, sx1 = px1 + dx;
    const bool row_ok = sy >= 0 && sy < a.H;
    const bool ok0 = row_ok && px0 < a.W && sx0 >= 0 && sx0 < a.W;
    const bool ok1 = row_ok && px1 < a.W && sx1 >= 0 && sx1 < a.W;
    // Element offsets of the two source pixels in in0 (possibly upsampled) and in1.
    const int y0 = a.up0 ? (sy >> 1) : sy;
    const size_t o00 = ok0 ? ((size_t(img) * H0 + y0) * W0 + (a.up0 ? (sx0 >> 1) : sx0)) * a.c0 : 0;
    const size_t o01 = ok1 ? ((size_t(img) * H0 + y0) * W0 + (a.up0 ? (sx1 >> 1) : sx1)) * a.c0 : 0;
    const size_t o10 = ok0 ? ((size_t(img) * a.H + sy) * a.W + sx0) * a.c1 : 0;
    const size_t o11 = ok1 ? ((size_t(img) * a.H + sy) * a.W + sx1) * a.c1 : 0;
    const T* wr0 = w + (size_t(m0 + col) * TAPS + tap) * cin + half * F::kLane;
    const T* wr1 = two_m ? wr0 + size_t(32) * TAPS * cin : wr0;
    for (int c = 0; c < cin; c += F::kStep) {
      V b0 = {}, b1 = {};
      if (c < a.c0) {
        if (ok0) b0 = F::load(in0 + o00 + c + half * F::kLane);
        if (ok1) b1 = F::load(in0 + o01 + c + half * F::kLane);
      } else {
        if (ok0) b0 = F::load(in1 + o10 + (c - a.c0) + half * F::kLane);
        if (ok1) b1 = F::load(in1 + o11 + (c - a.c0) + half * F::kLane);
      }
      const V w0 = F::load(wr0 + c);
      F::mma(w0, b0, acc00);
      F::mma(w0, b1, acc01);
      if (two_m) {
        const V w1 = F::load(wr1 + c);
        F::mma(w1, b0, acc10);
        F::mma(w1, b1, acc11);
      }
    }
  }
  // Lane (col, half) of block (mt, nt) holds pixel x0 + 32 nt + col and output channels
  // m0 + 32 mt + 8 j + 4 half + i (element 4 j + i).
  T* out = static_cast<T*>(a.out);
  const T* mask = static_cast<const T*>(a.mask);
  for (int mt = 0; mt < (two_m ? 2 : 1); ++mt) {
    for (int nt = 0; nt < 2; ++nt) {
      const int px = xb * 64 + nt * 32 + col;
      if (px >= a.W) continue;
      const f32x16& acc = mt == 0 ? (nt == 0 ? acc00 : acc01) : (nt == 0 ? acc10 : acc11);
      const size_t base = ((size_t(img) * a.H + y) * a.W + px) * a.cout;
      for (int j = 0; j < 4; ++j) {
        const int co = m0 + mt * 32 + 8 * j + 4 * half;
        float v[4];
        for (int i = 0; i < 4; ++i) {
          v[i] = acc[4 * j + i];
          if (a.bias) v[i] += a.bias[co + i];
          if (a.relu) v[i] = fmaxf(v[i], 0.0f);
        }
        if (mask) {
          float mv[4];
          F::load4(mask + base + co, mv);
          for (int i = 0; i < 4; ++i) v[i] = mv[i] > 0.0f ? v[i] : 0.0f;
        }
        F::store4(out + base + co, v[0], v[1], v[2], v[3]);
      }
    }
  }
}

// Tiled fp16 3x3 convolution for inference. A workgroup of ROWS waves computes a tile of ROWS
// output rows x 64 pixels x NM blocks of 32 output channels (output-channel group g of the
// layer's cout / (32 NM)); wave w owns row w of the tile. For each chunk of 16 input channels the
// workgroup stages in LDS the input halo of the tile (ROWS + 2 rows x 66 pixels) and the
// chunk's weights for all 9 taps, then every wave runs 9 taps x NM x 2 MFMAs (32x32x16) out of LDS:
// an input element is read from L2 once per tile and channel group instead of once per tap. The
// next chunk's global loads are issued into registers before the current chunk's MFMAs, so their
// latency overlaps the math. LDS layouts put the two 8-channel halves of a fragment in separate
// planes, so a half wave reads 32 consecutive 16-byte records (no bank conflicts): input
// [half][row][col][8], weights [half][tap][cout][8].
constexpr int kTileCols = 64;
constexpr int kHaloCols = kTileCols + 2;

// Network input channel c (0 .. kUnetInPhys - 1) from a raw feature record: log1p of the clamped
// radiance, albedo and normal as they are, 1 / (1 + depth) (0 for no hit), zero padding.
__device__ inline float input_channel(const float* s, int c) {
  if (c < 3) return log1pf(fmaxf(s[c], 0.0f));
  if (c < 9) return s[c];
  if (c == 9) {
    const float d = s[9];
    return (d > 0.0f && d < __int_as_float(0x7f800000)) ? 1.0f / (1.0f + d) : 0.0f;
  }
  return 0.0f;
}

// The 1x1 output convolution of one pixel: y = b + W x over cin channels, x read as 8-channel
// records rec(k) for k = 0, 8, ...; the products are accumulated in this fixed order.
template <typename Rec>
__device__ inline void output_conv(Rec rec, int cin, const _Float16* w, const float* bias, float (&y)[3]) {
  y[0] = bias[0];
  y[1] = bias[1];
  y[2] = bias[2];
  for (int k = 0; k < cin; k += 8) {
    const h8 v = rec(k);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const h8 wv = *reinterpret_cast<const h8*>(w + size_t(c) * cin + k);
#pragma unroll
      for (int e = 0; e < 8; ++e) y[c] = fmaf(float(wv[e]), float(v[e]), y[c]);
    }
  }
}

// The finish step: the log-space residual y added to log1p(radiance) and mapped back, alpha 1.
__device__ inline float4 finish_rgba(const float* raw_px, const float (&y)[3]) {
  float o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float z = log1pf(fmaxf(raw_px[c], 0.0f)) + y[c];
    o[c] = expm1f(fmaxf(z, 0.0f));
  }
  return make_float4(o[0], o[1], o[2], 1.0f);
}

// FUSE: the output layer and the finish step run on the tile (ConvArgs::out_w); a separate
// instantiation keeps the plain kernels' registers as they are.
template <int NM, int ROWS, bool FUSE = false>
__global__ __launch_bounds__(ROWS * 64) void conv_tile_kernel(ConvArgs a, int groups) {
  constexpr int kThreads = ROWS * 64;  // one wave per output row of the tile
  constexpr int kHRows = ROWS + 2;
  constexpr int kHItems = 2 * kHRows * kHaloCols;  // 16-byte records of the input halo
  constexpr int kHPer = (kHItems + kThreads - 1) / kThreads;
  constexpr int kCout = NM * 32;
  constexpr int kWItems = 2 * 9 * kCout;
  constexpr int kWPer = (kWItems + kThreads - 1) / kThreads;
  // One LDS block: the input halo and the weights while computing, then the output tile.
  constexpr int kOutPitch = 5;  // 16-byte records per staged pixel: 4 of data, 1 of padding
  constexpr int kBuf = kHItems + kWItems > ROWS * kTileCols * kOutPitch ? kHItems + kWItems : ROWS * kTileCols * kOutPitch;
  __shared__ h8 s_buf[kBuf];
  h8* const s_in = s_buf;
  h8* const s_w = s_buf + kHItems;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 31, half = lane >> 5;
  const int ntx = (a.W + kTileCols - 1) / kTileCols, nty = (a.H + ROWS - 1) / ROWS;
  int t = blockIdx.x;
  const int g = t % groups;
  t /= groups;
  const int tx = t % ntx;
  t /= ntx;
  const int ty = t % nty;
  const int img = t / nty;
  const int x0 = tx * kTileCols, y0 = ty * ROWS, m0 = g * kCout;
  const int cin = a.c0 + a.c1;
  const int H0 = a.up0 ? (a.H >> 1) : a.H, W0 = a.up0 ? (a.W >> 1) : a.W;
  const _Float16* in0 = static_cast<const _Float16*>(a.in0);
  const _Float16* in1 = static_cast<const _Float16*>(a.in1);
  const _Float16* w = static_cast<const _Float16*>(a.w);

  h8 r_in[kHPer], r_w[kWPer];
  const auto fetch = [&](int c) {
#pragma unroll
    for (int k = 0; k < kHPer; ++k) {
      const int i = tid + k * kThreads;
      h8 v = {};
      if (i < kHItems) {
        const int hf = i & 1, pix = i >> 1;
        const int r = pix / kHaloCols, cc = pix - r * kHaloCols;
        const int sy = y0 - 1 + r, sx = x0 - 1 + cc;
        if (sy >= 0 && sy < a.H && sx >= 0 && sx < a.W) {
          if (c < a.c0) {
            const int yy = a.up0 ? (sy >> 1) : sy, xx = a.up0 ? (sx >> 1) : sx;
            v = *reinterpret_cast<const h8*>(in0 + ((size_t(img) * H0 + yy) * W0 + xx) * a.c0 + c + 8 * hf);
          } else {
            v = *reinterpret_cast<const h8*>(in1 + ((size_t(img) * a.H + sy) * a.W + sx) * a.c1 + (c - a.c0) +
                                             8 * hf);
          }
        }
      }
      r_in[k] = v;
    }
#pragma unroll
    for (int k = 0; k < kWPer; ++k) {
      const int i = tid + k * kThreads;
      if (i < kWItems) {
        const int hf = i & 1, rest = i >> 1;
        const int tap = rest / kCout, co = rest - tap * kCout;
        r_w[k] = *reinterpret_cast<const h8*>(w + (size_t(m0 + co) * 9 + tap) * cin + c + 8 * hf);
      }
    }
  };
  const auto stage = [&]() {
#pragma unroll
    for (int k = 0; k < kHPer; ++k) {
      const int i = tid + k * kThreads;
      if (i < kHItems) {
        const int hf = i & 1, pix = i >> 1;
        const int r = pix / kHaloCols, cc = pix - r * kHaloCols;
        s_in[(hf * kHRows + r) * kHaloCols + cc] = r_in[k];
      }
    }
#pragma unroll
    for (int k = 0; k < kWPer; ++k) {
      const int i = tid + k * kThreads;
      if (i < kWItems) {
        // Consecutive threads take consecutive output channels of one tap, so a wave's records
        // land on consecutive LDS addresses of each plane (tap-major order put 9 lanes on one bank).
        const int hf = i & 1, rest = i >> 1;
        const int tap = rest / kCout, co = rest - tap * kCout;
        s_w[(hf * 9 + tap) * kCout + co] = r_w[k];
      }
    }
  };

  f32x16 acc[NM][2];
#pragma unroll
  for (int mb = 0; mb < NM; ++mb) acc[mb][0] = acc[mb][1] = f32x16{};

  fetch(0);
  for (int c = 0; c < cin; c += 16) {
    __syncthreads();  // the previous chunk's LDS reads are done
    stage();
    __syncthreads();
    if (c + 16 < cin) fetch(c + 16);
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap % 3;
      const int row = wave + dy;
      const h8 b0 = s_in[(half * kHRows + row) * kHaloCols + col + dx];
      const h8 b1 = s_in[(half * kHRows + row) * kHaloCols + 32 + col + dx];
#pragma unroll
      for (int mb = 0; mb < NM; ++mb) {
        const h8 av = s_w[(half * 9 + tap) * kCout + mb * 32 + col];
        acc[mb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, b0, acc[mb][0], 0, 0, 0);
        acc[mb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, b1, acc[mb][1], 0, 0, 0);
      }
    }
  }
  // Lane (col, half) of block (mb, nb) holds pixel x0 + 32 nb + col of row y0 + wave and output
  // channels m0 + 32 mb + 8 j + 4 half + i (element 4 j + i). Each block of 32 channels goes
  // through LDS as [row][pixel][4 records of 8 channels] (pixel pitch 5 records, which spreads a
  // half wave's 8-byte writes over the banks), so that the global stores write whole 16-byte
  // records of contiguous 64-byte pixel runs: a row of the tile is one 4 KB run when the layer
  // has 32 output channels, instead of 8-byte pieces 64 bytes apart.
  _Float16* out = static_cast<_Float16*>(a.out);
  __syncthreads();  // every wave is done with the halo and the weights
#pragma unroll
  for (int mb = 0; mb < NM; ++mb) {
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      const int px = nb * 32 + col;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = m0 + mb * 32 + 8 * j + 4 * half;
        h4 v;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float x = acc[mb][nb][4 * j + i];
          if (a.bias) x += a.bias[co + i];
          if (a.relu) x = fmaxf(x, 0.0f);
          v[i] = (_Float16)x;
        }
        *reinterpret_cast<h4*>(reinterpret_cast<char*>(s_buf + (wave * kTileCols + px) * kOutPitch + j) + 8 * half) = v;
      }
    }
    __syncthreads();
    if (FUSE) {
      // The output layer and the finish step fused in (cout == 32: this is the only block): one
      // pixel per thread from the staged tile, written as rgba; the 32 channels never reach HBM.
      const _Float16* ow = static_cast<const _Float16*>(a.out_w);
      for (int k = tid; k < ROWS * kTileCols; k += kThreads) {
        const int row = k / kTileCols, pix = k % kTileCols;
        const int y = y0 + row, px = x0 + pix;
        if (y < a.Hs && px < a.Ws) {
          float yv[3];
          output_conv([&](int kk) { return s_buf[(row * kTileCols + pix) * kOutPitch + kk / 8]; }, 32, ow, a.out_b,
                      yv);
          const size_t q = (size_t(img) * a.Hs + y) * a.Ws + px;
          reinterpret_cast<float4*>(a.rgba)[q] = finish_rgba(a.raw + q * kUnetInCh, yv);
        }
      }
    } else {
      constexpr int kRecords = ROWS * kTileCols * 4;
#pragma unroll
      for (int it = 0; it < kRecords / kThreads; ++it) {
        const int k = tid + it * kThreads;
        const int row = k / (kTileCols * 4), rem = k % (kTileCols * 4), pix = rem >> 2, q = rem & 3;
        const int y = y0 + row, px = x0 + pix;
        if (y < a.H && px < a.W)
          *reinterpret_cast<h8*>(out + ((size_t(img) * a.H + y) * a.W + px) * a.cout + m0 + mb * 32 + 8 * q) =
              s_buf[(row * kTileCols + pix) * kOutPitch + q];
      }
    }
    if (a.pool_out) {
      // The 2x2 average of the staged block (tiles start on even rows and columns): one 16-byte
      // record per thread, summed in pool_kernel's order.
      const int prow = tid / (kTileCols * 2), rem = tid % (kTileCols * 2), ppx = rem >> 2, q = rem & 3;
      const int yo = (y0 >> 1) + prow, xo = (x0 >> 1) + ppx;
      if (yo < (a.H >> 1) && xo < (a.W >> 1)) {
        const h8 v00 = s_buf[((2 * prow) * kTileCols + 2 * ppx) * kOutPitch + q];
        const h8 v01 = s_buf[((2 * prow) * kTileCols + 2 * ppx + 1) * kOutPitch + q];
        const h8 v10 = s_buf[((2 * prow + 1) * kTileCols + 2 * ppx) * kOutPitch + q];
        const h8 v11 = s_buf[((2 * prow + 1) * kTileCols + 2 * ppx + 1) * kOutPitch + q];
        h8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          o[e] = (_Float16)((float(v00[e]) + float(v01[e]) + float(v10[e]) + float(v11[e])) * 0.25f);
        *reinterpret_cast<h8*>(static_cast<_Float16*>(a.pool_out) +
                               ((size_t(img) * (a.H >> 1) + yo) * (a.W >> 1) + xo) * a.cout + m0 + mb * 32 + 8 * q) = o;
      }
    }
    if (mb + 1 < NM) __syncthreads();  // the next block's staging overwrites this one
  }
}

// One wave: dW block [m0, m0+32) x [n0, n0+32) of one tap over `chunk` pixels.
__global__ __launch_bounds__(kBlock) void wgrad_kernel(WgradArgs a, int taps) {
  const int lane = threadIdx.x & 63;
  const int wave = int((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const int col = lane & 31, half = lane >> 5;
  const int cin = a.c0 + a.c1;
  const int tm = a.cout / 32, tn = (cin + 31) / 32;
  const size_t npix = size_t(a.B) * a.H * a.W;
  const int nchunks = int((npix + a.chunk - 1) / a.chunk);
  const int total = tm * tn * taps * nchunks;
  if (wave >= total) return;
  int t = wave;
  const int mi = t % tm;
  t /= tm;
  const int ni = t % tn;
  t /= tn;
  const int tap = t % taps;
  const int chunk = t / taps;
  const int dy = taps == 9 ? tap / 3 - 1 : 0, dx = taps == 9 ? tap % 3 - 1 : 0;
  const int m0 = mi * 32, n0 = ni * 32;
  const int co = m0 + col, ci = n0 + col;
  const bool ci_ok = ci < cin;
  const int H0 = a.up0 ? (a.H >> 1) : a.H, W0 = a.up0 ? (a.W >> 1) : a.W;
  const size_t p_begin = size_t(chunk) * a.chunk;
  const size_t p_end = p_begin + a.chunk < npix ? p_begin + a.chunk : npix;
  f32x16 acc = {};
  for (size_t p = p_begin; p < p_end; p += 2) {
    const size_t pp = p + half;
    float av = 0.0f, bv = 0.0f;
    if (pp < p_end) {
      av = a.g[pp * a.cout + co];
      const int x = int(pp % a.W);
      const size_t r = pp / a.W;
      const int y = int(r % a.H);
      const int img = int(r / a.H);
      const int sy = y + dy, sx = x + dx;
      if (ci_ok && sy >= 0 && sy < a.H && sx >= 0 && sx < a.W) {
        if (ci < a.c0) {
          const int yy = a.up0 ? (sy >> 1) : sy, xx = a.up0 ? (sx >> 1) : sx;
          bv = a.in0[((size_t(img) * H0 + yy) * W0 + xx) * a.c0 + ci];
        } else {
          bv = a.in1[((size_t(img) * a.H + sy) * a.W + sx) * a.c1 + (ci - a.c0)];
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
  }
  if (!ci_ok) return;
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      const int row = m0 + 8 * j + 4 * half + i;
      unsafeAtomicAdd(a.dw + (size_t(row) * taps + tap) * cin + ci, acc[4 * j + i]);
    }
}

// db[co] += sum of g[p][co] over a chunk of 1024 pixels (one thread per channel).
__global__ __launch_bounds__(kBlock) void bgrad_kernel(const float* g, float* db, int cout, size_t npix) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  const size_t p0 = size_t(blockIdx.x) * 1024;
  const size_t p1 = p0 + 1024 < npix ? p0 + 1024 : npix;
  float s = 0.0f;
  for (size_t p = p0; p < p1; ++p) s += g[p * cout + co];
  unsafeAtomicAdd(db + co, s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void pool_kernel(const T* in, T* out, int B, int H, int W, int C) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const int Ho = H >> 1, Wo = W >> 1;
  const size_t n = size_t(B) * Ho * Wo * C;
  if (i >= n) return;
  const int c = int(i % C);
  size_t r = i / C;
  const int x = int(r % Wo);
  r /= Wo;
  const int y = int(r % Ho);
  const int b = int(r / Ho);
  const size_t row0 = (size_t(b) * H + 2 * y) * W, row1 = row0 + W;
  const float s = float(in[(row0 + 2 * x) * C + c]) + float(in[(row0 + 2 * x + 1) * C + c]) +
                  float(in[(row1 + 2 * x) * C + c]) + float(in[(row1 + 2 * x + 1) * C + c]);
  out[i] = T(s * 0.25f);
}

__global__ __launch_bounds__(kBlock) void grad_skip_pool_kernel(const float* dcat, int ctot, int off, const float* dP,
                                                                const float* E, float* dE, int B, int H, int W,
                                                                int C) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const size_t n = size_t(B) * H * W * C;
  if (i >= n) return;
  const int c = int(i % C);
  const size_t p = i / C;
  float v = dcat ? dcat[p * ctot + off + c] : 0.0f;
  if (dP) {
    const int x = int(p % W);
    const size_t r = p / W;
    const int y = int(r % H);
    const int b = int(r / H);
    v += 0.25f * dP[((size_t(b) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c];
  }
  dE[i] = E[i] > 0.0f ? v : 0.0f;
}

// Coarse element (b, y, x, c): sum of dcat over the 2x2 fine block, channels [0, C).
__global__ __launch_bounds__(kBlock) void grad_up_kernel(const float* dcat, int ctot, const float* D, float* dD, int B,
                                                         int H, int W, int C) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const int Hc = H >> 1, Wc = W >> 1;
  const size_t n = size_t(B) * Hc * Wc * C;
  if (i >= n) return;
  const int c = int(i % C);
  size_t r = i / C;
  const int x = int(r % Wc);
  r /= Wc;
  const int y = int(r % Hc);
  const int b = int(r / Hc);
  const size_t row0 = (size_t(b) * H + 2 * y) * W, row1 = row0 + W;
  const float s = dcat[(row0 + 2 * x) * ctot + c] + dcat[(row0 + 2 * x + 1) * ctot + c] +
                  dcat[(row1 + 2 * x) * ctot + c] + dcat[(row1 + 2 * x + 1) * ctot + c];
  dD[i] = D[i] > 0.0f ? s : 0.0f;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void prep_kernel(const float* raw, T* x, int B, int Hs, int Ws, int H, int W) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const size_t n = size_t(B) * H * W;
  if (i >= n) return;
  const int px = int(i % W);
  const size_t r = i / W;
  const int py = int(r % H);
  const int b = int(r / H);
  const int sx = px < Ws ? px : Ws - 1, sy = py < Hs ? py : Hs - 1;
  const float* s = raw + ((size_t(b) * Hs + sy) * Ws + sx) * kUnetInCh;
  T* o = x + i * kUnetInPhys;
  for (int c = 0; c < kUnetInPhys; ++c) o[c] = T(input_channel(s, c));
}

__global__ __launch_bounds__(kBlock) void loss_kernel(const float* x, const float* y, int yc, const float* target,
                                                      float* g, float* partial, size_t npix, float inv_n) {
  __shared__ float red[kBlock];
  const size_t p = size_t(blockIdx.x) * kBlock + threadIdx.x;
  float l = 0.0f;
  if (p < npix) {
    for (int c = 0; c < yc; ++c) {
      float gv = 0.0f;
      if (c < 3) {
        const float z = x[p * kUnetInPhys + c] + y[p * yc + c];
        const float tt = log1pf(fmaxf(target[p * 3 + c], 0.0f));
        const float d = z - tt;
        l += fabsf(d);
        gv = d > 0.0f ? inv_n : (d < 0.0f ? -inv_n : 0.0f);
      }
      g[p * yc + c] = gv;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}



__global__ __launch_bounds__(kBlock) void transpose_w_kernel(const float* w, float* wt, int cout, int taps, int cin) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const size_t n = size_t(cout) * taps * cin;
  if (i >= n) return;
  const int ci = int(i % cin);
  const size_t r = i / cin;
  const int tap = int(r % taps);
  const int co = int(r / taps);
  wt[(size_t(ci) * taps + (taps - 1 - tap)) * cout + co] = w[i];
}

__global__ __launch_bounds__(kBlock) void to_half_kernel(const float* in, _Float16* out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = (_Float16)in[i];
}

__global__ __launch_bounds__(kBlock) void adam_kernel(float* w, float* m, float* v, const float* g, size_t n, float lr,
                                                      float b1, float b2, float eps, float bc1, float bc2) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks_for(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }
inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace unk

template <typename T>
int launch_unet_conv(const ConvArgs& a, int taps, hipStream_t s) {
  const int step = unk::Frag<T>::kStep;
  if (a.cout % 32 || a.c0 % step || a.c1 % step || (a.up0 && (a.H % 2 || a.W % 2)) || (a.c1 && !a.in1) ||
      (taps != 9 && taps != 1))
    return -1;
  const bool fused = a.out_w != nullptr;
  if (fused && (!a.raw || !a.rgba || !a.out_b || a.cout != 32 || a.Hs < 1 || a.Ws < 1 || a.Hs > a.H || a.Ws > a.W))
    return -1;
  if (std::is_same<T, _Float16>::value && taps == 9 && !a.mask && a.cout <= 128 && a.c0 % 16 == 0 &&
      a.c1 % 16 == 0) {
    // Inference 3x3: the LDS-tiled kernel, one workgroup per 4 x 64 output tile and group of
    // 32 NM output channels. NM is the largest divisor of cout / 32 that still gives at least
    // 1024 workgroups (4 per CU), so the coarse levels split their channels across workgroups.
    // Where 8-row tiles of one 32-channel block still give 1024 workgroups (the two finest
    // levels at 1080p), they take the place of 4-row ones: the halo re-read falls from 1.5x to
    // 1.25x and each chunk's weights serve twice the pixels.
    const auto tiles_of = [&](int rows) {
      return unsigned(a.B) * unsigned((a.H + rows - 1) / rows) * unsigned((a.W + unk::kTileCols - 1) / unk::kTileCols);
    };
    const unsigned tiles = tiles_of(4);
    const int mt = a.cout / 32;
    int nm = 1;
    for (int d = mt; d >= 1; --d)
      if (mt % d == 0 && tiles * unsigned(mt / d) >= 1024u) {
        nm = d;
        break;
      }
    int groups = mt / nm;
    // One 32-channel block per workgroup keeps the 8-row kernel at 90 VGPRs (5 waves per SIMD;
    // two blocks take 150), so it is preferred wherever it alone gives the 1024 workgroups.
    const bool tall = tiles_of(8) * unsigned(mt) >= 1024u;
    if (tall) {
      nm = 1;
      groups = mt;
    }
    const dim3 grid((tall ? tiles_of(8) : tiles) * unsigned(groups));
    if (fused) {
      if (nm != 1 || groups != 1) return -1;  // one 32-channel block (the checks above)
      if (tall)
        hipLaunchKernelGGL((unk::conv_tile_kernel<1, 8, true>), grid, dim3(8 * 64), 0, s, a, groups);
      else
        hipLaunchKernelGGL((unk::conv_tile_kernel<1, 4, true>), grid, dim3(4 * 64), 0, s, a, groups);
    } else if (tall) {
      hipLaunchKernelGGL((unk::conv_tile_kernel<1, 8>), grid, dim3(8 * 64), 0, s, a, groups);
    } else {
      switch (nm) {
        case 1: hipLaunchKernelGGL((unk::conv_tile_kernel<1, 4>), grid, dim3(4 * 64), 0, s, a, groups); break;
        case 2: hipLaunchKernelGGL((unk::conv_tile_kernel<2, 4>), grid, dim3(4 * 64), 0, s, a, groups); break;
        case 3: hipLaunchKernelGGL((unk::conv_tile_kernel<3, 4>), grid, dim3(4 * 64), 0, s, a, groups); break;
        default: hipLaunchKernelGGL((unk::conv_tile_kernel<4, 4>), grid, dim3(4 * 64), 0, s, a, groups); break;
      }
    }
    return unk::ok();
  }
  if (fused) return -1;  // the fused output step exists in the fp16 tiled kernel only
  const size_t waves = size_t(a.B) * a.H * ((a.W + 63) / 64) * ((a.cout + 63) / 64);
  const unsigned grid = unsigned((waves * 64 + unk::kBlock - 1) / unk::kBlock);
  if (taps == 9)
    hipLaunchKernelGGL((unk::conv_kernel<T, 9>), dim3(grid), dim3(unk::kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((unk::conv_kernel<T, 1>), dim3(grid), dim3(unk::kBlock), 0, s, a);
  if (unk::ok() != 0) return -1;
  // This kernel does not pool: the pooled copy comes from the separate pass.
  if (a.pool_out)
    return launch_unet_pool<T>(static_cast<const T*>(a.out), static_cast<T*>(a.pool_out), a.B, a.H, a.W, a.cout, s);
  return 0;
}
template int launch_unet_conv<float>(const ConvArgs&, int, hipStream_t);
template int launch_unet_conv<_Float16>(const ConvArgs&, int, hipStream_t);

int launch_unet_wgrad(const WgradArgs& a, int taps, hipStream_t s) {
  if (a.cout % 32 || a.chunk < 2 || a.chunk % 2 || (a.c1 && !a.in1) || (a.up0 && (a.H % 2 || a.W % 2))) return -1;
  const size_t npix = size_t(a.B) * a.H * a.W;
  const size_t chunks = (npix + a.chunk - 1) / a.chunk;
  const size_t waves = size_t(a.cout / 32) * ((a.c0 + a.c1 + 31) / 32) * taps * chunks;
  const unsigned grid = unsigned((waves * 64 + unk::kBlock - 1) / unk::kBlock);
  hipLaunchKernelGGL(unk::wgrad_kernel, dim3(grid), dim3(unk::kBlock), 0, s, a, taps);
  return unk::ok();
}

int launch_unet_bgrad(const float* g, float* db, int cout, size_t pixels, hipStream_t s) {
  if (cout > unk::kBlock) return -1;
  const unsigned grid = unsigned((pixels + 1023) / 1024);
  hipLaunchKernelGGL(unk::bgrad_kernel, dim3(grid ? grid : 1), dim3(unk::kBlock), 0, s, g, db, cout, pixels);
  return unk::ok();
}

template <typename T>
int launch_unet_pool(const T* in, T* out, int B, int H, int W, int C, hipStream_t s) {
  const size_t n = size_t(B) * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(unk::pool_kernel<T>, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, in, out, B, H, W, C);
  return unk::ok();
}
template int launch_unet_pool<float>(const float*, float*, int, int, int, int, hipStream_t);
template int launch_unet_pool<_Float16>(const _Float16*, _Float16*, int, int, int, int, hipStream_t);

int launch_unet_grad_skip_pool(const float* dcat, int ctot, int off, const float* dP, const float* E, float* dE, int B,
                               int H, int W, int C, hipStream_t s) {
  const size_t n = size_t(B) * H * W * C;
  hipLaunchKernelGGL(unk::grad_skip_pool_kernel, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, dcat, ctot, off, dP,
                     E, dE, B, H, W, C);
  return unk::ok();
}

int launch_unet_grad_up(const float* dcat, int ctot, const float* D, float* dD, int B, int H, int W, int C,
                        hipStream_t s) {
  const size_t n = size_t(B) * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(unk::grad_up_kernel, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, dcat, ctot, D, dD, B, H, W,
                     C);
  return unk::ok();
}

template <typename T>
int launch_unet_prep(const float* raw, T* x, int B, int Hs, int Ws, int H, int W, hipStream_t s) {
  const size_t n = size_t(B) * H * W;
  hipLaunchKernelGGL(unk::prep_kernel<T>, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, raw, x, B, Hs, Ws, H, W);
  return unk::ok();
}
template int launch_unet_prep<float>(const float*, float*, int, int, int, int, int, hipStream_t);
template int launch_unet_prep<_Float16>(const float*, _Float16*, int, int, int, int, int, hipStream_t);

int launch_unet_loss(const float* x, const float* y, int yc, const float* target, float* g, float* partial, int B, int H,
                     int W, int* blocks, hipStream_t s) {
  const size_t npix = size_t(B) * H * W;
  const unsigned grid = unk::blocks_for(npix);
  *blocks = int(grid);
  hipLaunchKernelGGL(unk::loss_kernel, dim3(grid), dim3(unk::kBlock), 0, s, x, y, yc, target, g, partial, npix,
                     1.0f / float(npix * 3));
  return unk::ok();
}



int launch_unet_transpose_w(const float* w, float* wt, int cout, int taps, int cin, hipStream_t s) {
  const size_t n = size_t(cout) * taps * cin;
  hipLaunchKernelGGL(unk::transpose_w_kernel, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, w, wt, cout, taps, cin);
  return unk::ok();
}

int launch_unet_to_half(const float* in, _Float16* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(unk::to_half_kernel, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, in, out, n);
  return unk::ok();
}

int launch_unet_adam(float* w, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(unk::adam_kernel, dim3(unk::blocks_for(n)), dim3(unk::kBlock), 0, s, w, m, v, g, n, lr, b1, b2, eps,
                     bc1, bc2);
  return unk::ok();
}

}  // namespace mrt
