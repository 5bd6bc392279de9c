// Denoiser U-Net: layer table, tensor layout and kernel launchers shared by denoiser.cpp and
// unet.hip.
//
// Network: 5 levels of 32/48/64/96/128 channels; per level two 3x3 convolutions with bias and
// ReLU on the way down (2x2 average pooling between levels) and two on the way up (nearest
// upsampling of the coarser output concatenated with the level's encoder output); a 1x1 output
// convolution to 3 channels. Inputs are 10 features per pixel: log1p(radiance), albedo, normal and
// 1 / (1 + depth); the output is a log-space residual added to log1p(radiance).
//
// Tensors are NHWC with physical channel counts that are multiples of 32 for every convolution
// output (48 is stored as 64 with zero channels) and 16 for the input. Weights are stored
// [cout_phys][tap][cin_phys] with zeros at padded positions; the zeros stay zero under training
// because padded channels carry zero activations and zero gradients.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mrt {

constexpr int kUnetLevels = 5;
constexpr int kUnetLayers = 2 * kUnetLevels + 2 * (kUnetLevels - 1) + 1;  // 19
constexpr int kUnetLevelCh[kUnetLevels] = {32, 48, 64, 96, 128};
constexpr int kUnetInCh = 10;        // logical input features
constexpr int kUnetInPhys = 16;      // physical input channels
constexpr int kUnetOutCh = 3;

inline int unet_phys(int c) { return (c + 31) / 32 * 32; }

struct UnetLayer {
  int parts[2];       // logical input channel parts (concatenation order); parts[1] = 0 if one
  int parts_phys[2];  // physical
  int cin, cin_phys;  // totals
  int cout, cout_phys;
  int taps;           // 9 or 1
  int relu;
  size_t w_log, b_log;    // offsets in the packed logical parameter vector
  size_t w_phys, b_phys;  // offsets in the physical weight / bias arrays
};

// Convolution over one or two NHWC sources. in0 may be read at half resolution (nearest
// upsampling); in1, if present, follows in0 in the channel order.
struct ConvArgs {
  const void* in0;
  const void* in1;
  const void* w;        // [cout][taps][c0 + c1]
  const float* bias;    // [cout] or null
  void* out;            // [B][H][W][cout]
  const void* mask;     // if set, [B][H][W][cout]: the result is zeroed where mask <= 0
  void* pool_out;       // if set, also the 2x2 average of the result, [B][H/2][W/2][cout]
  int c0, c1, up0;
  int cout;
  int relu;
  int B, H, W;
  // fp16 3x3 tiles only: with out_w (and cout == 32) the 1x1 output convolution (out_w [3][32],
  // out_b [3]) and the finish step against the raw feature image raw [B][Hs][Ws][kUnetInCh] run on
  // the tile; rgba [B][Hs][Ws] is written and `out` is not.
  const float* raw;
  int Hs, Ws;
  const void* out_w;
  const float* out_b;
  float* rgba;
};

// Weight gradient dW[co][tap][ci] += sum over pixels of g[p][co] * x_tap[p][ci], where x_tap is the
// convolution input (ConvArgs addressing: in0/in1/up0) shifted by the tap.
struct WgradArgs {
  const float* g;       // [B][H][W][cout]
  const float* in0;
  const float* in1;
  float* dw;            // [cout][taps][c0 + c1]
  int c0, c1, up0;
  int cout;
  int B, H, W;
  int chunk;            // pixels per wave
};

template <typename T>
int launch_unet_conv(const ConvArgs& a, int taps, hipStream_t s);
int launch_unet_wgrad(const WgradArgs& a, int taps, hipStream_t s);
int launch_unet_bgrad(const float* g, float* db, int cout, size_t pixels, hipStream_t s);
template <typename T>
int launch_unet_pool(const T* in, T* out, int B, int H, int W, int C, hipStream_t s);
// dE = (dcat[..., off:off+C] + up(dP) / 4) * (E > 0), dP at half resolution (null: none).
int launch_unet_grad_skip_pool(const float* dcat, int ctot, int off, const float* dP, const float* E, float* dE, int B,
                               int H, int W, int C, hipStream_t s);
// dD (half resolution) = (sum of the 2x2 block of dcat[..., 0:C]) * (D > 0).
int launch_unet_grad_up(const float* dcat, int ctot, const float* D, float* dD, int B, int H, int W, int C,
                        hipStream_t s);
// Input tensor [B][H][W][16] from raw features [B][Hs][Ws][10] (rgb, albedo, normal, depth) read
// with clamped coordinates (replicated border where H, W exceed Hs, Ws).
template <typename T>
int launch_unet_prep(const float* raw, T* x, int B, int Hs, int Ws, int H, int W, hipStream_t s);
// Loss and output gradient: z = x[..., 0:3] + y, L1 against log1p(target); g = sign(z - t) / n.
// Writes per-block partial sums of |z - t| to partial[blocks].
int launch_unet_loss(const float* x, const float* y, int yc, const float* target, float* g, float* partial,
                     int B, int H, int W, int* blocks, hipStream_t s);
int launch_unet_transpose_w(const float* w, float* wt, int cout, int taps, int cin, hipStream_t s);
int launch_unet_to_half(const float* in, _Float16* out, size_t n, hipStream_t s);
int launch_unet_adam(float* w, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t s);

}  // namespace mrt
