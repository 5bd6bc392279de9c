// Denoiser state behind the mrt_denoiser handle (kernels in unet.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <vector>

#include "unet.hpp"

namespace mrt {

// A device buffer that only grows.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n);
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Activations of one forward pass at batch B, H x W (multiples of 16).
template <typename T>
struct UnetActs {
  int B = 0, H = 0, W = 0;
  DevBuf<T> x;                           // input [B][H][W][16]
  std::array<DevBuf<T>, kUnetLevels> p;  // pooled input of level l >= 1
  std::array<DevBuf<T>, kUnetLevels> ea, eb;
  std::array<DevBuf<T>, kUnetLevels - 1> da, db;
  DevBuf<T> y;                           // [B][H][W][32]
  int ensure(int B, int H, int W);
  void release();
};

struct Denoiser {
  std::array<UnetLayer, kUnetLayers> L;
  size_t n_logical = 0, n_w = 0, n_b = 0;   // parameter counts (logical; physical weights; biases)
  DevBuf<float> w, b, mw, vw, mb, vb, gw, gb, wt;
  DevBuf<_Float16> wh;                      // fp16 copy of w for inference
  bool half_dirty = true;
  float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  int step = 0;
  UnetActs<float> act;      // training
  UnetActs<_Float16> hact;  // inference
  // Training gradients.
  DevBuf<float> g_out, g_tmp_a, g_tmp_b, g_raw, g_tgt, loss_partial;
  std::array<DevBuf<float>, kUnetLevels - 1> g_cat;  // dX of decoder conv a: [up | skip] channels
  std::array<DevBuf<float>, kUnetLevels> g_pool;     // dX of encoder conv a of level l >= 1
  DevBuf<float> raw_dev, out_dev;
  hipStream_t stream = nullptr;

  Denoiser();
  ~Denoiser();
  int init(uint64_t seed);
  int set_logical(const float* p);
  int get_logical(float* p, bool grads);
  // Forward pass in T over raw features (device [B][Hs][Ws][10]) padded to H x W.
  // with_out = false stops before the 1x1 output layer. rgba (fp16 only, with_out = false): the
  // last convolution also runs the output layer and the finish step, writing the denoised image
  // [B][Hs][Ws][4] instead of its 32 channels.
  template <typename T>
  int forward(UnetActs<T>& A, const float* raw, int B, int Hs, int Ws, int H, int W, hipStream_t s,
              bool with_out = true, float* rgba = nullptr);
  // Loss, gradients into gw/gb (zeroed first) and, if lr > 0, one Adam step.
  int train(const float* d_raw, const float* d_target, int B, int H, int W, float lr_used, float* loss,
            hipStream_t s);
  // Denoised rgba of a device raw image (Hs x Ws x 10) into d_out (Hs x Ws x 4).
  int denoise(const float* d_raw, int Hs, int Ws, float* d_out, hipStream_t s);
  int ensure_half(hipStream_t s);
};

}  // namespace mrt
