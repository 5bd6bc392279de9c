// Denoiser host object and C ABI: the U-Net's layer table, forward schedules (fp32 for training,
// fp16 for inference), the backward schedule, Adam, and the logical <-> physical parameter layout.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/myrt.h"
#include "denoiser_internal.hpp"
#include "scene_internal.hpp"
#include "abi_guard.hpp"

namespace mrt {

template <typename T>
int DevBuf<T>::ensure(size_t n) {
  if (p && cap >= n) return MRT_OK;
  release();
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc denoiser buffer");
  cap = n;
  return MRT_OK;
}
template struct DevBuf<float>;
template struct DevBuf<_Float16>;

namespace {

constexpr int kEnc(int l, int s) { return 2 * l + s; }                          // s: 0 = a, 1 = b
constexpr int kDec(int l, int s) { return 2 * kUnetLevels + 2 * (kUnetLevels - 2 - l) + s; }
constexpr int kOut = kUnetLayers - 1;
constexpr int kYc = 32;  // physical output channels

inline int phys_level(int l) { return unet_phys(kUnetLevelCh[l]); }

std::array<UnetLayer, kUnetLayers> make_layers(size_t& n_log, size_t& n_w, size_t& n_b) {
  std::array<UnetLayer, kUnetLayers> L{};
  auto set = [](UnetLayer& u, int p0, int p0p, int p1, int p1p, int co, int taps, int relu) {
    u.parts[0] = p0;
    u.parts_phys[0] = p0p;
    u.parts[1] = p1;
    u.parts_phys[1] = p1p;
    u.cin = p0 + p1;
    u.cin_phys = p0p + p1p;
    u.cout = co;
    u.cout_phys = unet_phys(co);
    u.taps = taps;
    u.relu = relu;
  };
  for (int l = 0; l < kUnetLevels; ++l) {
    const int cin = l ? kUnetLevelCh[l - 1] : kUnetInCh, cinp = l ? phys_level(l - 1) : kUnetInPhys;
    set(L[kEnc(l, 0)], cin, cinp, 0, 0, kUnetLevelCh[l], 9, 1);
    set(L[kEnc(l, 1)], kUnetLevelCh[l], phys_level(l), 0, 0, kUnetLevelCh[l], 9, 1);
  }
  for (int l = kUnetLevels - 2; l >= 0; --l) {
    set(L[kDec(l, 0)], kUnetLevelCh[l + 1], phys_level(l + 1), kUnetLevelCh[l], phys_level(l), kUnetLevelCh[l], 9, 1);
    set(L[kDec(l, 1)], kUnetLevelCh[l], phys_level(l), 0, 0, kUnetLevelCh[l], 9, 1);
  }
  set(L[kOut], kUnetLevelCh[0], phys_level(0), 0, 0, kUnetOutCh, 1, 0);
  n_log = n_w = n_b = 0;
  for (UnetLayer& u : L) {
    u.w_log = n_log;
    u.b_log = n_log + size_t(u.cout) * u.taps * u.cin;
    n_log = u.b_log + u.cout;
    u.w_phys = n_w;
    n_w += size_t(u.cout_phys) * u.taps * u.cin_phys;
    u.b_phys = n_b;
    n_b += u.cout_phys;
  }
  return L;
}

inline size_t phys_index(const UnetLayer& u, int co, int tap, int ci) {
  const int cp = ci < u.parts[0] ? ci : u.parts_phys[0] + (ci - u.parts[0]);
  return u.w_phys + (size_t(co) * u.taps + tap) * u.cin_phys + cp;
}

uint64_t splitmix_next(uint64_t& x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int conv_call(const UnetLayer& u, const void* in0, int up0, const void* in1, const void* w, const float* bias,
              void* out, const void* mask, int relu, int B, int H, int W, bool half, hipStream_t s,
              void* pool_out = nullptr, const ConvArgs* fuse = nullptr) {
  ConvArgs a{};
  if (fuse) {  // the fused output step (raw, Hs, Ws, out_w, out_b, rgba)
    a.raw = fuse->raw;
    a.Hs = fuse->Hs;
    a.Ws = fuse->Ws;
    a.out_w = fuse->out_w;
    a.out_b = fuse->out_b;
    a.rgba = fuse->rgba;
  }
  a.pool_out = pool_out;
  a.in0 = in0;
  a.in1 = in1;
  a.w = w;
  a.bias = bias;
  a.out = out;
  a.mask = mask;
  a.c0 = u.parts_phys[0];
  a.c1 = u.parts_phys[1];
  a.up0 = up0;
  a.cout = u.cout_phys;
  a.relu = relu;
  a.B = B;
  a.H = H;
  a.W = W;
  const int rc = half ? launch_unet_conv<_Float16>(a, u.taps, s) : launch_unet_conv<float>(a, u.taps, s);
  return rc == 0 ? MRT_OK : fail(MRT_ERR_HIP, "denoiser convolution launch");
}

}  // namespace

template <typename T>
int UnetActs<T>::ensure(int B_, int H_, int W_) {
  if (B_ == B && H_ == H && W_ == W) return MRT_OK;
  B = B_;
  H = H_;
  W = W_;
  const size_t px = size_t(B) * H * W;
  int rc = x.ensure(px * kUnetInPhys);
  for (int l = 0; l < kUnetLevels && rc == MRT_OK; ++l) {
    const size_t pl = px >> (2 * l);
    const int c = unet_phys(kUnetLevelCh[l]);
    if (l > 0 && (rc = p[l].ensure(pl * unet_phys(kUnetLevelCh[l - 1]))) != MRT_OK) break;
    if ((rc = ea[l].ensure(pl * c)) != MRT_OK || (rc = eb[l].ensure(pl * c)) != MRT_OK) break;
    if (l < kUnetLevels - 1 && ((rc = da[l].ensure(pl * c)) != MRT_OK || (rc = db[l].ensure(pl * c)) != MRT_OK)) break;
  }
  if (rc == MRT_OK) rc = y.ensure(px * kYc);
  if (rc != MRT_OK) B = H = W = 0;
  return rc;
}

template <typename T>
void UnetActs<T>::release() {
  x.release();
  y.release();
  for (auto& b : p) b.release();
  for (auto& b : ea) b.release();
  for (auto& b : eb) b.release();
  for (auto& b : da) b.release();
  for (auto& b : db) b.release();
  B = H = W = 0;
}
template struct UnetActs<float>;
template struct UnetActs<_Float16>;

Denoiser::Denoiser() { L = make_layers(n_logical, n_w, n_b); }

Denoiser::~Denoiser() {
  for (DevBuf<float>* b : {&w, &b, &mw, &vw, &mb, &vb, &gw, &gb, &wt, &g_out, &g_tmp_a, &g_tmp_b, &g_raw, &g_tgt,
                           &loss_partial, &raw_dev, &out_dev})
    b->release();
  for (auto& x : g_cat) x.release();
  for (auto& x : g_pool) x.release();
  wh.release();
  act.release();
  hact.release();
}

int Denoiser::init(uint64_t seed) {
  // He-uniform over the logical parameters in packing order (one splitmix64 stream), biases 0.
  std::vector<float> p(n_logical, 0.0f);
  uint64_t x = seed;
  for (int li = 0; li < kUnetLayers; ++li) {
    const UnetLayer& u = L[li];
    float s = std::sqrt(6.0f / float(u.taps * u.cin));
    if (li == kOut) s = s * 0.1f;
    const size_t nw = size_t(u.cout) * u.taps * u.cin;
    for (size_t i = 0; i < nw; ++i) {
      const float r = float(splitmix_next(x) >> 40) * (1.0f / 16777216.0f);
      p[u.w_log + i] = (2.0f * r - 1.0f) * s;
    }
    for (int i = 0; i < u.cout; ++i) splitmix_next(x);  // biases take their stream slots
  }
  int rc;
  if ((rc = w.ensure(n_w)) || (rc = b.ensure(n_b)) || (rc = mw.ensure(n_w)) || (rc = vw.ensure(n_w)) ||
      (rc = mb.ensure(n_b)) || (rc = vb.ensure(n_b)) || (rc = gw.ensure(n_w)) || (rc = gb.ensure(n_b)) ||
      (rc = wh.ensure(n_w)))
    return rc;
  size_t max_layer = 0;
  for (const UnetLayer& u : L) max_layer = std::max(max_layer, size_t(u.cout_phys) * u.taps * u.cin_phys);
  if ((rc = wt.ensure(max_layer)) != MRT_OK) return rc;
  return set_logical(p.data());
}

int Denoiser::set_logical(const float* p) {
  std::vector<float> hw(n_w, 0.0f), hb(n_b, 0.0f);
  for (const UnetLayer& u : L) {
    for (int co = 0; co < u.cout; ++co) {
      for (int tap = 0; tap < u.taps; ++tap)
        for (int ci = 0; ci < u.cin; ++ci)
          hw[phys_index(u, co, tap, ci)] = p[u.w_log + (size_t(co) * u.taps + tap) * u.cin + ci];
      hb[u.b_phys + co] = p[u.b_log + co];
    }
  }
  hipError_t e = hipMemcpy(w.p, hw.data(), n_w * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b.p, hb.data(), n_b * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(mw.p, 0, n_w * 4);
  if (e == hipSuccess) e = hipMemset(vw.p, 0, n_w * 4);
  if (e == hipSuccess) e = hipMemset(mb.p, 0, n_b * 4);
  if (e == hipSuccess) e = hipMemset(vb.p, 0, n_b * 4);
  if (e != hipSuccess) return hip_fail(e, "uploading denoiser parameters");
  step = 0;
  half_dirty = true;
  return MRT_OK;
}

int Denoiser::get_logical(float* p, bool grads) {
  std::vector<float> hw(n_w), hb(n_b);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(hw.data(), grads ? gw.p : w.p, n_w * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hb.data(), grads ? gb.p : b.p, n_b * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading denoiser parameters");
  for (const UnetLayer& u : L)
    for (int co = 0; co < u.cout; ++co) {
      for (int tap = 0; tap < u.taps; ++tap)
        for (int ci = 0; ci < u.cin; ++ci)
          p[u.w_log + (size_t(co) * u.taps + tap) * u.cin + ci] = hw[phys_index(u, co, tap, ci)];
      p[u.b_log + co] = hb[u.b_phys + co];
    }
  return MRT_OK;
}

int Denoiser::ensure_half(hipStream_t s) {
  if (!half_dirty) return MRT_OK;
  if (launch_unet_to_half(w.p, wh.p, n_w, s)) return fail(MRT_ERR_HIP, "weight conversion");
  half_dirty = false;
  return MRT_OK;
}

template <typename T>
int Denoiser::forward(UnetActs<T>& A, const float* raw, int B, int Hs, int Ws, int H, int W, hipStream_t s,
                      bool with_out, float* rgba) {
  constexpr bool half = sizeof(T) == 2;
  if (H % 16 || W % 16 || B < 1 || H < 16 || W < 16) return fail(MRT_ERR_INVALID, "denoiser tiles are 16-pixel multiples");
  int rc = A.ensure(B, H, W);
  if (rc != MRT_OK) return rc;
  if (half && (rc = ensure_half(s)) != MRT_OK) return rc;
  const void* wbase = half ? static_cast<const void*>(wh.p) : static_cast<const void*>(w.p);
  auto wl = [&](int li) -> const void* {
    return half ? static_cast<const void*>(wh.p + L[li].w_phys) : static_cast<const void*>(w.p + L[li].w_phys);
  };
  (void)wbase;
  const bool fused = rgba != nullptr;
  if (fused && (!half || with_out)) return fail(MRT_ERR_INTERNAL, "fused denoiser steps need the fp16 pass");
  ConvArgs fout{};
  fout.raw = raw;
  fout.Hs = Hs;
  fout.Ws = Ws;
  fout.out_w = wl(kOut);
  fout.out_b = b.p + L[kOut].b_phys;
  fout.rgba = rgba;
  if (launch_unet_prep<T>(raw, A.x.p, B, Hs, Ws, H, W, s)) return fail(MRT_ERR_HIP, "denoiser input");
  for (int l = 0; l < kUnetLevels; ++l) {
    const int Hl = H >> l, Wl = W >> l;
    // The level's second convolution also writes the 2x2 average that feeds the next level.
    const T* src = l > 0 ? A.p[l].p : A.x.p;
    void* pooled = l + 1 < kUnetLevels ? static_cast<void*>(A.p[l + 1].p) : nullptr;
    if ((rc = conv_call(L[kEnc(l, 0)], src, 0, nullptr, wl(kEnc(l, 0)), b.p + L[kEnc(l, 0)].b_phys, A.ea[l].p, nullptr,
                        1, B, Hl, Wl, half, s)) != MRT_OK ||
        (rc = conv_call(L[kEnc(l, 1)], A.ea[l].p, 0, nullptr, wl(kEnc(l, 1)), b.p + L[kEnc(l, 1)].b_phys, A.eb[l].p,
                        nullptr, 1, B, Hl, Wl, half, s, pooled)) != MRT_OK)
      return rc;
  }
  for (int l = kUnetLevels - 2; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l;
    const T* src0 = l == kUnetLevels - 2 ? A.eb[l + 1].p : A.db[l + 1].p;
    if ((rc = conv_call(L[kDec(l, 0)], src0, 1, A.eb[l].p, wl(kDec(l, 0)), b.p + L[kDec(l, 0)].b_phys, A.da[l].p,
                        nullptr, 1, B, Hl, Wl, half, s)) != MRT_OK ||
        (rc = conv_call(L[kDec(l, 1)], A.da[l].p, 0, nullptr, wl(kDec(l, 1)), b.p + L[kDec(l, 1)].b_phys, A.db[l].p,
                        nullptr, 1, B, Hl, Wl, half, s, nullptr, fused && l == 0 ? &fout : nullptr)) != MRT_OK)
      return rc;
  }
  if (!with_out) return MRT_OK;
  return conv_call(L[kOut], A.db[0].p, 0, nullptr, wl(kOut), b.p + L[kOut].b_phys, A.y.p, nullptr, 0, B, H, W, half, s);
}
template int Denoiser::forward<float>(UnetActs<float>&, const float*, int, int, int, int, int, hipStream_t, bool,
                                      float*);
template int Denoiser::forward<_Float16>(UnetActs<_Float16>&, const float*, int, int, int, int, int, hipStream_t, bool,
                                         float*);

int Denoiser::train(const float* d_raw, const float* d_target, int B, int H, int W, float lr_used, float* loss,
                    hipStream_t s) {
  int rc = forward<float>(act, d_raw, B, H, W, H, W, s);
  if (rc != MRT_OK) return rc;
  const size_t px = size_t(B) * H * W;
  size_t max_cat = 0;
  for (int l = 0; l < kUnetLevels - 1; ++l)
    max_cat = std::max(max_cat, (px >> (2 * l)) * size_t(phys_level(l + 1) + phys_level(l)));
  if ((rc = g_out.ensure(px * kYc)) || (rc = g_tmp_a.ensure(px * 128)) || (rc = g_tmp_b.ensure(px * 128)) ||
      (rc = loss_partial.ensure(px / 256 + 1)))
    return rc;
  for (int l = 0; l < kUnetLevels - 1; ++l)
    if ((rc = g_cat[l].ensure((px >> (2 * l)) * size_t(phys_level(l + 1) + phys_level(l)))) != MRT_OK) return rc;
  for (int l = 1; l < kUnetLevels; ++l)
    if ((rc = g_pool[l].ensure((px >> (2 * l)) * size_t(phys_level(l - 1)))) != MRT_OK) return rc;
  (void)max_cat;
  hipError_t e = hipMemsetAsync(gw.p, 0, n_w * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(gb.p, 0, n_b * 4, s);
  if (e != hipSuccess) return hip_fail(e, "clearing gradients");
  int blocks = 0;
  if (launch_unet_loss(act.x.p, act.y.p, kYc, d_target, g_out.p, loss_partial.p, B, H, W, &blocks, s))
    return fail(MRT_ERR_HIP, "denoiser loss");

  // Weight and bias gradients of layer li from the masked output gradient g over input (in0, up0, in1).
  auto grads = [&](int li, const float* g, const float* in0, int up0, const float* in1, int Hl, int Wl) {
    const UnetLayer& u = L[li];
    WgradArgs a{};
    a.g = g;
    a.in0 = in0;
    a.in1 = in1;
    a.dw = gw.p + u.w_phys;
    a.c0 = u.parts_phys[0];
    a.c1 = u.parts_phys[1];
    a.up0 = up0;
    a.cout = u.cout_phys;
    a.B = B;
    a.H = Hl;
    a.W = Wl;
    a.chunk = 1024;
    if (launch_unet_wgrad(a, u.taps, s)) return fail(MRT_ERR_HIP, "denoiser weight gradient");
    if (launch_unet_bgrad(g, gb.p + u.b_phys, u.cout_phys, size_t(B) * Hl * Wl, s))
      return fail(MRT_ERR_HIP, "denoiser bias gradient");
    return int(MRT_OK);
  };
  // Input gradient of layer li: conv of g with the transposed, tap-mirrored weights.
  auto dx = [&](int li, const float* g, float* out, const float* mask, int Hl, int Wl) {
    const UnetLayer& u = L[li];
    if (launch_unet_transpose_w(w.p + u.w_phys, wt.p, u.cout_phys, u.taps, u.cin_phys, s))
      return fail(MRT_ERR_HIP, "weight transpose");
    UnetLayer t{};
    t.parts_phys[0] = u.cout_phys;
    t.parts_phys[1] = 0;
    t.cout_phys = u.cin_phys;
    t.taps = u.taps;
    return conv_call(t, g, 0, nullptr, wt.p, nullptr, out, mask, 0, B, Hl, Wl, false, s);
  };

  // Output layer and decoder, fine to coarse.
  if ((rc = grads(kOut, g_out.p, act.db[0].p, 0, nullptr, H, W)) != MRT_OK) return rc;
  if ((rc = dx(kOut, g_out.p, g_tmp_a.p, act.db[0].p, H, W)) != MRT_OK) return rc;
  for (int l = 0; l < kUnetLevels - 1; ++l) {
    const int Hl = H >> l, Wl = W >> l;
    const float* src0 = l == kUnetLevels - 2 ? act.eb[l + 1].p : act.db[l + 1].p;
    if ((rc = grads(kDec(l, 1), g_tmp_a.p, act.da[l].p, 0, nullptr, Hl, Wl)) != MRT_OK) return rc;
    if ((rc = dx(kDec(l, 1), g_tmp_a.p, g_tmp_b.p, act.da[l].p, Hl, Wl)) != MRT_OK) return rc;
    if ((rc = grads(kDec(l, 0), g_tmp_b.p, src0, 1, act.eb[l].p, Hl, Wl)) != MRT_OK) return rc;
    if ((rc = dx(kDec(l, 0), g_tmp_b.p, g_cat[l].p, nullptr, Hl, Wl)) != MRT_OK) return rc;
    if (launch_unet_grad_up(g_cat[l].p, phys_level(l + 1) + phys_level(l), src0, g_tmp_a.p, B, Hl, Wl,
                            phys_level(l + 1), s))
      return fail(MRT_ERR_HIP, "upsampling gradient");
  }
  // Encoder, coarse to fine; g_tmp_a holds the masked gradient of the level's conv b output.
  for (int l = kUnetLevels - 1; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l;
    if (l < kUnetLevels - 1 &&
        launch_unet_grad_skip_pool(g_cat[l].p, phys_level(l + 1) + phys_level(l), phys_level(l + 1), g_pool[l + 1].p,
                                   act.eb[l].p, g_tmp_a.p, B, Hl, Wl, phys_level(l), s))
      return fail(MRT_ERR_HIP, "skip gradient");
    if ((rc = grads(kEnc(l, 1), g_tmp_a.p, act.ea[l].p, 0, nullptr, Hl, Wl)) != MRT_OK) return rc;
    if ((rc = dx(kEnc(l, 1), g_tmp_a.p, g_tmp_b.p, act.ea[l].p, Hl, Wl)) != MRT_OK) return rc;
    const float* in = l ? act.p[l].p : act.x.p;
    if ((rc = grads(kEnc(l, 0), g_tmp_b.p, in, 0, nullptr, Hl, Wl)) != MRT_OK) return rc;
    if (l > 0 && (rc = dx(kEnc(l, 0), g_tmp_b.p, g_pool[l].p, nullptr, Hl, Wl)) != MRT_OK) return rc;
  }
  if (lr_used > 0.f) {
    ++step;
    const float bc1 = 1.0f - std::pow(beta1, float(step)), bc2 = 1.0f - std::pow(beta2, float(step));
    if (launch_unet_adam(w.p, mw.p, vw.p, gw.p, n_w, lr_used, beta1, beta2, eps, bc1, bc2, s) ||
        launch_unet_adam(b.p, mb.p, vb.p, gb.p, n_b, lr_used, beta1, beta2, eps, bc1, bc2, s))
      return fail(MRT_ERR_HIP, "denoiser Adam");
    half_dirty = true;
  }
  if (loss) {
    std::vector<float> part(blocks);
    e = hipMemcpyAsync(part.data(), loss_partial.p, size_t(blocks) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "reading the denoiser loss");
    double sum = 0;
    for (float v : part) sum += v;
    *loss = float(sum / double(px * 3));
  }
  return MRT_OK;
}

int Denoiser::denoise(const float* d_raw, int Hs, int Ws, float* d_out, hipStream_t s) {
  const int H = (Hs + 15) / 16 * 16, W = (Ws + 15) / 16 * 16;
  // The output layer and the finish step run inside the last convolution (the same arithmetic as
  // the separate output pass, on the fp16 values that pass would read).
  return forward<_Float16>(hact, d_raw, 1, Hs, Ws, H, W, s, false, d_out);
}

}  // namespace mrt

using namespace mrt;

namespace {

int upload_to(DevBuf<float>& dst, const float* src, size_t n) {
  int rc = dst.ensure(n);
  if (rc != MRT_OK) return rc;
  hipError_t e = hipMemcpy(dst.p, src, n * 4, hipMemcpyHostToDevice);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "uploading denoiser inputs");
}

bool bad_batch(int B, int H, int W) {
  return B < 1 || H < 16 || W < 16 || H % 16 || W % 16 || size_t(B) * H * W > (size_t(1) << 26);
}

}  // namespace

extern "C" {

int mrt_denoiser_create(mrt_denoiser* out, uint64_t seed) {
  MRT_GUARD_BEGIN
  if (!out) return fail(MRT_ERR_INVALID, "null output");
  auto* d = new Denoiser;
  const int rc = d->init(seed);
  if (rc != MRT_OK) {
    delete d;
    return rc;
  }
  *out = reinterpret_cast<mrt_denoiser>(d);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_denoiser_destroy(mrt_denoiser d) {
  MRT_GUARD_BEGIN
  delete reinterpret_cast<Denoiser*>(d);
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_denoiser_num_params(size_t* n) {
  MRT_GUARD_BEGIN
  if (!n) return fail(MRT_ERR_INVALID, "null output");
  Denoiser d0;  // the layer table only; no device memory
  *n = d0.n_logical;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_denoiser_get_params(mrt_denoiser dh, float* out, size_t n) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !out || n != d->n_logical) return fail(MRT_ERR_INVALID, "parameter count mismatch");
  return d->get_logical(out, false);
  MRT_GUARD_END
}

int mrt_denoiser_set_params(mrt_denoiser dh, const float* in, size_t n) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !in || n != d->n_logical) return fail(MRT_ERR_INVALID, "parameter count mismatch");
  return d->set_logical(in);
  MRT_GUARD_END
}

int mrt_denoiser_set_adam(mrt_denoiser dh, float lr, float beta1, float beta2, float eps) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return fail(MRT_ERR_INVALID, "bad Adam parameters");
  d->lr = lr;
  d->beta1 = beta1;
  d->beta2 = beta2;
  d->eps = eps;
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_denoiser_forward(mrt_denoiser dh, const float* raw, int B, int H, int W, int half, float* y) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !raw || !y) return fail(MRT_ERR_INVALID, "null argument");
  if (bad_batch(B, H, W)) return fail(MRT_ERR_INVALID, "batch dimensions must be multiples of 16");
  const size_t px = size_t(B) * H * W;
  int rc = upload_to(d->raw_dev, raw, px * kUnetInCh);
  if (rc != MRT_OK) return rc;
  std::vector<float> yp(px * kYc);
  if (half) {
    rc = d->forward<_Float16>(d->hact, d->raw_dev.p, B, H, W, H, W, nullptr);
    if (rc != MRT_OK) return rc;
    std::vector<_Float16> yh(px * kYc);
    hipError_t e = hipMemcpy(yh.data(), d->hact.y.p, yh.size() * 2, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the denoiser output");
    for (size_t i = 0; i < yh.size(); ++i) yp[i] = float(yh[i]);
  } else {
    rc = d->forward<float>(d->act, d->raw_dev.p, B, H, W, H, W, nullptr);
    if (rc != MRT_OK) return rc;
    hipError_t e = hipMemcpy(yp.data(), d->act.y.p, yp.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "reading the denoiser output");
  }
  for (size_t p = 0; p < px; ++p)
    for (int c = 0; c < 3; ++c) y[p * 3 + c] = yp[p * kYc + c];
  return MRT_OK;
  MRT_GUARD_END
}

int mrt_denoiser_gradient(mrt_denoiser dh, const float* raw, const float* target, int B, int H, int W, float* grad,
                          float* loss) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !raw || !target || !grad) return fail(MRT_ERR_INVALID, "null argument");
  if (bad_batch(B, H, W)) return fail(MRT_ERR_INVALID, "batch dimensions must be multiples of 16");
  const size_t px = size_t(B) * H * W;
  int rc = upload_to(d->raw_dev, raw, px * kUnetInCh);
  if (rc == MRT_OK) rc = upload_to(d->g_tgt, target, px * 3);
  if (rc == MRT_OK) rc = d->train(d->raw_dev.p, d->g_tgt.p, B, H, W, 0.f, loss, nullptr);
  if (rc == MRT_OK) rc = d->get_logical(grad, true);
  return rc;
  MRT_GUARD_END
}

int mrt_denoiser_train_device(mrt_denoiser dh, const float* d_raw, const float* d_target, int B, int H, int W,
                              float* loss, void* stream) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !d_raw || !d_target) return fail(MRT_ERR_INVALID, "null argument");
  if (bad_batch(B, H, W)) return fail(MRT_ERR_INVALID, "batch dimensions must be multiples of 16");
  return d->train(d_raw, d_target, B, H, W, d->lr, loss, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}

int mrt_denoiser_train(mrt_denoiser dh, const float* raw, const float* target, int B, int H, int W, float* loss) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !raw || !target) return fail(MRT_ERR_INVALID, "null argument");
  if (bad_batch(B, H, W)) return fail(MRT_ERR_INVALID, "batch dimensions must be multiples of 16");
  const size_t px = size_t(B) * H * W;
  int rc = upload_to(d->raw_dev, raw, px * kUnetInCh);
  if (rc == MRT_OK) rc = upload_to(d->g_tgt, target, px * 3);
  if (rc == MRT_OK) rc = d->train(d->raw_dev.p, d->g_tgt.p, B, H, W, d->lr, loss, nullptr);
  if (rc == MRT_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "denoiser training");
  }
  return rc;
  MRT_GUARD_END
}

int mrt_denoiser_denoise(mrt_denoiser dh, const float* raw, int H, int W, float* rgba) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !raw || !rgba) return fail(MRT_ERR_INVALID, "null argument");
  if (H < 1 || W < 1 || size_t(H) * W > (size_t(1) << 26)) return fail(MRT_ERR_INVALID, "bad image size");
  const size_t px = size_t(H) * W;
  int rc = upload_to(d->raw_dev, raw, px * kUnetInCh);
  if (rc == MRT_OK) rc = d->out_dev.ensure(px * 4);
  if (rc == MRT_OK) rc = d->denoise(d->raw_dev.p, H, W, d->out_dev.p, nullptr);
  if (rc != MRT_OK) return rc;
  hipError_t e = hipMemcpy(rgba, d->out_dev.p, px * 16, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MRT_OK : hip_fail(e, "reading the denoised image");
  MRT_GUARD_END
}

int mrt_denoiser_denoise_device(mrt_denoiser dh, const float* d_raw, int H, int W, float* d_rgba, void* stream) {
  MRT_GUARD_BEGIN
  Denoiser* d = reinterpret_cast<Denoiser*>(dh);
  if (!d || !d_raw || !d_rgba) return fail(MRT_ERR_INVALID, "null argument");
  if (H < 1 || W < 1 || size_t(H) * W > (size_t(1) << 26)) return fail(MRT_ERR_INVALID, "bad image size");
  return d->denoise(d_raw, H, W, d_rgba, static_cast<hipStream_t>(stream));
  MRT_GUARD_END
}

}  // extern "C"
