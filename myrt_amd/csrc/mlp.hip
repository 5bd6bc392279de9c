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

#include <algorithm>
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

// sin(2 pi r) of r revolutions: the hardware's transcendental unit takes its argument in
// revolutions; the fractional part keeps it in its domain. The features are rounded to bf16
// next, far coarser than the unit's error.
__device__ inline float sin_rev(float r) { return __builtin_amdgcn_sinf(__builtin_amdgcn_fractf(r)); }

// The 64 network inputs of a query, by input slot. The query's 13 components are
//   U = (position in the cache box (3), direction (3), normal (3), roughness, albedo (3)),
// and an input is sin(2^k pi U[c]) or cos(2^k pi U[c]) (2^(k-1) U[c] revolutions, cos a quarter
// revolution later), or U[c] itself. Position takes octaves 0-5, direction and normal 0-1. The
// slots are ordered so that lane halves h = 0 and h = 1 of a B fragment (slots 16 kb + 8 h + i)
// evaluate the same octave and phase on different components in six of the eight slot rows:
//   slots  0- 7 position x, octaves 2-5      slots  8-15 position y, octaves 2-5
//          16-23 position x, y, octaves 0-1        24-31 position z, direction x, octaves 0-1
//          32-39 direction y, z, octaves 0-1       40-47 normal x, y, octaves 0-1
//          48-55 position z, octaves 2-5           56-63 normal z, octaves 0-1; roughness; albedo
struct Slot {
  int c;       // component of U
  int k;       // octave
  bool cos;
  bool ident;  // U[c] itself
};
__host__ __device__ constexpr Slot lo_pair(int c0, int c1, int i) { return Slot{i < 4 ? c0 : c1, (i >> 1) & 1, (i & 1) != 0, false}; }
__host__ __device__ constexpr Slot slot_of(int s) {
  return s < 8    ? Slot{0, 2 + (s >> 1), (s & 1) != 0, false}
         : s < 16 ? Slot{1, 2 + ((s - 8) >> 1), (s & 1) != 0, false}
         : s < 24 ? lo_pair(0, 1, s - 16)
         : s < 32 ? lo_pair(2, 3, s - 24)
         : s < 40 ? lo_pair(4, 5, s - 32)
         : s < 48 ? lo_pair(6, 7, s - 40)
         : s < 56 ? Slot{2, 2 + ((s - 48) >> 1), (s & 1) != 0, false}
         : s < 60 ? lo_pair(8, 8, s - 56)
                  : Slot{9 + (s - 60), 0, false, true};
}

typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

__device__ inline uint32_t pk_bf16(float lo, float hi) {
  typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
  const bf16x2 v = {(__bf16)lo, (__bf16)hi};  // v_cvt_pk_bf16_f32: round to nearest even
  return __builtin_bit_cast(uint32_t, v);
}

// Encodes the query of lane (s, h) (query index qi) into its four B fragments: lane h holds
// input slots 16 kb + 8 h + (0..7) of fragment kb. Queries at or past n encode as zeros.
__device__ inline void encode_fragments(const CacheQuery* qs, uint32_t qi, uint32_t n, const CacheParams& P,
                                        bf16x8 (&b)[4], int lane) {
  const bool hi = (lane >> 5) != 0;
  float4 q0 = make_float4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
  const bool live = qi < n;
  if (live) {
    const float4* qp = reinterpret_cast<const float4*>(qs + qi);
    q0 = qp[0];
    q1 = qp[1];
    q2 = qp[2];
    q3 = qp[3];
  }
  const float U[13] = {(q0.x - P.lo[0]) * P.inv_ext[0], (q0.y - P.lo[1]) * P.inv_ext[1],
                       (q0.z - P.lo[2]) * P.inv_ext[2], q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, q0.w, q3.x, q3.y, q3.z};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
      float y[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const Slot a = slot_of(16 * kb + i + e), b8 = slot_of(16 * kb + 8 + i + e);
        const float x = hi ? U[b8.c] : U[a.c];
        const float ma = float(1 << a.k) * 0.5f, mb = float(1 << b8.k) * 0.5f;
        const float pa = a.cos ? 0.25f : 0.0f, pb = b8.cos ? 0.25f : 0.0f;
        const float t = (ma == mb && pa == pb) ? fmaf(x, ma, pa) : fmaf(x, hi ? mb : ma, hi ? pb : pa);
        float v = sin_rev(t);
        if (a.ident || b8.ident) v = (hi ? b8.ident : a.ident) ? x : v;
        y[e] = live ? v : 0.0f;
      }
      w[i >> 1] = pk_bf16(y[0], y[1]);
    }
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    b[kb] = __builtin_bit_cast(bf16x8, v);
  }
}

// Copies `count` 16-byte records to LDS with THREADS threads, src(rec) giving each record's global
// address: every load of a thread is issued before the first store waits, so the staging costs one
// memory round trip instead of one per record.
template <int THREADS, int COUNT, typename Src>
__device__ inline void stage_records(u32x4* dst, Src src) {
  constexpr int kPer = (COUNT + THREADS - 1) / THREADS;
  u32x4 v[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int rec = int(threadIdx.x) + i * THREADS;
    if (rec < COUNT) v[i] = *src(rec);
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int rec = int(threadIdx.x) + i * THREADS;
    if (rec < COUNT) dst[rec] = v[i];
  }
}

// Global address of forward fragment record (l, mb, kb, lane): W[l][32 mb + lane % 32][16 kb + 8 (lane / 32) ..+8].
__device__ inline const u32x4* forward_record_src(const CacheParams& P, int rec) {
  const int lane = rec & 63, kb = (rec >> 6) & 3, lm = rec >> 8, l = lm >> 1, mb = lm & 1;
  return reinterpret_cast<const u32x4*>(P.w + (size_t(l) * kW + 32 * mb + (lane & 31)) * kW + 16 * kb + 8 * (lane >> 5));
}

// ---- inference: weights in LDS, activations in registers -----------------------------------
//
// A wave runs two tiles of 32 queries at once. The B operand of a product (8 input neurons of one
// query per lane) is built in registers: the previous layer's D holds, in lane (s, h), neurons
// 8j + 4h + (0..3) of query s, so lane h = 0 has rows 0-3 and 8-11 of every 16-row block and
// lane h = 1 rows 4-7 and 12-15. One v_permlane32_swap per register pair gives lane h = 0 rows
// 0-7 and lane h = 1 rows 8-15 of the block: exactly the next layer's B fragment. Weights sit in
// LDS in fragment order (one 16-byte record per lane, consecutive lanes consecutive records), 44 KB
// per workgroup: five 64 x 64 layers and the output layer's first 32 rows.
constexpr int kFragRecords = (5 * 8 + 4) * 64;  // 16-byte A-fragment records of all layers

// Record index of the A fragment (layer l, row block mb, k block kb) of lane `lane`.
__device__ inline int frag_record(int l, int mb, int kb, int lane) { return ((l * 2 + mb) * 4 + kb) * 64 + lane; }

// ReLU on two packed bf16: a bf16 with its sign bit set is a negative int16, so a signed 16-bit
// max with 0 zeroes exactly the negative values (and -0).
__device__ inline uint32_t relu_pk(uint32_t x) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  const s16x2 v = __builtin_bit_cast(s16x2, x);
  const s16x2 z = {0, 0};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, z));
}

// D (64 neurons x 32 queries, in d0 = rows 0-31 and d1 = rows 32-63) to the next layer's four B
// fragments: MODE 0 as is, 1 through ReLU.
template <int MODE = 1>
__device__ inline void to_fragments(const f32x16& d0, const f32x16& d1, bf16x8 (&b)[4]) {
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const f32x16& d = kb < 2 ? d0 : d1;
    const int v = 8 * (kb & 1);  // d[v..v+3]: rows 16kb' + 4h + (0..3); d[v+4..v+7]: rows 16kb' + 8 + 4h + (0..3)
    uint32_t x0 = pk_bf16(d[v + 0], d[v + 1]), x1 = pk_bf16(d[v + 2], d[v + 3]);
    uint32_t y0 = pk_bf16(d[v + 4], d[v + 5]), y1 = pk_bf16(d[v + 6], d[v + 7]);
    if (MODE == 1) {
      x0 = relu_pk(x0);
      x1 = relu_pk(x1);
      y0 = relu_pk(y0);
      y1 = relu_pk(y1);
    }
    // Lanes 32-63 of x trade with lanes 0-31 of y: h = 0 keeps x (rows +0..3) and gets the other
    // half's x (rows +4..7); h = 1 gets the other half's y (rows +8..11) and keeps its y (+12..15).
    const auto s0 = __builtin_amdgcn_permlane32_swap(x0, y0, false, false);
    const auto s1 = __builtin_amdgcn_permlane32_swap(x1, y1, false, false);
    const u32x4 w = {s0[0], s1[0], s0[1], s1[1]};
    b[kb] = __builtin_bit_cast(bf16x8, w);
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void infer_kernel(CacheParams P, const CacheQuery* qs, uint32_t n_host,
                                                    const uint32_t* n_dev, float* out) {
  __shared__ u32x4 wlds[kFragRecords];
  // Weights to LDS in fragment order: record (l, mb, kb, lane) = W[l][32 mb + lane % 32][16 kb + 8 (lane / 32) ..+8].
  stage_records<256, kFragRecords>(wlds, [&](int rec) { return forward_record_src(P, rec); });
  __syncthreads();
  const int lane = __lane_id();
  const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;  // a device count is capped by the buffer size
  const uint32_t pairs = (n + 63) / 64;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), waves = gridDim.x * 4;
  for (uint32_t pr = wave; pr < pairs; pr += waves) {
    // The weight records do not change inside the loop; an opaque zero per layer keeps the
    // compiler from hoisting all 44 fragment reads (176 registers) out of it.
    int wofs = 0;
    bf16x8 b[2][4];
    encode_fragments(qs, pr * 64 + uint32_t(lane & 31), n, P, b[0], lane);
    encode_fragments(qs, pr * 64 + 32 + uint32_t(lane & 31), n, P, b[1], lane);
    f32x16 d[2][2];
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
      asm volatile("" : "+v"(wofs));
      const u32x4* wl = wlds + wofs;
      const int mbs = l < kLayers - 1 ? 2 : 1;  // the output layer needs rows 0-2 only
#pragma unroll
      for (int t = 0; t < 2; ++t) d[t][0] = d[t][1] = f32x16{};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int mb = 0; mb < mbs; ++mb) {
          const bf16x8 a = __builtin_bit_cast(bf16x8, wl[frag_record(l, mb, kb, lane)]);
#pragma unroll
          for (int t = 0; t < 2; ++t) d[t][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[t][kb], d[t][mb], 0, 0, 0);
        }
      if (l < kLayers - 1) {
#pragma unroll
        for (int t = 0; t < 2; ++t) to_fragments(d[t][0], d[t][1], b[t]);
      }
    }
    // Outputs are neurons 0-2: lanes 0-31 hold them in d[t][0][0..2] for query lane.
    if (lane < 32) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const uint32_t qi = pr * 64 + 32 * t + uint32_t(lane);
        if (qi < n) {
          out[3 * size_t(qi) + 0] = d[t][0][0];
          out[3 * size_t(qi) + 1] = d[t][0][1];
          out[3 * size_t(qi) + 2] = d[t][0][2];
        }
      }
    }
  }
}

// ---- training: forward, backward and weight gradients of one tile per wave ----------------------
//
// Weights W (forward, 44 KB) and W^T (backward: layers 1-5, 34 KB) sit in LDS in fragment order.
// The forward and backward chains keep their operands in registers (to_fragments), and so do the
// layer inputs H_l (the forward B fragments), whose nonzero entries are also the ReLU derivative
// (keep_active). In the backward pass each layer's weight gradient dW_l = G_l H_l^T over the
// workgroup's 8 tiles (256 queries) is one more set of products: every wave writes its H_l and G_l
// transposed, [neuron][query], into the workgroup's two LDS tiles, where the run of 8 queries of
// one neuron is the 16-byte operand that v_mfma_f32_16x16x32_bf16 takes along K; then each wave
// sums two 16 x 16 blocks of dW_l over all 256 queries (the output layer: one block each for waves
// 0-3) and adds them to the workgroup's partial in HBM. No block is summed by two waves, so the
// reduction needs no atomics, and nothing per query goes to HBM: the optimiser step sums one partial
// per workgroup.
constexpr int kTransRecords = (4 * 8 + 2) * 64;  // W^T fragments: layers 1-4 full, layer 5 k block 0
constexpr int kTrainWaves = 8;
constexpr int kTrainThreads = 64 * kTrainWaves;
// Row stride (bf16) of the transposed tiles: 256 queries + 8. Rows 132 dwords apart put the 8 rows
// of a 16-byte read phase in distinct banks, and rows 4, 8 and 12 apart (the row groups of one
// 32-bit write) 16, 32 and 48 banks apart.
constexpr int kTQ = kTrainWaves * 32 + 8;
constexpr int kOutRows = 3;  // rows of the output layer

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// Offset (floats) of dW block (rows 16 ob.., columns 16 ib..) in a layer's partial: 16 blocks of
// 64 lane records of 4 floats (lane (r, g) holds rows 16 ob + 4 g + (0..3) of column 16 ib + r).
__device__ inline int partial_block(int ob, int ib) { return (ob * 4 + ib) * 256; }

__device__ inline int trans_record(int l, int mb, int kb, int lane) {
  return l < 5 ? (((l - 1) * 2 + mb) * 4 + kb) * 64 + lane : (32 + mb) * 64 + lane;
}

// The ReLU derivative: keeps the values of G (packed bf16 fragments) whose layer input H (the
// same positions of the forward fragments, after ReLU, so never negative) is nonzero. A positive
// pre-activation rounds to a nonzero bf16 unless it is below 2^-133.
__device__ inline void keep_active(bf16x8 (&g)[4], const bf16x8 (&hf)[4]) {
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    u32x4 gw = __builtin_bit_cast(u32x4, g[kb]);
    const u32x4 hw = __builtin_bit_cast(u32x4, hf[kb]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const u16x2 one = {1, 1};
      // A scalar copy first: clang's __builtin_bit_cast of a vector element reads element 0.
      const uint32_t hv = hw[i];
      const u16x2 nz = __builtin_elementwise_min(__builtin_bit_cast(u16x2, hv), one);  // 0 or 1 per half
      gw[i] &= __builtin_bit_cast(uint32_t, u16x2{0, 0} - nz);                              // 0 or 0xffff
    }
    g[kb] = __builtin_bit_cast(bf16x8, gw);
  }
}

// Writes a tile's four B fragments (lane (s, h): neurons 16 kb + 8 h + (0..7) of query s) to
// t[neuron][query] at the queries 32 wave + s, as 32-bit pairs of queries: lanes s and s ^ 1 trade
// halves so that the even lane writes neurons +0..3 and the odd lane neurons +4..7 of the pair
// (s & ~1, s | 1). With rows 132 dwords apart, the four row groups of one write (even and odd lane,
// h = 0 and 1: rows +0, +4, +8, +12) fall on four disjoint quarters of the banks.
__device__ inline void write_transposed(const bf16x8 (&b)[4], uint16_t* t, int wave, int lane) {
  const int s = lane & 31, h = lane >> 5;
  const bool odd = (lane & 1) != 0;
  // v_perm selectors over (src0 = received word, src1 = own word): the low halves (first word) and
  // the high halves (second word) of the pair, own query first on the even lane, second on the odd.
  const uint32_t sel_lo = odd ? 0x01000504u : 0x05040100u;
  const uint32_t sel_hi = odd ? 0x03020706u : 0x07060302u;
  uint32_t* row0 = reinterpret_cast<uint32_t*>(t + (8 * h + 4 * (odd ? 1 : 0)) * kTQ + 32 * wave + (s & ~1));
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const u32x4 w = __builtin_bit_cast(u32x4, b[kb]);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      // Word p holds neurons 2p, 2p + 1; word p + 2 neurons 2p + 4, 2p + 5.
      const uint32_t send = odd ? w[p] : w[p + 2];
      const uint32_t own = odd ? w[p + 2] : w[p];
      const uint32_t recv = uint32_t(__builtin_amdgcn_mov_dpp(int(send), 0xB1, 0xF, 0xF, false));  // lane ^ 1
      uint32_t* r = row0 + (16 * kb + 2 * p) * (kTQ / 2);
      r[0] = __builtin_amdgcn_perm(recv, own, sel_lo);
      r[kTQ / 2] = __builtin_amdgcn_perm(recv, own, sel_hi);
    }
  }
}

__global__ __launch_bounds__(kTrainThreads) void train_kernel(CacheParams P, const CacheQuery* qs, const float* targets,
                                                             uint32_t n, float grad_scale, float* loss_out,
                                                             float* partial) {
  __shared__ u32x4 wf[kFragRecords];
  __shared__ u32x4 wb[kTransRecords];
  __shared__ __attribute__((aligned(16))) uint16_t th[kW * kTQ];  // H_l^T
  __shared__ __attribute__((aligned(16))) uint16_t tg[kW * kTQ];  // G_l^T
  stage_records<kTrainThreads, kFragRecords>(wf, [&](int rec) { return forward_record_src(P, rec); });
  stage_records<kTrainThreads, kTransRecords>(wb, [&](int rec) {
    const int lane = rec & 63;
    int l, mb, kb;
    if (rec < 32 * 64) {
      kb = (rec >> 6) & 3;
      const int lm = rec >> 8;
      l = 1 + (lm >> 1);
      mb = lm & 1;
    } else {
      l = 5;
      kb = 0;
      mb = (rec >> 6) - 32;
    }
    return reinterpret_cast<const u32x4*>(P.wt + (size_t(l) * kW + 32 * mb + (lane & 31)) * kW + 16 * kb +
                                          8 * (lane >> 5));
  });
  __syncthreads();
  // The wave index in a scalar register: addresses and branches that depend on it are uniform.
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const uint32_t tiles = (n + 31) / 32;
  // Every wave runs the same number of rounds (the workgroup synchronises inside each); a wave past
  // the last tile encodes zeros, whose gradients are zero.
  const uint32_t per_round = gridDim.x * kTrainWaves;
  const uint32_t rounds = (tiles + per_round - 1) / per_round;
  float* part = partial + size_t(blockIdx.x) * kLayers * kW * kW;
  for (uint32_t round = 0; round < rounds; ++round) {
    const uint32_t t = round * per_round + blockIdx.x * kTrainWaves + uint32_t(wave);
    int wofs = 0;
    // The lane index through an opaque copy: lane-dependent constants (the input encoding's
    // per-half choices, LDS addresses) are formed in each round instead of kept live across rounds.
    int lane = __lane_id();
    asm volatile("" : "+v"(lane));
    const int s = lane & 31, h = lane >> 5;
    const uint32_t qi = t * 32 + uint32_t(s);
    bf16x8 b[4];
    encode_fragments(qs, qi, n, P, b, lane);
    bf16x8 hin[kLayers][4];       // H_l, the input of layer l, as B fragments (H_0 is encoded again
                                  // when the backward pass reaches it, rather than held throughout)
    f32x16 d0, d1;
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
      if (l > 0) {
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) hin[l][kb] = b[kb];
      }
      // An opaque zero per layer keeps each layer's fragment reads in place instead of hoisted to
      // the top of the loop.
      asm volatile("" : "+v"(wofs));
      const u32x4* wl = wf + wofs;
      const int mbs = l < kLayers - 1 ? 2 : 1;
      d0 = d1 = f32x16{};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wl[frag_record(l, 0, kb, lane)]), b[kb],
                                                      d0, 0, 0, 0);
        if (mbs > 1)
          d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wl[frag_record(l, 1, kb, lane)]),
                                                        b[kb], d1, 0, 0, 0);
      }
      if (l < kLayers - 1) to_fragments(d0, d1, b);
    }
    // Loss: relative L2 (y - t)^2 / (sg(y)^2 + 0.01) per channel; dL/dy = 2 (y - t) / (y^2 + 0.01).
    // Outputs y (rows 0-2) are d0[0..2] of lanes h = 0; G_5 as fragments: kb 0, lane h = 0 slots 0-2.
    float loss = 0.f;
    float g[3] = {0.f, 0.f, 0.f};
    if (h == 0 && qi < n) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float y = d0[c], tg_ = targets[3 * size_t(qi) + c];
        const float den = y * y + 0.01f;
        const float e = y - tg_;
        loss += e * e / den;
        g[c] = 2.0f * e / den * grad_scale;
      }
    }
    // Butterfly sum over the wave (ds_bpermute addressed from the round's lane copy).
    for (int off = 32; off > 0; off >>= 1)
      loss += __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ off) << 2, __float_as_int(loss)));
    if (lane == 0 && t < tiles) loss_out[t] = loss;
    {
      const u32x4 z = {0u, 0u, 0u, 0u};
      const u32x4 g0 = {pk_bf16(g[0], g[1]), pk_bf16(g[2], 0.f), 0u, 0u};
      b[0] = __builtin_bit_cast(bf16x8, h == 0 ? g0 : z);
      b[1] = b[2] = b[3] = __builtin_bit_cast(bf16x8, z);
    }
    // Backward, l = 5..0: dW_l = G_l H_l^T over the workgroup's queries, and G_(l-1) = (W_l^T G_l)
    // masked by the ReLU pattern of H_l.
#pragma unroll
    for (int l = kLayers - 1; l >= 0; --l) {
      // A fresh opaque lane copy per layer, so that this layer's addresses are formed here and not
      // held in registers through the forward pass.
      int ln = __lane_id();
      asm volatile("" : "+v"(ln));
      const int r = ln & 15, q8 = 8 * (ln >> 4);  // 16x16x32 operand lane: row r, queries q8..q8+7
      if (l == 0) encode_fragments(qs, t * 32 + uint32_t(ln & 31), n, P, hin[0], ln);
      write_transposed(hin[l], th, wave, ln);
      write_transposed(b, tg, wave, ln);
      if (l > 0) {
        asm volatile("" : "+v"(wofs));
        const u32x4* wt = wb + wofs;
        const int kbs = l == kLayers - 1 ? 1 : 4;  // G_5 has rows 0-2 only
        d0 = d1 = f32x16{};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
          if (kb >= kbs) break;
          d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wt[trans_record(l, 0, kb, ln)]),
                                                        b[kb], d0, 0, 0, 0);
          d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wt[trans_record(l, 1, kb, ln)]),
                                                        b[kb], d1, 0, 0, 0);
        }
        to_fragments<0>(d0, d1, b);
        keep_active(b, hin[l]);
      }
      __syncthreads();
      // This wave's blocks of dW_l: rows 16 ob + (0..15), columns 16 ib + (0..15) and, for the
      // hidden layers, 16 (ib + 1) + (0..15). D of the 16x16 product: lane (r, g) holds rows
      // 16 ob + 4 g + v, column r.
      const bool out_layer = l == kLayers - 1;
      if (!out_layer || wave < 4) {
        const int ob = out_layer ? 0 : wave >> 1;
        const int ib = out_layer ? wave : 2 * (wave & 1);
        const uint16_t* ga = tg + (16 * ob + r) * kTQ + q8;
        const uint16_t* hb = th + (16 * ib + r) * kTQ + q8;
        // The partial of this wave's blocks: records of 16 bytes in the product's own lane layout
        // (one coalesced store per block); the optimiser step maps them back to [row][column].
        // Earlier rounds' sums are loaded before the products, so that no branch separates the
        // products from the add and the store that read their results (a conditional branch there
        // has come out of the compiler without the wait states the results need).
        f32x4* p0 = reinterpret_cast<f32x4*>(part + l * kW * kW + partial_block(ob, ib)) + ln;
        f32x4 prev0 = f32x4{}, prev1 = f32x4{};
        if (round > 0) {
          prev0 = p0[0];
          if (!out_layer) prev1 = p0[64];
        }
        asm volatile("" : "+v"(prev0), "+v"(prev1));
        f32x4 c0 = f32x4{}, c1 = f32x4{};
#pragma unroll
        for (int w2 = 0; w2 < kTrainWaves; ++w2) {
          // Operand reads in two groups of four tiles (a compiler-level fence between them) bound
          // the registers they occupy.
          if (w2 == kTrainWaves / 2) asm volatile("" ::: "memory");
          const bf16x8 a = ld8(ga + 32 * w2);
          c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ld8(hb + 32 * w2), c0, 0, 0, 0);
          if (!out_layer) c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ld8(hb + 16 * kTQ + 32 * w2), c1, 0, 0, 0);
        }
        p0[0] = c0 + prev0;
        if (!out_layer) p0[64] = c1 + prev1;
      }
      __syncthreads();
    }
  }
}

// Sums the training workgroups' partials of 64 parameters per workgroup (4 slices of the partials
// per parameter, combined in LDS) and takes the Adam step.
// Launched over the partial slots of layers 0-4 and of the output layer's first four blocks (rows
// 0-15, of which rows 0-2 are parameters; the rest of the output layer is padding and stays zero).
// Slot (block b, lane, v) of a layer is parameter row 16 (b / 4) + 4 (lane / 16) + v, column
// 16 (b % 4) + lane % 16.
__global__ __launch_bounds__(256) void adam_kernel(CacheParams P, const float* partial, uint32_t chunks, float lr,
                                                   float beta1, float beta2, float eps, float bc1, float bc2,
                                                   float* grad_out) {
  __shared__ float part[4][64];
  const uint32_t q = blockIdx.x * 64 + (threadIdx.x & 63), slice = threadIdx.x >> 6;
  const uint32_t l = q / (kW * kW), slot = q % (kW * kW), blk = slot >> 8, ln = (slot >> 2) & 63;
  const uint32_t row = 16 * (blk >> 2) + 4 * (ln >> 4) + (slot & 3), col = 16 * (blk & 3) + (ln & 15);
  const bool padding = l == kLayers - 1 && row >= kOutRows;
  float g = 0.f;
  if (!padding) {
#pragma unroll 8
    for (uint32_t c = slice; c < chunks; c += 4) g += partial[size_t(c) * kLayers * kW * kW + q];
  }
  part[slice][threadIdx.x & 63] = g;
  __syncthreads();
  if (slice != 0 || padding) return;
  const uint32_t i = l * kW * kW + row * kW + col;
  g = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
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

int cache_infer_blocks_per_cu() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, mlpk::infer_kernel, 256, 0) != hipSuccess || blocks < 1)
    blocks = 1;
  return blocks;
}

int launch_cache_infer(const CacheParams& P, const CacheQuery* q, uint32_t n, const uint32_t* n_dev, float* out,
                       int grid, hipStream_t s) {
  hipLaunchKernelGGL(mlpk::infer_kernel, dim3(grid), dim3(256), 0, s, P, q, n, n_dev, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cache_train_groups() {
  static int max_groups = 0;  // one resident workgroup per CU (LDS-limited): a persistent grid
  if (!max_groups) {
    int dev = 0, cus = 0, per_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mlpk::train_kernel, mlpk::kTrainThreads, 0) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    max_groups = std::min(kCacheMaxTrainGroups, std::max(1, cus * per_cu));
  }
  return max_groups;
}

int launch_cache_train(const CacheParams& P, const CacheQuery* q, const float* targets, uint32_t n, float grad_scale,
                       float* loss_tiles, float* partial, float lr, float beta1, float beta2, float eps, float bc1,
                       float bc2, float* grad_out, hipStream_t s) {
  const uint32_t tiles = (n + 31) / 32;
  const uint32_t groups = std::min<uint32_t>(uint32_t(cache_train_groups()),
                                             (tiles + mlpk::kTrainWaves - 1) / mlpk::kTrainWaves);
  hipLaunchKernelGGL(mlpk::train_kernel, dim3(groups), dim3(mlpk::kTrainThreads), 0, s, P, q, targets, n, grad_scale,
                     loss_tiles, partial);
  constexpr uint32_t kSlots = (mlpk::kLayers - 1) * mlpk::kW * mlpk::kW + 4 * 256;  // hidden layers, output blocks
  hipLaunchKernelGGL(mlpk::adam_kernel, dim3(kSlots / 64), dim3(256), 0, s, P, partial, groups, lr, beta1, beta2,
                     eps, bc1, bc2, grad_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cache_sync(const CacheParams& P, hipStream_t s) {
  hipLaunchKernelGGL(mlpk::sync_bf16_kernel, dim3((6 * 64 * 64 + 255) / 256), dim3(256), 0, s, P);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mrt
