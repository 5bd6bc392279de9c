// Culling grids of SDF union chains: the host build (at mrt_scene_push_sdf) and the same build as
// two GPU passes (after mrt_scene_set_sdf_param changes a program's constants). Both run the one
// set of __host__ __device__ routines below in binary64 with contraction off, so a grid rebuilt
// on the GPU holds the masks the host would have built, bit for bit.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <functional>
#include <thread>
#include <vector>

#include "sdf_grid.hpp"

namespace mrt {
namespace {

__host__ __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double dabs(double a) { return a < 0.0 ? -a : a; }

// Exact signed distance (binary64) of a chain primitive.
__host__ __device__ double chain_prim_distance(const SdfInsn& in, double x, double y, double z) {
  if (in.op & kSdfFuseTranslate) {
    x -= in.t[0];
    y -= in.t[1];
    z -= in.t[2];
  }
  switch (in.op & 0xffu) {
    case kSdfSphere:
      return sqrt(x * x + y * y + z * z) - in.p[0];
    case kSdfBox: {
      const double qx = dabs(x) - in.p[0], qy = dabs(y) - in.p[1], qz = dabs(z) - in.p[2];
      const double ox = dmax(qx, 0.0), oy = dmax(qy, 0.0), oz = dmax(qz, 0.0);
      return sqrt(ox * ox + oy * oy + oz * oz) + dmin(dmax(qx, dmax(qy, qz)), 0.0);
    }
    case kSdfTorus: {
      const double qx = sqrt(x * x + z * z) - in.p[0];
      return sqrt(qx * qx + y * y) - in.p[1];
    }
    default: {  // capsule
      const double yy = y - dmin(dmax(y, 0.0), double(in.p[0]));
      return sqrt(x * x + yy * yy + z * z) - in.p[1];
    }
  }
}

// Object-space box of a chain primitive.
__host__ __device__ void chain_prim_box(const SdfInsn& in, double lo[3], double hi[3]) {
  double h[3], c[3] = {0, 0, 0};
  switch (in.op & 0xffu) {
    case kSdfSphere:
      h[0] = h[1] = h[2] = in.p[0];
      break;
    case kSdfBox:
      for (int a = 0; a < 3; ++a) h[a] = dabs(in.p[a]);
      break;
    case kSdfTorus:
      h[0] = h[2] = double(in.p[0]) + in.p[1];
      h[1] = in.p[1];
      break;
    default:  // capsule: y in [-r, height + r]
      h[0] = h[2] = in.p[1];
      h[1] = 0.5 * in.p[0] + in.p[1];
      c[1] = 0.5 * in.p[0];
      break;
  }
  for (int a = 0; a < 3; ++a) {
    const double t = (in.op & kSdfFuseTranslate) ? in.t[a] : 0.0;
    lo[a] = t + c[a] - h[a];
    hi[a] = t + c[a] + h[a];
  }
}

// Primitives whose distance is exact (1-Lipschitz and at least the distance to their box).
bool chain_prim_ok(const SdfInsn& in) {
  switch (in.op & 0xffu) {
    case kSdfSphere:
      return in.p[0] > 0.f;
    case kSdfBox:
      return true;
    case kSdfTorus:
      return in.p[0] >= 0.f && in.p[1] > 0.f;
    case kSdfCapsule:
      return in.p[0] >= 0.f && in.p[1] > 0.f;
    default:
      return false;
  }
}

// Cell [lo, hi] of a grid with `dim` cells per axis, widened by a slack: the cell the device
// computes in binary32 may reach a little past the exact one.
__host__ __device__ inline void cell_box(const double mn[3], const double ext[3], int dim, const int ci[3],
                                         double clo[3], double chi[3]) {
  for (int a = 0; a < 3; ++a) {
    const double size = ext[a] / dim, slack = 1e-4 * ext[a];
    clo[a] = mn[a] + size * ci[a] - slack;
    chi[a] = mn[a] + size * (ci[a] + 1) + slack;
  }
}

// Mask of one cell [clo, chi] (slack included): bit j is set unless primitive j provably leaves
// the running value unchanged everywhere in the cell. Only the primitives in `candidates` are
// considered; the others are known to drop over the cell. A dropped primitive never lowers the
// bound on the running value (its distance anywhere in the cell exceeds the bound), so leaving the
// non-candidates out of the bound changes no mask.
__host__ __device__ uint64_t cell_mask(const SdfInsn* prog, uint32_t n, const double* blo, const double* bhi,
                                       const double clo[3], const double chi[3], uint64_t candidates) {
  double ctr[3], half2 = 0;
  for (int a = 0; a < 3; ++a) {
    ctr[a] = 0.5 * (clo[a] + chi[a]);
    half2 += 0.25 * (chi[a] - clo[a]) * (chi[a] - clo[a]);
  }
  const double half = sqrt(half2);
  uint64_t mask = 1;
  double upper = chain_prim_distance(prog[0], ctr[0], ctr[1], ctr[2]) + half;
  for (uint32_t j = 1; j < n; ++j) {
    if (!((candidates >> j) & 1u)) continue;
    {
      double gap2 = 0;
      bool overlap = true;
      for (int a = 0; a < 3; ++a) {
        const double gap = dmax(0.0, dmax(blo[3 * j + a] - chi[a], clo[a] - bhi[3 * j + a]));
        if (gap > 0) overlap = false;
        gap2 += gap * gap;
      }
      const bool smooth = ((prog[j].op >> 8) & 0xffu) == kSdfSmoothUnion;
      const double margin = (smooth ? double(prog[j].k) * (1.0 + 1e-3) : 0.0) + 1e-4 * (1.0 + dabs(upper));
      bool drop = !overlap && sqrt(gap2) > upper + margin;
      if ((prog[j].op & 0xffu) == kSdfSphere) {
        // Tighter and valid everywhere: |p - c| - r >= (distance from the cell to c) - r.
        double c2 = 0;
        for (int a = 0; a < 3; ++a) {
          const double c = (prog[j].op & kSdfFuseTranslate) ? double(prog[j].t[a]) : 0.0;
          const double gap = dmax(0.0, dmax(c - chi[a], clo[a] - c));
          c2 += gap * gap;
        }
        drop = sqrt(c2) - double(prog[j].p[0]) > upper + margin;
      }
      if (!drop) mask |= uint64_t(1) << j;
    }
    upper = dmin(upper, chain_prim_distance(prog[j], ctr[0], ctr[1], ctr[2]) + half);
  }
  return mask;
}

// GPU passes. Each block stages the program and its primitive boxes in LDS.
struct GridArgs {
  const SdfInsn* prog;
  uint32_t n;
  int dim;           // cells per axis of this pass
  int parent_ratio;  // fine pass: fine cells per coarse cell and axis (0: coarse pass)
  double mn[3], ext[3];
  const uint64_t* parent;  // coarse masks (fine pass)
  uint64_t* out;
};

__global__ void __launch_bounds__(256) grid_kernel(GridArgs g) {
  __shared__ SdfInsn prog[kSdfChainMax];
  __shared__ double blo[3 * kSdfChainMax], bhi[3 * kSdfChainMax];
  for (uint32_t j = threadIdx.x; j < g.n; j += blockDim.x) {
    prog[j] = g.prog[j];
    chain_prim_box(prog[j], &blo[3 * j], &bhi[3 * j]);
  }
  __syncthreads();
  const size_t cells = size_t(g.dim) * g.dim * g.dim;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= cells) return;
  const int cx = int(id % g.dim), cy = int((id / g.dim) % g.dim), cz = int(id / (size_t(g.dim) * g.dim));
  const int ci[3] = {cx, cy, cz};
  double clo[3], chi[3];
  cell_box(g.mn, g.ext, g.dim, ci, clo, chi);
  uint64_t cand = g.n >= 64 ? ~0ull : ((uint64_t(1) << g.n) - 1);
  if (g.parent_ratio) {
    const int C = g.dim / g.parent_ratio, R = g.parent_ratio;
    cand = g.parent[(size_t(cz / R) * C + cy / R) * C + cx / R];
  }
  g.out[id] = cell_mask(prog, g.n, blo, bhi, clo, chi, cand);
}

}  // namespace

int sdf_grid_dim(uint32_t n) { return n <= 8 ? 32 : (n <= 24 ? 64 : 128); }

bool sdf_grid_applies(const SdfInsn* prog, uint32_t n, const Box& box) {
  if (n < 2 || n > uint32_t(kSdfChainMax)) return false;
  if (!chain_prim_ok(prog[0]) || (prog[0].op & ~(0xffu | kSdfFuseTranslate)) != 0) return false;
  for (uint32_t j = 1; j < n; ++j) {
    const uint32_t fused = (prog[j].op >> 8) & 0xffu;
    if (!chain_prim_ok(prog[j]) || (prog[j].op >> 17) != 0) return false;
    if (fused != kSdfUnion && !(fused == kSdfSmoothUnion && prog[j].k > 0.f)) return false;
  }
  for (int a = 0; a < 3; ++a) {
    const double ext = double(box.mx[a]) - double(box.mn[a]);
    if (!(ext > 0.0) || !std::isfinite(ext)) return false;
  }
  return true;
}

void sdf_grid_frame(const Box& box, uint32_t n, SdfGrid& g) {
  const int D = sdf_grid_dim(n);
  g.dim = uint32_t(D);
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = box.mn[a];
    g.inv[a] = float(double(D) / (double(box.mx[a]) - double(box.mn[a])));
  }
}

// A coarse grid of 32 cells per axis comes first; each fine cell then tests only the bits of the
// coarse cell that contains it, since a primitive that provably drops over a cell drops over every
// part of it. Slabs of cells are built in parallel.
void build_sdf_grid(const SdfInsn* prog, uint32_t n, const Box& box, SdfGrid& g, std::vector<uint64_t>& masks) {
  g = SdfGrid{};
  if (!sdf_grid_applies(prog, n, box)) return;
  double mn[3], ext[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = double(box.mn[a]);
    ext[a] = double(box.mx[a]) - double(box.mn[a]);
  }
  std::vector<double> blo(3 * n), bhi(3 * n);
  for (uint32_t j = 0; j < n; ++j) chain_prim_box(prog[j], &blo[3 * j], &bhi[3 * j]);
  const int D = sdf_grid_dim(n), C = kSdfGridCoarse, R = D / C;
  const uint64_t all = n >= 64 ? ~0ull : ((uint64_t(1) << n) - 1);
  std::vector<uint64_t> coarse(size_t(C) * C * C), fine(size_t(D) * D * D);
  auto run_slabs = [](int slabs, const std::function<void(int)>& body) {
    const int threads = int(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&] {
        for (int z; (z = next.fetch_add(1)) < slabs;) body(z);
      });
    for (auto& th : pool) th.join();
  };
  run_slabs(C, [&](int cz) {
    for (int cy = 0; cy < C; ++cy)
      for (int cx = 0; cx < C; ++cx) {
        const int ci[3] = {cx, cy, cz};
        double clo[3], chi[3];
        cell_box(mn, ext, C, ci, clo, chi);
        coarse[(size_t(cz) * C + cy) * C + cx] = cell_mask(prog, n, blo.data(), bhi.data(), clo, chi, all);
      }
  });
  run_slabs(D, [&](int cz) {
    for (int cy = 0; cy < D; ++cy)
      for (int cx = 0; cx < D; ++cx) {
        const int ci[3] = {cx, cy, cz};
        double clo[3], chi[3];
        cell_box(mn, ext, D, ci, clo, chi);
        const uint64_t parent = coarse[(size_t(cz / R) * C + cy / R) * C + cx / R];
        fine[(size_t(cz) * D + cy) * D + cx] = R == 1 ? parent : cell_mask(prog, n, blo.data(), bhi.data(), clo, chi, parent);
      }
  });
  sdf_grid_frame(box, n, g);
  g.mask_offset = uint32_t(masks.size());
  masks.insert(masks.end(), fine.begin(), fine.end());
}

int launch_sdf_grid(const SdfInsn* d_prog, uint32_t n, const Box& box, uint64_t* d_coarse, uint64_t* d_masks,
                    hipStream_t stream) {
  const int D = sdf_grid_dim(n), C = kSdfGridCoarse;
  GridArgs g{};
  g.prog = d_prog;
  g.n = n;
  for (int a = 0; a < 3; ++a) {
    g.mn[a] = double(box.mn[a]);
    g.ext[a] = double(box.mx[a]) - double(box.mn[a]);
  }
  const bool two_pass = D != C;
  g.dim = C;
  g.parent_ratio = 0;
  g.out = two_pass ? d_coarse : d_masks;
  size_t cells = size_t(C) * C * C;
  grid_kernel<<<unsigned((cells + 255) / 256), 256, 0, stream>>>(g);
  if (hipGetLastError() != hipSuccess) return 1;
  if (!two_pass) return 0;
  g.dim = D;
  g.parent_ratio = D / C;
  g.parent = d_coarse;
  g.out = d_masks;
  cells = size_t(D) * D * D;
  grid_kernel<<<unsigned((cells + 255) / 256), 256, 0, stream>>>(g);
  return hipGetLastError() != hipSuccess ? 1 : 0;
}

}  // namespace mrt
