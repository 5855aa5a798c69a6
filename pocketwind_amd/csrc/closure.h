// closure.h -- the learned wall closure's kernels over the wall-cell list: the closure kernel
// (features, and the network's stress when one is attached) and the per-wall-cell forces.
// The features are defined in closure_features.h.
#pragma once

#include "closure_features.h"
#include "force_kernels.h"

namespace pw {
namespace closure {

// Wall cells per wavefront: one 16-row matrix-core tile. The wall cells are few (thousands),
// so one tile per wavefront spreads them over many more wavefronts than four tiles would, and
// each wavefront runs one network pass instead of four in sequence.
constexpr int kCellsPerWave = 16;

// One wavefront per 16 wall cells: features (optionally written raw to `records`), then, when
// a network is attached (P non-null), its stress for every cell into tau (w = 1 valid, 0 use
// the wall law). Lanes 0-15 own the cells; all 64 lanes run the network's tile.
template <class S, bool BOUZIDI>
__global__ __launch_bounds__(64) void closure_kernel(const typename S::T* __restrict__ in,
                                                     const uint32_t* __restrict__ geo, StepParams p,
                                                     const WallGeo* __restrict__ walls, uint32_t n_cells,
                                                     const mlp::Packed* __restrict__ P, Norm norm,
                                                     float4* __restrict__ tau, float* __restrict__ records) {
  __shared__ __attribute__((aligned(16))) uint16_t xs[kCellsPerWave * mlp::kInPad];
  __shared__ __attribute__((aligned(16))) uint16_t h1[16 * mlp::kHidden];
  __shared__ __attribute__((aligned(16))) uint16_t h2[16 * mlp::kHidden];
  __shared__ float ys[kCellsPerWave * 4];
  const int l = threadIdx.x;
  const uint32_t k = blockIdx.x * kCellsPerWave + l;
  const bool valid = l < kCellsPerWave && k < n_cells;
  float f[kFeatures], law[3];
  bool in_range = valid;
  if (valid) {
    wall_features<S, BOUZIDI>(in, geo, p, norm.u_ref, load_wall_geo(walls, k), f, law);
    if (records) {
      for (int j = 0; j < kFeatures; ++j) records[(size_t)k * kRecord + j] = f[j];
      for (int j = 0; j < 3; ++j) records[(size_t)k * kRecord + kFeatures + j] = law[j];
    }
  }
  if (!P) return;
  if (l < kCellsPerWave)
    for (int j = 0; j < mlp::kInPad; ++j) {
      float v = 0.0f;
      if (valid && j < kFeatures) {
        v = (f[j] - norm.mean[j]) * norm.scale[j];
        in_range = in_range && fabsf(v) <= norm.limit;
      }
      xs[l * mlp::kInPad + j] = mlp::bf16_bits(v);
    }
  mlp::wave_sync();
  const int r = l & 15, gq = l >> 4;
  const mlp::f32x4_t y = mlp::forward16(P, xs, h1, h2);
  if (r < 3)
    for (int j = 0; j < 4; ++j) ys[(4 * gq + j) * 4 + r] = y[j];
  mlp::wave_sync();
  if (!valid) return;
  const float tx = ys[l * 4] * norm.out_scale, ty = ys[l * 4 + 1] * norm.out_scale, tz = ys[l * 4 + 2] * norm.out_scale;
  const bool ok = in_range && isfinite(tx) && isfinite(ty) && isfinite(tz);
  tau[k] = make_float4(tx, ty, tz, ok ? 1.0f : 0.0f);
}

// Momentum-exchange force on the body from each wall cell, the mean of both states (the
// per-cell terms of force_partials_kernel, not reduced).
template <class S, bool BOUZIDI>
__global__ __launch_bounds__(256) void wall_force_kernel(const typename S::T* __restrict__ state_a,
                                                         const typename S::T* __restrict__ state_b,
                                                         const uint32_t* __restrict__ geo, StepParams p,
                                                         const uint32_t* __restrict__ cells, uint32_t n_cells,
                                                         double* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_cells) return;
  const uint32_t n = cells[k];
  const uint32_t flag = geo[n];
  uint32_t x, y, z;
  cell_coords(n, p, x, y, z);
  const Neighbours nb(x, y, z, p);
  const uint32_t self = pop_x(p, x) + pop_row(p, y + p.ny * z);
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
  cell_exchange<S, BOUZIDI>(state_a, geo, p, nb, n, self, flag, ax, ay, az);
  cell_exchange<S, BOUZIDI>(state_b, geo, p, nb, n, self, flag, bx, by, bz);
  out[3 * (size_t)k] = 0.5 * (ax + bx);
  out[3 * (size_t)k + 1] = 0.5 * (ay + by);
  out[3 * (size_t)k + 2] = 0.5 * (az + bz);
}

}  // namespace closure
}  // namespace pw
