// closure.h -- the learned wall closure's kernels over the wall-cell list: the closure kernel
// (features, and the network's stress when one is attached) and the per-wall-cell forces.
// The features are defined in closure_features.h.
#pragma once

#include "closure_features.h"
#include "force_kernels.h"

namespace pw {
namespace closure {

// The closure kernel: one wavefront per 16-cell tile of the wall-cell list.
template <class S, bool BOUZIDI>
__global__ __launch_bounds__(64) void closure_kernel(const typename S::T* __restrict__ in,
                                                     const uint32_t* __restrict__ geo, StepParams p,
                                                     const WallGeo* __restrict__ walls, uint32_t n_cells,
                                                     const mlp::Packed* __restrict__ P, Norm norm,
                                                     float4* __restrict__ tau, float* __restrict__ records) {
  __shared__ __attribute__((aligned(16))) char lds[kFusedWaveBytes];
  closure_tile<S, BOUZIDI>(in, geo, p, walls, n_cells, P, norm, tau, records, blockIdx.x, lds);
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
