// force_kernels.h -- FP64 momentum-exchange drag over the boundary-cell list.
//
// Semantics follow src/lbm/shaders/force.wgsl:1-40 and reduce.wgsl:1-18: for every listed
// fluid cell and every wall link whose pull source is SOLID, add c_j (g*_j + wall value of
// incoming i), j = opp(i), over BOTH ping-pong states, times 1/2 (the mean of the two latest
// steps). Accumulation is FP64 with a fixed-order two-stage reduction (wavefront butterfly,
// then one block over the partials), so a sample is bit-identical from run to run.
#pragma once

#include "step_kernels.h"

namespace pw {

constexpr int kForceBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

template <class S, bool BOUZIDI>
__device__ __forceinline__ void cell_exchange(const typename S::T* __restrict__ state,
                                              const uint32_t* __restrict__ geo, const StepParams& p,
                                              const Neighbours& nb, uint32_t n, uint32_t self, uint32_t flag,
                                              double& fx, double& fy, double& fz) {
  const size_t P = p.run;
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (!has_wall(flag, i)) continue;
    const uint32_t source = nb.index(-cx(i), -cy(i), -cz(i));
    if ((geo[source] & kTypeMask) != kSolid) continue;
    const int j = opp(i);
    const double outgoing = S::load(state + (size_t)j * P + self);
    const double incoming = wall_value<S, BOUZIDI, double>(state, geo, p, nb, n, self, flag, i);
    const double e = outgoing + incoming;
    fx += cx(j) * e;
    fy += cy(j) * e;
    fz += cz(j) * e;
  }
}

template <class S, bool BOUZIDI>
__global__ __launch_bounds__(kForceBlock) void force_partials_kernel(
    const typename S::T* __restrict__ state_a, const typename S::T* __restrict__ state_b,
    const uint32_t* __restrict__ geo, const uint32_t* __restrict__ boundary, uint32_t n_boundary,
    double* __restrict__ partials, StepParams p) {
  const uint32_t item = blockIdx.x * kForceBlock + threadIdx.x;
  double fx = 0, fy = 0, fz = 0;
  if (item < n_boundary) {
    const uint32_t n = boundary[item];
    const uint32_t flag = geo[n];
    uint32_t x, y, z;
    cell_coords(n, p, x, y, z);
    const Neighbours nb(x, y, z, p);
    const uint32_t self = pop_x(p, x) + pop_row(p, y + p.ny * z);
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    cell_exchange<S, BOUZIDI>(state_a, geo, p, nb, n, self, flag, ax, ay, az);
    cell_exchange<S, BOUZIDI>(state_b, geo, p, nb, n, self, flag, bx, by, bz);
    fx = 0.5 * (ax + bx);
    fy = 0.5 * (ay + by);
    fz = 0.5 * (az + bz);
  }
  __shared__ double lds[3][kForceBlock / 64];
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    lds[0][wave] = fx;
    lds[1][wave] = fy;
    lds[2][wave] = fz;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kForceBlock / 64; ++w) s += lds[threadIdx.x][w];
    partials[3 * (size_t)blockIdx.x + threadIdx.x] = s;
  }
}

// One block strides the partials in a fixed order, then reduces across its lanes.
__device__ __forceinline__ void force_reduce_block(const double* __restrict__ partials, uint32_t n_partials,
                                                   double* __restrict__ result) {
  double s[3] = {0, 0, 0};
  for (uint32_t k = threadIdx.x; k < n_partials; k += kForceBlock)
    for (int c = 0; c < 3; ++c) s[c] += partials[3 * (size_t)k + c];
  __shared__ double lds[3][kForceBlock / 64];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int c = 0; c < 3; ++c) {
    const double v = wave_sum(s[c]);
    if (lane == 0) lds[c][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0;
    for (int w = 0; w < kForceBlock / 64; ++w) t += lds[threadIdx.x][w];
    result[threadIdx.x] = t;
  }
}

}  // namespace pw
