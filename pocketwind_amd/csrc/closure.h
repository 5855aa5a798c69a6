// closure.h -- features of the learned wall closure and its kernels over the wall-cell list.
//
// A wall cell is a fluid cell with at least one wall link. Its 12 features, all from the state
// a step is about to read:
//   0-2   velocity of the cell / u_ref (its populations gathered exactly as the step gathers
//         them, wall links included)
//   3-5   velocity / u_ref of the fluid cell one step off the wall along the rounded wall
//         normal (the cell itself when that neighbour is not fluid)
//   6     mean wall-link fraction q of the cell's links from the packed geometry (1/2 when the
//         geometry carries none)
//   7     log10 of the molecular viscosity
//   8-10  unit wall normal, n = sum over wall links of w_i c_i normalised
//   11    6 sum over wall links of w_i (1 for one flat face)
// The network maps the normalised features to the wall shear stress vector (lattice units);
// the wall collision uses its component along the slip in place of the wall law's
// rho u_tau^2 (stream_collide.wgsl:18-80). Cells whose normalised features leave
// [-limit, limit], or whose output is not finite, keep the wall law.
#pragma once

#include "force_kernels.h"
#include "mlp.h"

namespace pw {
namespace closure {

constexpr int kFeatures = 12;
constexpr int kRecord = 15;  // features, then the wall law's stress vector

struct Norm {
  float mean[kFeatures];
  float scale[kFeatures];
  float out_scale;  // stress = network output * out_scale
  float limit;      // largest |normalised feature| the network is trusted with
  float u_ref;      // velocity unit of features 0-5
};

template <class S, bool BOUZIDI>
__device__ __forceinline__ void gather_cell(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                            const StepParams& p, uint32_t n, float* g) {
  const size_t P = p.run;
  const uint32_t flag = geo[n];
  uint32_t x, y, z;
  cell_coords(n, p, x, y, z);
  const Neighbours nb(x, y, z, p);
  const PopNeighbours<> pnb(x, y, z, p);
  const uint32_t self = pnb.at(0, 0, 0);
  g[0] = S::load(in + self);
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    g[i] = has_wall(flag, i) ? wall_value<S, BOUZIDI>(in, geo, p, nb, n, self, flag, i)
                             : S::load(in + (size_t)i * P + pnb.at(-cx(i), -cy(i), -cz(i)));
}

__device__ __forceinline__ void velocity_of(const float* g, float& rho, float& ux, float& uy, float& uz) {
  float rd = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    rd += g[i];
    mx += cx(i) * g[i];
    my += cy(i) * g[i];
    mz += cz(i) * g[i];
  }
  rho = 1.0f + rd;
  ux = mx / rho;
  uy = my / rho;
  uz = mz / rho;
}

template <class S, bool BOUZIDI>
__device__ __forceinline__ void wall_features(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                              const StepParams& p, float u_ref, uint32_t n, float* f, float* law) {
  const size_t N = p.cells;
  const uint32_t flag = geo[n];
  float g[kQ], rho, ux, uy, uz;
  gather_cell<S, BOUZIDI>(in, geo, p, n, g);
  velocity_of(g, rho, ux, uy, uz);
  float nx = 0.0f, ny = 0.0f, nz = 0.0f, exposed = 0.0f, qsum = 0.0f;
  int links = 0;
  const uint32_t base = geo[N + n];
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (!has_wall(flag, i)) continue;
    nx += wf(i) * cx(i);
    ny += wf(i) * cy(i);
    nz += wf(i) * cz(i);
    exposed += wf(i);
    float q = 0.5f;  // the geometry's link fraction when it carries one, whatever the wall rule
    if (base != kNoQ) q = as_float(geo[2 * N + base + __builtin_popcount((flag >> 3) & ((1u << i) - 1u))]);
    qsum += q;
    ++links;
  }
  const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
  const float ax = nl > 1e-3f ? nx / nl : 0.0f, ay = nl > 1e-3f ? ny / nl : 0.0f, az = nl > 1e-3f ? nz / nl : 0.0f;
  // Second sample: the fluid cell one step along the rounded normal.
  const int dx = (int)rintf(ax), dy = (int)rintf(ay), dz = (int)rintf(az);
  float vx = ux, vy = uy, vz = uz;
  if (dx | dy | dz) {
    uint32_t x, y, z;
    cell_coords(n, p, x, y, z);
    const uint32_t m = Neighbours(x, y, z, p).index(dx, dy, dz);
    if ((geo[m] & kTypeMask) == kFluid) {
      float h[kQ], r2;
      gather_cell<S, BOUZIDI>(in, geo, p, m, h);
      velocity_of(h, r2, vx, vy, vz);
    }
  }
  const float inv = 1.0f / u_ref;
  f[0] = ux * inv;
  f[1] = uy * inv;
  f[2] = uz * inv;
  f[3] = vx * inv;
  f[4] = vy * inv;
  f[5] = vz * inv;
  f[6] = links ? qsum / links : 0.5f;
  f[7] = log10f(p.nu);
  f[8] = ax;
  f[9] = ay;
  f[10] = az;
  f[11] = 6.0f * exposed;
  // The wall law's stress vector rho u_tau^2 t (stream_collide.wgsl:18-80), for comparison.
  float sx = ux, sy = uy, sz = uz;
  if (nl > 1e-3f) {
    const float d = ux * ax + uy * ay + uz * az;
    sx -= d * ax;
    sy -= d * ay;
    sz -= d * az;
  }
  const float speed = sqrtf(sx * sx + sy * sy + sz * sz);
  law[0] = law[1] = law[2] = 0.0f;
  if (speed >= 1e-7f) {
    const float ut = friction_velocity(p.nu, speed);
    const float st = rho * ut * ut / speed;
    law[0] = st * sx;
    law[1] = st * sy;
    law[2] = st * sz;
  }
}

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
                                                     const uint32_t* __restrict__ cells, uint32_t n_cells,
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
    wall_features<S, BOUZIDI>(in, geo, p, norm.u_ref, cells[k], f, law);
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
