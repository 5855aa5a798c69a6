// step_kernels.h -- fused pull-stream + collide step, initialisation and outlet kernels.
//
// One lane owns one cell n = x + nx (y + ny z). Populations live in two complete sets of
// 19 direction-major planes (plane i of a set starts at i * cells), so every lane's loads
// of one plane and its stores to one plane are unit-stride across the wavefront.
// Semantics follow src/lbm/shaders/stream_collide.wgsl:95-136 (and the collision,
// wall and outlet helpers it assembles); arithmetic is f32 as in the reference.
#pragma once

#include "lattice.h"
#include "solver_internal.h"

namespace pw {

enum { kOpTrt = 0, kOpReg = 1 };

struct Stress {
  float xx, yy, zz, xy, xz, yz;
};

// common.wgsl:84-88: equilibrium with the density deviation kept apart from unit density.
__host__ __device__ __forceinline__ float geq(int i, float rho_delta, float ux, float uy, float uz,
                                     float uu) {
  const float cu = cx(i) * ux + cy(i) * uy + cz(i) * uz;
  return wf(i) * (rho_delta + (1.0f + rho_delta) * (3.0f * cu + 4.5f * cu * cu - 1.5f * uu));
}

// common.wgsl:100-113
__host__ __device__ __forceinline__ Stress neq_stress(const float* g, const float* eq) {
  Stress s{0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const float d = g[i] - eq[i];
    if (cx(i)) s.xx += d;
    if (cy(i)) s.yy += d;
    if (cz(i)) s.zz += d;
    if (cx(i) * cy(i)) s.xy += cx(i) * cy(i) * d;
    if (cx(i) * cz(i)) s.xz += cx(i) * cz(i) * d;
    if (cy(i) * cz(i)) s.yz += cy(i) * cz(i) * d;
  }
  return s;
}

// common.wgsl:115-122 (closed-form Smagorinsky tau; docs/numerics.md "LES").
template <bool LES>
__host__ __device__ __forceinline__ float relaxation_time(const StepParams& p, const Stress& s, float rho) {
  if (!LES) return p.tau0;
  const float q = s.xx * s.xx + s.yy * s.yy + s.zz * s.zz +
                  2.0f * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz);
  return 0.5f * (p.tau0 + sqrtf(p.tau0 * p.tau0 + p.cs_term * sqrtf(q) / rho));
}

__host__ __device__ __forceinline__ float projected(int i, const Stress& s, float third) {
  const float x = cx(i), y = cy(i), z = cz(i);
  return x * x * s.xx + y * y * s.yy + z * z * s.zz + 2.0f * (x * y * s.xy + x * z * s.xz + y * z * s.yz) -
         third;
}

// collide-trt.wgsl:1-15 and collide-reg.wgsl:3-18; writes post into g.
template <int OP, bool LES>
__host__ __device__ __forceinline__ void collide(const StepParams& p, float* g, const float* eq, float rho) {
  Stress s{0, 0, 0, 0, 0, 0};
  if (LES || OP == kOpReg) s = neq_stress(g, eq);
  const float tau = relaxation_time<LES>(p, s, rho);
  const float wp = 1.0f / tau;
  if (OP == kOpReg) {
    const float third = (s.xx + s.yy + s.zz) / 3.0f;
    const float k = (1.0f - wp) * 4.5f;
#pragma unroll
    for (int i = 0; i < kQ; ++i) g[i] = eq[i] + k * wf(i) * projected(i, s, third);
  } else {
    const float wm = 1.0f / (p.lambda / (tau - 0.5f) + 0.5f);
    g[0] = g[0] - wp * (g[0] - eq[0]);
#pragma unroll
    for (int i = 1; i < kQ; i += 2) {
      const int j = i + 1;
      const float even = 0.5f * wp * (g[i] + g[j] - eq[i] - eq[j]);
      const float odd = 0.5f * wm * (g[i] - g[j] - eq[i] + eq[j]);
      const float gi = g[i] - even - odd;
      const float gj = g[j] - even + odd;
      g[i] = gi;
      g[j] = gj;
    }
  }
}

// stream_collide.wgsl:8-12: Werner-Wengle friction velocity at y = 1/2.
__host__ __device__ __forceinline__ float friction_velocity(float nu, float speed) {
  const float viscous = sqrtf(nu * speed / 0.5f);
  if (0.5f * viscous / nu <= 11.81f) return viscous;
  return powf(speed * powf(nu / 0.5f, 1.0f / 7.0f) / 8.3f, 1.0f / (1.0f + 1.0f / 7.0f));
}

// stream_collide.wgsl:18-80: prescribe the smallest post-collision stress whose bounce-back
// flux along the slip is the wall-law shear; falls back to the bulk collision.
template <int OP>
__host__ __device__ __forceinline__ void collide_wall(const StepParams& p, float* g, const float* eq, float rho,
                                             float ux, float uy, float uz, uint32_t flag) {
  float nx = 0, ny = 0, nz = 0, exposed = 0;
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if (has_wall(flag, i)) {
      nx += wf(i) * cx(i);
      ny += wf(i) * cy(i);
      nz += wf(i) * cz(i);
      exposed += wf(i);
    }
  float sx = ux, sy = uy, sz = uz;
  const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
  const bool projected_slip = nl > 1e-3f;
  if (projected_slip) {
    const float ax = nx / nl, ay = ny / nl, az = nz / nl;
    const float d = ux * ax + uy * ay + uz * az;
    sx -= d * ax;
    sy -= d * ay;
    sz -= d * az;
  }
  const float speed = sqrtf(sx * sx + sy * sy + sz * sz);
  if (speed < 1e-7f) {
    collide<OP, true>(p, g, eq, rho);
    return;
  }
  const float tx = sx / speed, ty = sy / speed, tz = sz / speed;
  const float ut = friction_velocity(p.nu, speed);
  const float shear = 6.0f * exposed * rho * ut * ut;
  float flux = 0;
  Stress b{0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (!has_wall(flag, i)) continue;
    const int j = opp(i);
    const float x = cx(j), y = cy(j), z = cz(j);
    const float along = x * tx + y * ty + z * tz;
    // The weight part of the flux, sum 2 along_j w_j = -2 n.t, is exactly zero once the slip
    // has had its normal component removed; evaluated in f32 it would only add the rounding
    // residual of t's normal component, amplified by 1 / |slip| when the slip is small.
    flux += 2.0f * along * (projected_slip ? eq[j] : eq[j] + wf(j));
    const float s = 9.0f * wf(j) * along;
    b.xx += s * (x * x - 1.0f / 3.0f);
    b.yy += s * (y * y - 1.0f / 3.0f);
    b.zz += s * (z * z - 1.0f / 3.0f);
    b.xy += s * x * y;
    b.xz += s * x * z;
    b.yz += s * y * z;
  }
  const float norm = b.xx * b.xx + b.yy * b.yy + b.zz * b.zz + 2.0f * (b.xy * b.xy + b.xz * b.xz + b.yz * b.yz);
  if (norm < 0.05f) {
    collide<OP, true>(p, g, eq, rho);
    return;
  }
  const float scale = 4.5f * (shear - flux) / norm;
  const float third = (b.xx + b.yy + b.zz) / 3.0f;
  float post[kQ];
  bool admissible = true;
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    post[i] = eq[i] + scale * wf(i) * projected(i, b, third);
    admissible = admissible && (post[i] + wf(i) >= 0.0f);
  }
  if (!admissible) {
    collide<OP, true>(p, g, eq, rho);
    return;
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = post[i];
}

// Wrapped neighbour offsets of a cell: the flat index of (x + dx, y + dy, z + dz) is
// xo[dx+1] + yo[dy+1] + zo[dz+1] with every axis periodic (common.wgsl:64-68).
struct Neighbours {
  uint32_t xo[3], yo[3], zo[3];
  __host__ __device__ __forceinline__ Neighbours(uint32_t x, uint32_t y, uint32_t z, const StepParams& p) {
    const uint32_t row = p.nx, plane = p.nx * p.ny;
    xo[0] = x == 0 ? p.nx - 1 : x - 1;
    xo[1] = x;
    xo[2] = x + 1 == p.nx ? 0 : x + 1;
    yo[0] = (y == 0 ? p.ny - 1 : y - 1) * row;
    yo[1] = y * row;
    yo[2] = (y + 1 == p.ny ? 0 : y + 1) * row;
    zo[0] = (z == 0 ? p.nz - 1 : z - 1) * plane;
    zo[1] = z * plane;
    zo[2] = (z + 1 == p.nz ? 0 : z + 1) * plane;
  }
  __host__ __device__ __forceinline__ uint32_t at(int dx, int dy, int dz) const {
    return xo[dx + 1] + yo[dy + 1] + zo[dz + 1];
  }
};

// walls.wgsl:6-21: halfway bounce-back or linear Bouzidi for incoming direction i, evaluated
// in R (f32 inside the step, f64 inside the drag reduction).
template <class S, bool BOUZIDI, class R = float>
__host__ __device__ __forceinline__ R wall_value(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                        const StepParams& p, const Neighbours& nb, uint32_t n,
                                        uint32_t flag, int i) {
  const size_t N = p.cells;
  const int j = opp(i);
  const R reflected = S::load(in + (size_t)j * N + n);
  if (!BOUZIDI) return reflected;
  const uint32_t base = geo[N + n];
  if (base == kNoQ) return reflected;
  const uint32_t rank = __builtin_popcount((flag >> 3) & ((1u << i) - 1u));
  const R q = pw::as_float(geo[2 * N + base + rank]);
  if (q == R(0.5)) return reflected;
  if (q < R(0.5)) {
    const uint32_t second = nb.at(cx(i), cy(i), cz(i));
    if ((geo[second] & kTypeMask) != kFluid) return reflected;
    return R(2) * q * reflected + (R(1) - R(2) * q) * R(S::load(in + (size_t)j * N + second));
  }
  return (reflected + (R(2) * q - R(1)) * R(S::load(in + (size_t)i * N + n))) / (R(2) * q);
}

template <class S, bool NT = false>
__host__ __device__ __forceinline__ void store_state(typename S::T* __restrict__ out, size_t N, uint32_t n, const float* g) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) st<NT>(out + (size_t)i * N + n, S::encode(g[i], (uint32_t)i, n));
}

// common.wgsl:145-150 with rho - 1 = 0.
template <class S>
__host__ __device__ __forceinline__ void store_equilibrium(typename S::T* __restrict__ out, size_t N, uint32_t n,
                                                  float ux, float uy, float uz) {
  const float uu = ux * ux + uy * uy + uz * uz;
  float g[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = geq(i, 0.0f, ux, uy, uz, uu);
  store_state<S>(out, N, n, g);
}

// init.wgsl:1-8: walls at rest, every other cell at the inflow equilibrium.
template <class S>
__global__ __launch_bounds__(256) void init_kernel(typename S::T* __restrict__ out,
                                                   const uint32_t* __restrict__ geo,
                                                   float4* __restrict__ macro, StepParams p) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;
  const uint32_t kind = geo[n] & kTypeMask;
  const float ux = is_wall(kind) ? 0.0f : p.u_in;
  store_equilibrium<S>(out, p.cells, n, ux, 0.0f, 0.0f);
  macro[n] = make_float4(1.0f, ux, 0.0f, 0.0f);
}

// Moments, equilibrium and collision of one gathered fluid cell (stream_collide.wgsl:114-133);
// g becomes the post-collision state; returns (rho, u).
template <int OP, bool LES>
__host__ __device__ __forceinline__ float4 relax(const StepParams& p, float* g, uint32_t flag) {
  float rho_delta = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    rho_delta += g[i];
    mx += cx(i) * g[i];
    my += cy(i) * g[i];
    mz += cz(i) * g[i];
  }
  const float rho = 1.0f + rho_delta;
  const float ux = mx / rho, uy = my / rho, uz = mz / rho;
  const float uu = ux * ux + uy * uy + uz * uz;
  float eq[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) eq[i] = geq(i, rho_delta, ux, uy, uz, uu);
  if (LES && (flag >> 3) != 0)
    collide_wall<OP>(p, g, eq, rho, ux, uy, uz, flag);
  else
    collide<OP, LES>(p, g, eq, rho);
  return make_float4(rho, ux, uy, uz);
}

// stream_collide.wgsl:95-136, one cell per lane.
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, bool NT = false>
__host__ __device__ __forceinline__ void stream_collide_cell(const typename S::T* __restrict__ in,
                                                             typename S::T* __restrict__ out,
                                                             const uint32_t* __restrict__ geo,
                                                             float4* __restrict__ macro,
                                                             const StepParams& p, uint32_t n) {
  const size_t N = p.cells;
  const uint32_t flag = geo[n];
  const uint32_t kind = flag & kTypeMask;
  if (is_wall(kind)) return;
  const uint32_t x = n % p.nx;
  const uint32_t r = n / p.nx;
  const uint32_t y = r % p.ny;
  const uint32_t z = r / p.ny;
  if (kind == kEq) {
    float ux = p.u_in, uy = 0.0f, uz = 0.0f;
    if (x == p.nx - 1 && x > 0) {
      // stream_collide.wgsl:83-93: pressure outlet takes the upstream velocity.
      float mass = 1.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const float v = S::load(in + (size_t)i * N + (n - 1));
        mass += v;
        mx += cx(i) * v;
        my += cy(i) * v;
        mz += cz(i) * v;
      }
      ux = mx / mass;
      uy = my / mass;
      uz = mz / mass;
    }
    store_equilibrium<S>(out, N, n, ux, uy, uz);
    if (MACRO) macro[n] = make_float4(1.0f, ux, uy, uz);
    return;
  }
  const Neighbours nb(x, y, z, p);
  float g[kQ];
  g[0] = load_pop<S, NT>(in + n);
  if ((flag >> 3) == 0) {
#pragma unroll
    for (int i = 1; i < kQ; ++i) g[i] = load_pop<S, NT>(in + (size_t)i * N + nb.at(-cx(i), -cy(i), -cz(i)));
  } else {
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      g[i] = has_wall(flag, i) ? wall_value<S, BOUZIDI>(in, geo, p, nb, n, flag, i)
                               : S::load(in + (size_t)i * N + nb.at(-cx(i), -cy(i), -cz(i)));
  }
  const float4 m = relax<OP, LES>(p, g, flag);
  store_state<S, NT>(out, N, n, g);
  if (MACRO) macro[n] = m;
}

template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, bool NT>
__global__ __launch_bounds__(256) void stream_collide_kernel(const typename S::T* __restrict__ in,
                                                             typename S::T* __restrict__ out,
                                                             const uint32_t* __restrict__ geo,
                                                             float4* __restrict__ macro, StepParams p) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;
  stream_collide_cell<S, OP, LES, BOUZIDI, MACRO, NT>(in, out, geo, macro, p, n);
}

// Two x-adjacent cells per lane. A pair whose cells are both plain fluid (no wall links) and
// whose pulls do not wrap in x gathers and stores each direction with one wide access;
// every other pair takes the per-cell path. Needs an even nx.
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, bool NT = false>
__host__ __device__ __forceinline__ void stream_collide_pair(const typename S::T* __restrict__ in,
                                                             typename S::T* __restrict__ out,
                                                             const uint32_t* __restrict__ geo,
                                                             float4* __restrict__ macro,
                                                             const StepParams& p, uint32_t n) {
  const uint2 f = *reinterpret_cast<const uint2*>(geo + n);
  const uint32_t x = n % p.nx;
  if ((f.x | f.y) != 0 || x == 0 || x + 2 == p.nx) {
    stream_collide_cell<S, OP, LES, BOUZIDI, MACRO, NT>(in, out, geo, macro, p, n);
    stream_collide_cell<S, OP, LES, BOUZIDI, MACRO, NT>(in, out, geo, macro, p, n + 1);
    return;
  }
  const size_t N = p.cells;
  const uint32_t r = n / p.nx;
  const Neighbours nb(x, r % p.ny, r / p.ny, p);
  float a[kQ], b[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) load_pair<S, NT>(in + (size_t)i * N + nb.at(-cx(i), -cy(i), -cz(i)), a[i], b[i]);
  const float4 ma = relax<OP, LES>(p, a, 0u);
  const float4 mb = relax<OP, LES>(p, b, 0u);
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    store_pair<S, NT>(out + (size_t)i * N + n, S::encode(a[i], (uint32_t)i, n), S::encode(b[i], (uint32_t)i, n + 1));
  if (MACRO) {
    macro[n] = ma;
    macro[n + 1] = mb;
  }
}

template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, bool NT>
__global__ __launch_bounds__(256) void stream_collide_pair_kernel(const typename S::T* __restrict__ in,
                                                                  typename S::T* __restrict__ out,
                                                                  const uint32_t* __restrict__ geo,
                                                                  float4* __restrict__ macro, StepParams p) {
  const uint32_t n = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (n >= p.cells) return;
  stream_collide_pair<S, OP, LES, BOUZIDI, MACRO, NT>(in, out, geo, macro, p, n);
}

}  // namespace pw
