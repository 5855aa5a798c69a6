// step_kernels.h -- fused pull-stream + collide step, initialisation and outlet kernels.
//
// One lane owns one cell n = x + nx (y + ny z). Populations live in two complete sets of
// 19 direction-major planes (plane i of a set starts at i * cells), so every lane's loads
// of one plane and its stores to one plane are unit-stride across the wavefront.
// Semantics follow src/lbm/shaders/stream_collide.wgsl:95-136 (and the collision,
// wall and outlet helpers it assembles); arithmetic is f32 as in the reference.
//
// The collision is evaluated in pair form. Directions i = 2p + 1 and j = i + 1 are
// opposite (c_j = -c_i), so
//   * the moments need only the pair sums g_i + g_j (density, second moment) and the pair
//     differences g_i - g_j (momentum);
//   * the equilibria share an even part a_p and flip an odd part b_p: eq_i = a_p + b_p,
//     eq_j = a_p - b_p;
//   * the non-equilibrium stress sum_i c c (g_i - eq_i) is the second moment of g minus its
//     closed form at equilibrium, (rho - 1)/3 I + rho u u (exact for D3Q19, whose weights
//     are isotropic to fourth order);
//   * the regularised post-collision state is even in c, so one correction per pair.
// This is the reference's arithmetic regrouped: the same quantities in f32, summed in
// another order.
#pragma once

#include <type_traits>

#include "lattice.h"
#include "solver_internal.h"

namespace pw {

enum { kOpTrt = 0, kOpReg = 1 };

// Reciprocal and square root to about one ulp (v_rcp_f32 / v_sqrt_f32) on the device; the
// host pass (CPU checks of this kernel source) uses the correctly rounded operations.
__host__ __device__ __forceinline__ float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ float fast_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return __builtin_sqrtf(x);
#endif
}

struct Stress {
  float xx, yy, zz, xy, xz, yz;
};

// A sum whose terms are known at compile time to be present or absent (products with the
// D3Q19 velocity components, which are -1, 0 or 1). The first present term starts it, so no
// product with a zero component and no addition of an initial zero is ever evaluated: IEEE
// arithmetic cannot fold either away (0 * x is not 0 for every x, nor is 0 + x always x).
struct Sum {
  float v = 0.0f;
  bool any = false;
  __host__ __device__ __forceinline__ void add(float t) {
    v = any ? v + t : t;
    any = true;
  }
  __host__ __device__ __forceinline__ void sub(float t) {
    v = any ? v - t : -t;
    any = true;
  }
  // c t for a velocity component c in {-1, 0, 1}.
  __host__ __device__ __forceinline__ void add(int c, float t) {
    if (c > 0) add(t);
    else if (c < 0) sub(t);
  }
};

// c_i . (x, y, z) from the nonzero components of c_i only.
__host__ __device__ __forceinline__ float cdot(int i, float x, float y, float z) {
  Sum r;
  r.add(cx(i), x);
  r.add(cy(i), y);
  r.add(cz(i), z);
  return r.v;
}

// common.wgsl:84-88: equilibrium with the density deviation kept apart from unit density.
__host__ __device__ __forceinline__ float geq(int i, float rho_delta, float ux, float uy, float uz,
                                              float uu) {
  const float cu = cdot(i, ux, uy, uz);
  return wf(i) * (rho_delta + (1.0f + rho_delta) * (3.0f * cu + 4.5f * cu * cu - 1.5f * uu));
}

constexpr int kPairs = 9;

// Moments and pair-form equilibrium of one gathered cell (stream_collide.wgsl:115-126).
struct CellState {
  float rd, rho, inv_rho, ux, uy, uz;
  float eq0;
  float a[kPairs], b[kPairs];      // eq_{2p+1} = a + b, eq_{2p+2} = a - b
  float sum[kPairs], dif[kPairs];  // g_{2p+1} + g_{2p+2}, g_{2p+1} - g_{2p+2}
};

__host__ __device__ __forceinline__ void cell_moments(const float* g, CellState& c) {
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    c.sum[p] = g[2 * p + 1] + g[2 * p + 2];
    c.dif[p] = g[2 * p + 1] - g[2 * p + 2];
  }
  float rd = g[0];
  Sum mx, my, mz;
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    const int i = 2 * p + 1;
    rd += c.sum[p];
    mx.add(cx(i), c.dif[p]);
    my.add(cy(i), c.dif[p]);
    mz.add(cz(i), c.dif[p]);
  }
  c.rd = rd;
  c.rho = 1.0f + rd;
  c.inv_rho = fast_rcp(c.rho);
  c.ux = mx.v * c.inv_rho;
  c.uy = my.v * c.inv_rho;
  c.uz = mz.v * c.inv_rho;
}

// A cell whose state is prescribed (an EQ cell, or a wall at rest) keeps its gathered sums only
// as placeholders: unit density and the velocity (ux, uy, uz) replace its moments.
__host__ __device__ __forceinline__ void prescribe_moments(CellState& c, float ux, float uy, float uz) {
  c.rd = 0.0f;
  c.rho = 1.0f;
  c.inv_rho = 1.0f;
  c.ux = ux;
  c.uy = uy;
  c.uz = uz;
}

// The output scale 2^SC of a collision: its post-collision values leave multiplied by the
// storage format's scale (2^15 for f16) so the stores convert them directly. A power of two
// commutes exactly with every rounding here, so the stored bits are those of scaling after.
template <int SC>
constexpr float kOut = (float)(1u << SC);

// The pair-form equilibrium of the cell's moments, scaled by 2^SC.
template <int SC = 0>
__host__ __device__ __forceinline__ void cell_equilibrium(CellState& c) {
  const float rd = c.rd;
  const float h = 1.5f * (c.ux * c.ux + c.uy * c.uy + c.uz * c.uz);
  c.eq0 = (wf(0) * kOut<SC>) * (rd - c.rho * h);
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    const int i = 2 * p + 1;
    const float cu = cdot(i, c.ux, c.uy, c.uz);
    const float rw = (wf(i) * kOut<SC>) * c.rho;
    c.a[p] = rw * (4.5f * cu * cu - h) + (wf(i) * kOut<SC>) * rd;
    c.b[p] = (3.0f * rw) * cu;
  }
}

__host__ __device__ __forceinline__ void cell_state(const float* g, CellState& c) {
  cell_moments(g, c);
  cell_equilibrium(c);
}

__host__ __device__ __forceinline__ float eq_of(const CellState& c, int i) {
  if (i == 0) return c.eq0;
  const int p = (i - 1) / 2;
  return (i & 1) ? c.a[p] + c.b[p] : c.a[p] - c.b[p];
}

// common.wgsl:100-113: non-equilibrium stress, here as the second moment of g minus its
// equilibrium value (rho - 1)/3 I + rho u u.
__host__ __device__ __forceinline__ Stress neq_stress(const CellState& c) {
  Sum xx, yy, zz, xy, xz, yz;
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    const int i = 2 * p + 1;
    const float v = c.sum[p];
    xx.add(cx(i) * cx(i), v);
    yy.add(cy(i) * cy(i), v);
    zz.add(cz(i) * cz(i), v);
    xy.add(cx(i) * cy(i), v);
    xz.add(cx(i) * cz(i), v);
    yz.add(cy(i) * cz(i), v);
  }
  Stress s{xx.v, yy.v, zz.v, xy.v, xz.v, yz.v};
  const float third = c.rd * (1.0f / 3.0f);
  const float rx = c.rho * c.ux, ry = c.rho * c.uy, rz = c.rho * c.uz;
  s.xx -= rx * c.ux + third;
  s.yy -= ry * c.uy + third;
  s.zz -= rz * c.uz + third;
  s.xy -= rx * c.uy;
  s.xz -= rx * c.uz;
  s.yz -= ry * c.uz;
  return s;
}

// common.wgsl:115-122 (closed-form Smagorinsky tau; docs/numerics.md "LES").
__host__ __device__ __forceinline__ float les_tau(const StepParams& p, const Stress& s, float inv_rho) {
  const float q = s.xx * s.xx + s.yy * s.yy + s.zz * s.zz + 2.0f * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz);
  return 0.5f * (p.tau0 + fast_sqrt(p.tau0 * p.tau0 + p.cs_term * fast_sqrt(q) * inv_rho));
}

// c_i c_i : s - trace(s) / 3 (third), from the components c_i has.
__host__ __device__ __forceinline__ float projected(int i, const Stress& s, float third) {
  Sum diag, off;
  diag.add(cx(i) * cx(i), s.xx);
  diag.add(cy(i) * cy(i), s.yy);
  diag.add(cz(i) * cz(i), s.zz);
  off.add(cx(i) * cy(i), s.xy);
  off.add(cx(i) * cz(i), s.xz);
  off.add(cy(i) * cz(i), s.yz);
  return (off.any ? diag.v + 2.0f * off.v : diag.v) - third;
}

// collide-trt.wgsl:1-15 and collide-reg.wgsl:3-18; writes the post-collision state into g.
// With `prescribed` (REG only) the non-equilibrium stress is zero, so the cell leaves with its
// equilibrium: the state of an EQ cell, or the rest state of a wall.
// c holds the equilibrium scaled by 2^SC (cell_equilibrium<SC>), and g receives the state at the
// same scale; the TRT relaxation works on the gathered values and takes SC = 0.
// The REG collision regrouped so that it needs only the cell's moments (cell_moments), not its
// per-direction equilibrium: with M = rho u u + (1 - omega) Pi, direction i leaves with
//   w_i (rho - 1) + 4.5 w_i (c_i c_i - I/3) : M  +-  3 w_i c_i . (rho u)
// (a pair shares the even part and flips the odd one): collide-reg.wgsl:3-18 with
// rho (4.5 (c.u)^2 - 1.5 u.u) = 4.5 (c c - I/3) : rho u u. It takes fewer vector instructions
// than forming the nine pairs' equilibrium first: the f16 step on the 256x128x128 body ran
// 0.7-0.9 us faster (same-box comparisons). Its f32 roundings differ from the equilibrium
// form's by about an ulp, far below the 16-bit stores' 11 (f16) or 8 (bf16) bits, so the 16-bit
// storages with the library's rounding take it (S::kRegroupedReg). f32 storage keeps the
// equilibrium form, whose density the FP64-oracle parity holds to three f32 ulps on a transient
// that amplifies either form's roundings (this form lands 3.8 there), and so does the f16
// storage with the reference's rounding, whose stored bits are pinned against the reference's
// encoding of the f32 step's values.
template <int SC>
__host__ __device__ __forceinline__ void collide_reg_regrouped(float* g, const CellState& c, const Stress& s,
                                                               float wp) {
  const float om = 1.0f - wp;
  const float rx = c.rho * c.ux, ry = c.rho * c.uy, rz = c.rho * c.uz;
  const float mxx = rx * c.ux + om * s.xx, myy = ry * c.uy + om * s.yy, mzz = rz * c.uz + om * s.zz;
  const float mxy = rx * c.uy + om * s.xy, mxz = rx * c.uz + om * s.xz, myz = ry * c.uz + om * s.yz;
  const float t3 = (mxx + myy + mzz) * (1.0f / 3.0f);
  g[0] = (wf(0) * kOut<SC>) * c.rd - (4.5f * wf(0) * kOut<SC>) * t3;
  const float e1 = (wf(1) * kOut<SC>) * c.rd, e2 = (wf(7) * kOut<SC>) * c.rd;
  constexpr float w1 = 4.5f * wf(1) * kOut<SC>, w2 = 4.5f * wf(7) * kOut<SC>;
  constexpr float o1 = 3.0f * wf(1) * kOut<SC>, o2 = 3.0f * wf(7) * kOut<SC>;
  const float pxy = mxx + myy - t3, pxz = mxx + mzz - t3, pyz = myy + mzz - t3;
#pragma unroll
  for (int q = 0; q < kPairs; ++q) {
    const int i = 2 * q + 1;
    float even, odd;
    if (wf(i) == wf(1)) {
      even = e1 + w1 * ((cx(i) ? mxx : cy(i) ? myy : mzz) - t3);
      odd = o1 * cdot(i, rx, ry, rz);
    } else {
      // (c_i c_i - I/3) : M for (+-1, +-1) in one plane: the plane's diagonal sum less the trace
      // third, + 2 c_a c_b m_ab.
      const float sum = cz(i) == 0 ? pxy : cy(i) == 0 ? pxz : pyz;
      const float m = cz(i) == 0 ? mxy : cy(i) == 0 ? mxz : myz;
      const int sign = cz(i) == 0 ? cx(i) * cy(i) : cy(i) == 0 ? cx(i) * cz(i) : cy(i) * cz(i);
      even = e2 + w2 * (sign > 0 ? sum + 2.0f * m : sum - 2.0f * m);
      odd = o2 * cdot(i, rx, ry, rz);
    }
    g[i] = even + odd;
    g[i + 1] = even - odd;
  }
}

// Whether a collision reads the per-direction equilibrium (cell_equilibrium): TRT, and REG in
// its equilibrium form.
template <int OP, bool REGROUPED>
constexpr bool kNeedsEquilibrium = !(OP == kOpReg && REGROUPED);

template <int OP, bool LES, int SC = 0, bool REGROUPED = false>
__host__ __device__ __forceinline__ void collide(const StepParams& p, float* g, const CellState& c,
                                                 bool prescribed = false) {
  static_assert(OP == kOpReg || SC == 0, "the TRT relaxation works unscaled");
  Stress s{0, 0, 0, 0, 0, 0};
  if (LES || OP == kOpReg) s = neq_stress(c);
  if (OP == kOpReg && prescribed) s = Stress{0, 0, 0, 0, 0, 0};
  const float tau = LES ? les_tau(p, s, c.inv_rho) : p.tau0;
  const float wp = LES ? fast_rcp(tau) : p.inv_tau0;
  if (OP == kOpReg && REGROUPED) {
    collide_reg_regrouped<SC>(g, c, s, wp);
  } else if (OP == kOpReg) {
    const float third = (s.xx + s.yy + s.zz) * (1.0f / 3.0f);
    const float k = (1.0f - wp) * 4.5f;
    g[0] = c.eq0 - k * (wf(0) * kOut<SC>) * third;
#pragma unroll
    for (int q = 0; q < kPairs; ++q) {
      const int i = 2 * q + 1;
      const float t = c.a[q] + k * (wf(i) * kOut<SC>) * projected(i, s, third);
      g[i] = t + c.b[q];
      g[i + 1] = t - c.b[q];
    }
  } else {
    const float wm = LES ? fast_rcp(p.lambda * fast_rcp(tau - 0.5f) + 0.5f) : p.wm0;
    g[0] = g[0] - wp * (g[0] - c.eq0);
#pragma unroll
    for (int q = 0; q < kPairs; ++q) {
      const int i = 2 * q + 1;
      const float even = wp * (0.5f * c.sum[q] - c.a[q]);
      const float odd = wm * (0.5f * c.dif[q] - c.b[q]);
      const float gi = g[i] - even - odd;
      const float gj = g[i + 1] - even + odd;
      g[i] = gi;
      g[i + 1] = gj;
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
// `learned`: the learned closure's wall shear stress vector for this cell (w > 0 when valid),
// used in place of the wall law; w = 0 keeps the law.
template <int OP, int SC = 0, bool REGROUPED = false>
__host__ __device__ __forceinline__ void collide_wall(const StepParams& p, float* g, const CellState& c,
                                                     uint32_t flag, const float4& learned) {
  float nx = 0, ny = 0, nz = 0, exposed = 0;
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if (has_wall(flag, i)) {
      nx += wf(i) * cx(i);
      ny += wf(i) * cy(i);
      nz += wf(i) * cz(i);
      exposed += wf(i);
    }
  float sx = c.ux, sy = c.uy, sz = c.uz;
  const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
  const bool projected_slip = nl > 1e-3f;
  if (projected_slip) {
    const float ax = nx / nl, ay = ny / nl, az = nz / nl;
    const float d = c.ux * ax + c.uy * ay + c.uz * az;
    sx -= d * ax;
    sy -= d * ay;
    sz -= d * az;
  }
  const float speed = sqrtf(sx * sx + sy * sy + sz * sz);
  if (speed < p.slip_floor) {  // stream_collide.wgsl:40 (1e-7 with f32 storage)
    collide<OP, true, SC, REGROUPED>(p, g, c);
    return;
  }
  const float tx = sx / speed, ty = sy / speed, tz = sz / speed;
  // The wall shear stress along the slip: the learned closure's where it gave a valid value
  // for this cell, else the wall law's rho u_tau^2.
  float stress = -1.0f;
  if (learned.w > 0.0f) stress = fmaxf(learned.x * tx + learned.y * ty + learned.z * tz, 0.0f);
  if (stress < 0.0f) {
    const float ut = friction_velocity(p.nu, speed);
    stress = c.rho * ut * ut;
  }
  const float shear = 6.0f * exposed * stress;
  float flux = 0;
  Stress b{0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (!has_wall(flag, i)) continue;
    const int j = opp(i);
    const float x = cx(j), y = cy(j), z = cz(j);
    const float along = cdot(j, tx, ty, tz);
    // The weight part of the flux, sum 2 along_j w_j = -2 n.t, is exactly zero once the slip
    // has had its normal component removed; evaluated in f32 it would only add the rounding
    // residual of t's normal component, amplified by 1 / |slip| when the slip is small.
    const float eqj = eq_of(c, j) * (1.0f / kOut<SC>);
    flux += 2.0f * along * (projected_slip ? eqj : eqj + wf(j));
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
    collide<OP, true, SC, REGROUPED>(p, g, c);
    return;
  }
  const float scale = 4.5f * (shear - flux) / norm;
  const float third = (b.xx + b.yy + b.zz) / 3.0f;
  float post[kQ];
  bool admissible = true;
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    post[i] = eq_of(c, i) + scale * (wf(i) * kOut<SC>) * projected(i, b, third);
    admissible = admissible && (post[i] + wf(i) * kOut<SC> >= 0.0f);
  }
  if (!admissible) {
    collide<OP, true, SC, REGROUPED>(p, g, c);
    return;
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = post[i];
}

// Cell coordinates of a flat index.
__host__ __device__ __forceinline__ void cell_coords(uint32_t n, const StepParams& p, uint32_t& x, uint32_t& y,
                                                     uint32_t& z) {
  const uint32_t r = fast_div(n, p.div_nx);
  x = n - r * p.nx;
  z = fast_div(r, p.div_ny);
  y = r - z * p.ny;
}

// Neighbour coordinate lo (c - 1) and hi (c + 1) of c on an axis of n cells: periodic
// (common.wgsl:64-68), or held at c where `clamp_lo` / `clamp_hi` holds (free-slip sides, the
// zero-gradient outlet).
__host__ __device__ __forceinline__ uint32_t below(uint32_t c, uint32_t n, bool clamp) {
  return c == 0 ? (clamp ? c : n - 1) : c - 1;
}
__host__ __device__ __forceinline__ uint32_t above(uint32_t c, uint32_t n, bool clamp) {
  return c + 1 == n ? (clamp ? c : 0) : c + 1;
}

// Flat indices of a cell's neighbours (cell arrays such as the geometry): the index of
// (x + dx, y + dy, z + dz) is xo[dx+1] + yo[dy+1] + zo[dz+1], every axis periodic except
// where p.bc holds the source at the boundary.
struct Neighbours {
  uint32_t xo[3], yo[3], zo[3];
  __host__ __device__ __forceinline__ Neighbours(uint32_t x, uint32_t y, uint32_t z, const StepParams& p)
      : Neighbours(x, y, z, p, p.bc) {}
  // bc: the boundary options in force (p.bc, or 0 where the caller knows them to be off).
  __host__ __device__ __forceinline__ Neighbours(uint32_t x, uint32_t y, uint32_t z, const StepParams& p,
                                                 uint32_t bc) {
    const uint32_t row = p.nx, plane = p.nx * p.ny;
    const bool slip = bc & kBcFreeSlip;
    xo[0] = below(x, p.nx, false);
    xo[1] = x;
    xo[2] = above(x, p.nx, bc & kBcZeroGradient);
    yo[0] = below(y, p.ny, slip) * row;
    yo[1] = y * row;
    yo[2] = above(y, p.ny, slip) * row;
    zo[0] = below(z, p.nz, slip) * plane;
    zo[1] = z * plane;
    zo[2] = above(z, p.nz, slip) * plane;
  }
  __host__ __device__ __forceinline__ uint32_t index(int dx, int dy, int dz) const {
    return xo[dx + 1] + yo[dy + 1] + zo[dz + 1];
  }
};

// Population element offsets (direction 0) of the layout described with StepParams::run:
// x + x_tile floor(x / 64) within a row, 19 nx per row.
__host__ __device__ __forceinline__ uint32_t pop_x(const StepParams& p, uint32_t x) { return x + p.x_tile * (x >> 6); }
__host__ __device__ __forceinline__ uint32_t pop_row(const StepParams& p, uint32_t r) { return r * (kQ * p.nx); }

// The same for population elements, scaled by 2^SHIFT (element bytes for byte addressing).
// The layout is additive in x and in the row, and a set is below 4 GiB (checked at create), so
// a byte offset fits a u32.
template <int SHIFT = 0>
struct PopNeighbours {
  uint32_t xo[3], yo[3], zo[3];
  __host__ __device__ __forceinline__ PopNeighbours(uint32_t x, uint32_t y, uint32_t z, const StepParams& p)
      : PopNeighbours(x, y, z, p, p.bc) {}
  __host__ __device__ __forceinline__ PopNeighbours(uint32_t x, uint32_t y, uint32_t z, const StepParams& p,
                                                    uint32_t bc) {
    const uint32_t row = kQ * p.nx, plane = row * p.ny;
    const bool slip = bc & kBcFreeSlip;
    xo[0] = pop_x(p, below(x, p.nx, false)) << SHIFT;
    xo[1] = pop_x(p, x) << SHIFT;
    xo[2] = pop_x(p, above(x, p.nx, bc & kBcZeroGradient)) << SHIFT;
    yo[0] = (below(y, p.ny, slip) * row) << SHIFT;
    yo[1] = (y * row) << SHIFT;
    yo[2] = (above(y, p.ny, slip) * row) << SHIFT;
    zo[0] = (below(z, p.nz, slip) * plane) << SHIFT;
    zo[1] = (z * plane) << SHIFT;
    zo[2] = (above(z, p.nz, slip) * plane) << SHIFT;
  }
  __host__ __device__ __forceinline__ uint32_t at(int dx, int dy, int dz) const {
    return xo[dx + 1] + yo[dy + 1] + zo[dz + 1];
  }
};

template <class T>
__host__ __device__ __forceinline__ const T* byte_offset(const T* base, uint32_t bytes) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + bytes);
}
template <class T>
__host__ __device__ __forceinline__ T* byte_offset(T* base, uint32_t bytes) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + bytes);
}

// walls.wgsl:6-21: halfway bounce-back or linear Bouzidi for incoming direction i, evaluated
// in R (f32 inside the step, f64 inside the drag reduction). n is the cell's flat index, self
// its population element offset and nb its flat neighbours.
template <class S, bool BOUZIDI, class R = float>
__host__ __device__ __forceinline__ R wall_value(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                                 const StepParams& p, const Neighbours& nb, uint32_t n, uint32_t self,
                                                 uint32_t flag, int i) {
  const size_t N = p.cells, P = p.run;
  const int j = opp(i);
  const R reflected = S::load(in + (size_t)j * P + self);
  if (!BOUZIDI) return reflected;
  const uint32_t base = geo[N + n];
  if (base == kNoQ) return reflected;
  const uint32_t rank = __builtin_popcount((flag >> 3) & ((1u << i) - 1u));
  const R q = pw::as_float(geo[2 * N + base + rank]);
  if (q == R(0.5)) return reflected;
  if (q < R(0.5)) {
    const uint32_t second = nb.index(cx(i), cy(i), cz(i));
    if ((geo[second] & kTypeMask) != kFluid) return reflected;
    const uint32_t r = fast_div(second, p.div_nx);
    const uint32_t second_elem = pop_x(p, second - r * p.nx) + pop_row(p, r);
    return R(2) * q * reflected + (R(1) - R(2) * q) * R(S::load(in + (size_t)j * P + second_elem));
  }
  return (reflected + (R(2) * q - R(1)) * R(S::load(in + (size_t)i * P + self))) / (R(2) * q);
}

// Population element offset of a flat cell index.
__host__ __device__ __forceinline__ uint32_t pop_cell(const StepParams& p, uint32_t n) {
  const uint32_t r = fast_div(n, p.div_nx);
  return pop_x(p, n - r * p.nx) + pop_row(p, r);
}

// The output scale exponent a step uses: the format's for the regularised collision (whose
// equilibrium and reconstruction carry the scale in their weights), 0 for TRT.
template <class S, int OP>
constexpr int kOutLog2 = OP == kOpReg ? S::kScaleLog2 : 0;

// Stores of one cell's 19 values; P is the direction stride (StepParams::run), `self` the
// cell's byte offset within direction 0 and h its random word for this step.
// SCALED: g holds the values already multiplied by the format's scale 2^S::kScaleLog2.
template <class S, bool SCALED>
__host__ __device__ __forceinline__ typename S::T encode_out(float v, uint32_t r) {
  if constexpr (SCALED) return S::encode_scaled(v, r);
  else return S::encode(v, r);
}

template <class S, bool NT = false, bool SCALED = false>
__host__ __device__ __forceinline__ void store_state(typename S::T* __restrict__ out, size_t P, uint32_t self,
                                                     const float* g, uint32_t h) {
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    st<NT>(byte_offset(out + (size_t)i * P, self), encode_out<S, SCALED>(g[i], S::dir_word(h, i)));
}

// common.wgsl:145-150 with rho - 1 = 0, in pair form: eq_i = w (4.5 cu^2 - 1.5 u^2) + 3 w cu.
// Evaluated without contraction into FMAs, so that the result does not depend on where it is
// inlined (the EQ cells of the row and cell paths, and of batched tunnels, agree bit for bit).
__host__ __device__ __forceinline__ void equilibrium_values(float ux, float uy, float uz, float* g) {
#pragma clang fp contract(off)
  const float h15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  g[0] = wf(0) * (0.0f - h15);  // +0 at rest, as the stored rest state of walls
#pragma unroll
  for (int q = 0; q < kPairs; ++q) {
    const int i = 2 * q + 1;
    const float cu = cdot(i, ux, uy, uz);
    const float a = wf(i) * (4.5f * cu * cu - h15);
    const float b = (3.0f * wf(i)) * cu;
    g[i] = a + b;
    g[i + 1] = a - b;
  }
}

template <class S, bool NT = false>
__host__ __device__ __forceinline__ void store_equilibrium(typename S::T* __restrict__ out, const StepParams& p,
                                                           uint32_t n, float ux, float uy, float uz, uint32_t h) {
  float g[kQ];
  equilibrium_values(ux, uy, uz, g);
  store_state<S, NT>(out, p.run, pop_cell(p, n) << S::kShift, g, h);
}

// Between a set in the population layout and the reference's planes (19 planes of N cells,
// direction-major; docs/numerics.md "Population storage"). T holds the raw storage bits.
template <class T>
__global__ __launch_bounds__(256) void to_planes_kernel(const T* __restrict__ set, T* __restrict__ planes,
                                                        StepParams p) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;
  const uint32_t e = pop_cell(p, n);
#pragma unroll
  for (int i = 0; i < kQ; ++i) planes[(size_t)i * p.cells + n] = set[(size_t)i * p.run + e];
}

template <class T>
__global__ __launch_bounds__(256) void from_planes_kernel(const T* __restrict__ planes, T* __restrict__ set,
                                                          StepParams p) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;
  const uint32_t e = pop_cell(p, n);
#pragma unroll
  for (int i = 0; i < kQ; ++i) set[(size_t)i * p.run + e] = planes[(size_t)i * p.cells + n];
}

// Copy a set out to (or in from) host planes through a device staging buffer, on `stream`.
inline hipError_t set_to_host_planes(const void* set, void* host, const StepParams& p, uint32_t elem, hipStream_t stream) {
  const size_t bytes = (size_t)kQ * p.cells * elem;
  void* stage = nullptr;
  hipError_t e = hipMallocAsync(&stage, bytes, stream);
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)((p.cells + 255) / 256);
  if (elem == 4)
    hipLaunchKernelGGL(to_planes_kernel<uint32_t>, dim3(blocks), dim3(256), 0, stream, (const uint32_t*)set,
                       (uint32_t*)stage, p);
  else
    hipLaunchKernelGGL(to_planes_kernel<uint16_t>, dim3(blocks), dim3(256), 0, stream, (const uint16_t*)set,
                       (uint16_t*)stage, p);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, stage, bytes, hipMemcpyDeviceToHost, stream);
  (void)hipFreeAsync(stage, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return e;
}

inline hipError_t set_from_host_planes(void* set, const void* host, const StepParams& p, uint32_t elem,
                                       hipStream_t stream) {
  const size_t bytes = (size_t)kQ * p.cells * elem;
  void* stage = nullptr;
  hipError_t e = hipMallocAsync(&stage, bytes, stream);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(stage, host, bytes, hipMemcpyHostToDevice, stream);
  const unsigned blocks = (unsigned)((p.cells + 255) / 256);
  if (e == hipSuccess) {
    if (elem == 4)
      hipLaunchKernelGGL(from_planes_kernel<uint32_t>, dim3(blocks), dim3(256), 0, stream, (const uint32_t*)stage,
                         (uint32_t*)set, p);
    else
      hipLaunchKernelGGL(from_planes_kernel<uint16_t>, dim3(blocks), dim3(256), 0, stream, (const uint16_t*)stage,
                         (uint16_t*)set, p);
    e = hipGetLastError();
  }
  (void)hipFreeAsync(stage, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return e;
}

// Walls keep the rest state (all deviations zero; stream_collide.wgsl:96 skips them). They
// store it with their wavefront anyway: a wavefront whose lanes all store writes whole cache
// lines, while a skipped lane leaves a partly written line that HBM must read, merge and
// write back.
__host__ __device__ __forceinline__ void rest_state(float* g) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = 0.0f;
}

// init.wgsl:1-8: walls at rest, every other cell at the inflow equilibrium.
template <class S>
__global__ __launch_bounds__(256) void init_kernel(typename S::T* __restrict__ out,
                                                   const uint32_t* __restrict__ geo,
                                                   float4* __restrict__ macro, StepParams p) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;
  const uint32_t kind = geo[n] & kTypeMask;
  const bool wall = is_wall(kind);
  float g[kQ];
  // Walls store exact zeros (the rest state) rather than an equilibrium at zero velocity,
  // whose direction-0 term could come out as -0 depending on how the compiler folds it.
  if (wall) rest_state(g);
  else equilibrium_values(p.u_in, 0.0f, 0.0f, g);
  store_state<S, false>(out, p.run, pop_cell(p, n) << S::kShift, g, S::cell_word(n, kInitSeed));
  macro[n] = make_float4(1.0f, wall ? 0.0f : p.u_in, 0.0f, 0.0f);
}

// The learned closure's output for cell n from the closure kernel's buffer (none: the law).
__host__ __device__ __forceinline__ float4 buffered_stress(const StepParams& p, uint32_t flag, uint32_t n) {
  if (p.closure_tau && (flag & kTypeMask) == kFluid && (flag >> 3) != 0) return p.closure_tau[p.closure_slot[n]];
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Moments, equilibrium and collision of one gathered fluid cell (stream_collide.wgsl:114-133);
// g becomes the post-collision state; returns (rho, u).
template <int OP, bool LES>
__host__ __device__ __forceinline__ float4 relax(const StepParams& p, float* g, uint32_t flag,
                                                 const float4& learned) {
  CellState c;
  cell_state(g, c);
  if (LES && (flag >> 3) != 0)
    collide_wall<OP>(p, g, c, flag, learned);
  else
    collide<OP, LES>(p, g, c);
  return make_float4(c.rho, c.ux, c.uy, c.uz);
}

// relax for any lane of a REG step (TRT steps call it for fluid lanes only): fluid lanes relax their gathered populations; a prescribed
// lane (EQ or wall) takes unit density and the velocity (ux, uy, uz) as its moments and zero
// non-equilibrium stress, so it leaves with that equilibrium (a wall with the rest state, exactly
// zero) through the same arithmetic as every fluid lane instead of a separate branch.
template <int OP, bool LES, int SC = 0, bool REGROUPED = false>
__host__ __device__ __forceinline__ float4 relax_prescribed(const StepParams& p, float* g, uint32_t flag,
                                                            const float4& learned, bool prescribed, float ux,
                                                            float uy, float uz) {
  CellState c;
  cell_moments(g, c);
  if (prescribed) prescribe_moments(c, ux, uy, uz);
  // The wall collision reads the per-direction equilibrium; the regrouped REG collision does not.
  if (LES && (flag >> 3) != 0) {
    cell_equilibrium<SC>(c);
    collide_wall<OP, SC, REGROUPED>(p, g, c, flag, learned);
  } else {
    if (kNeedsEquilibrium<OP, REGROUPED>) cell_equilibrium<SC>(c);
    collide<OP, LES, SC, REGROUPED>(p, g, c, prescribed);
  }
  return make_float4(c.rho, c.ux, c.uy, c.uz);
}

// The outlet plane x = nx - 1 takes its velocity from the stored state of the cell upstream,
// x = nx - 2 (stream_collide.wgsl:102-107): (1 + sum g, sum c g) of its 19 stored values, in
// direction order.
struct RowMoments {
  float mass, mx, my, mz;
};

__host__ __device__ __forceinline__ void add_moment(RowMoments& m, int i, float v) {
  m.mass += v;
  if (cx(i)) m.mx += cx(i) * v;
  if (cy(i)) m.my += cy(i) * v;
  if (cz(i)) m.mz += cz(i) * v;
}

template <class S>
__host__ __device__ __forceinline__ RowMoments stored_moments(const typename S::T* __restrict__ in,
                                                              const StepParams& p, uint32_t n) {
  RowMoments m{1.0f, 0.0f, 0.0f, 0.0f};
  const uint32_t e = pop_cell(p, n);
#pragma unroll
  for (int i = 0; i < kQ; ++i) add_moment(m, i, S::load(in + (size_t)i * p.run + e));
  return m;
}

// The step that stores a row's x = nx - 2 state also stores those moments, one float4 per row
// (y + ny z) in the set's allocation after its 19 N elements; the outlet reads that one value
// instead of gathering 19 at the end of its wavefront, whose latency nothing would hide.
__host__ __device__ __forceinline__ size_t edge_offset_bytes(size_t cells, size_t elem_bytes) {
  return (kQ * cells * elem_bytes + 255) & ~(size_t)255;
}

template <class T>
__host__ __device__ __forceinline__ float4* edge_rows(T* set, const StepParams& p) {
  return reinterpret_cast<float4*>(reinterpret_cast<char*>(set) + edge_offset_bytes(p.cells, sizeof(T)));
}

template <class T>
__host__ __device__ __forceinline__ const float4* edge_rows(const T* set, const StepParams& p) {
  return reinterpret_cast<const float4*>(reinterpret_cast<const char*>(set) + edge_offset_bytes(p.cells, sizeof(T)));
}

__host__ __device__ __forceinline__ RowMoments row_moments(const float4& e) { return RowMoments{e.x, e.y, e.z, e.w}; }

// Moments of the values a store of g with the random word h writes.
template <class S, bool SCALED = false>
__host__ __device__ __forceinline__ RowMoments encoded_moments(const float* g, uint32_t h) {
  RowMoments m{1.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < kQ; ++i) add_moment(m, i, S::decode(encode_out<S, SCALED>(g[i], S::dir_word(h, i))));
  return m;
}

// stream_collide.wgsl:83-107: an EQ cell takes the equilibrium at unit density and the
// inlet speed or, on the outlet plane x = nx - 1 (nx > 1), the velocity of the upstream
// moments `up`. Writes the post-step values into g and returns (1, u).
__host__ __device__ __forceinline__ void eq_velocity(const StepParams& p, uint32_t x, const RowMoments& up, float& ux,
                                                    float& uy, float& uz) {
  ux = p.u_in;
  uy = 0.0f;
  uz = 0.0f;
  if (x == p.nx - 1 && x > 0) {
    const float inv = fast_rcp(up.mass);
    ux = up.mx * inv;
    uy = up.my * inv;
    uz = up.mz * inv;
  }
}

__host__ __device__ __forceinline__ float4 eq_from(const StepParams& p, uint32_t x, const RowMoments& up, float* g) {
  float ux, uy, uz;
  eq_velocity(p, x, up, ux, uy, uz);
  equilibrium_values(ux, uy, uz, g);
  return make_float4(1.0f, ux, uy, uz);
}

// The same with the upstream state gathered from the input set.
template <class S>
__host__ __device__ __forceinline__ float4 eq_state(const typename S::T* __restrict__ in, const StepParams& p,
                                                    uint32_t n, uint32_t x, float* g) {
  RowMoments up{1.0f, 0.0f, 0.0f, 0.0f};
  if (x == p.nx - 1 && x > 0) up = stored_moments<S>(in, p, n - 1);
  return eq_from(p, x, up, g);
}

// The upstream moments of every row from a set's stored state (after init or an upload).
template <class S>
__global__ __launch_bounds__(256) void edge_moments_kernel(typename S::T* __restrict__ set, StepParams p) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= p.ny * p.nz || p.nx < 2) return;
  const RowMoments m = stored_moments<S>(set, p, r * p.nx + (p.nx - 2));
  edge_rows(set, p)[r] = make_float4(m.mass, m.mx, m.my, m.mz);
}


// nan_guard.wgsl:4-17 fused into every macro-writing step: a density that is not finite or
// outside [0.5, 1.5] (NaN fails both comparisons) sets the guard word, the first word after
// the cells' macros (macro buffers hold cells + 1 entries). Every writer stores the same 1, so
// a plain store suffices; the word lives beside the macros so that the step needs no further
// kernel argument.
__host__ __device__ __forceinline__ void guard_density(float4* macro, const StepParams& p, float rho) {
  if (!(rho >= 0.5f && rho <= 1.5f)) *reinterpret_cast<uint32_t*>(macro + p.cells) = 1u;
}

// A cell in a free-slip side row (y = 0, ny - 1 or z = 0, nz - 1).
__host__ __device__ __forceinline__ bool on_side(uint32_t y, uint32_t z, const StepParams& p) {
  return y == 0 || y + 1 == p.ny || z == 0 || z + 1 == p.nz;
}

// Plane that direction i pulls from in a free-slip side row: the specular image of i in each
// side the pull would cross (ylo: y = 0, and so on). The source offsets already hold that axis
// at the cell, so the pull reads the mirrored direction of (x - c_x, y, z - c_z) at y = 0, etc.
template <int I>
__host__ __device__ __forceinline__ int slip_plane(bool ylo, bool yhi, bool zlo, bool zhi) {
  const bool my = cy(I) > 0 ? ylo : (cy(I) < 0 ? yhi : false);
  const bool mz = cz(I) > 0 ? zlo : (cz(I) < 0 ? zhi : false);
  constexpr int y_image = mirror_dir(I, true, false), z_image = mirror_dir(I, false, true),
                yz_image = mirror_dir(I, true, true);
  return my ? (mz ? yz_image : y_image) : (mz ? z_image : I);
}

template <int I, int END, class F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < END) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, END>(f);
  }
}

// Free-slip side row, cell path: after the plain gather, the pulls that cross a side are read
// again from their mirrored planes (the source offsets already hold the crossed axis at the
// cell). Side rows hold no wall links (walls keep two cells from a free-slip side).
template <class S, bool NT>
__host__ __device__ __forceinline__ void slip_cell(const typename S::T* __restrict__ in, const StepParams& p,
                                                   const PopNeighbours<S::kShift>& pnb, uint32_t y, uint32_t z,
                                                   float* g) {
  const bool ylo = y == 0, yhi = y + 1 == p.ny, zlo = z == 0, zhi = z + 1 == p.nz;
  static_for<1, kQ>([&](auto c) {
    constexpr int i = decltype(c)::value;
    if constexpr (cy(i) != 0 || cz(i) != 0) {
      const int j = slip_plane<i>(ylo, yhi, zlo, zhi);
      if (j != i) g[i] = load_pop<S, NT>(byte_offset(in + (size_t)j * p.run, pnb.at(-cx(i), -cy(i), -cz(i))));
    }
  });
}

// Cache policy of the population streams: cached, non-temporal loads and stores, or cached
// loads with non-temporal stores. Flags and wall-path loads are always cached.
enum { kCached = 0, kNtAll = 1, kNtStores = 2 };

// stream_collide.wgsl:95-136, one cell per lane; `seed` selects the step's rounding words.
// EDGE: the outlet reads the upstream moments stored with the input set and the row's
// x = nx - 2 cell stores its own with the output set (the solver's sets carry them); without
// it the outlet gathers the upstream state (batched tunnels).
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, int POL = kCached, bool EDGE = false, bool BC = false>
__host__ __device__ __forceinline__ void stream_collide_cell(const typename S::T* __restrict__ in,
                                                             typename S::T* __restrict__ out,
                                                             const uint32_t* __restrict__ geo,
                                                             float4* __restrict__ macro,
                                                             const StepParams& p, uint32_t n, uint32_t seed) {
  const size_t P = p.run;
  const uint32_t flag = geo[n];
  const uint32_t kind = flag & kTypeMask;
  uint32_t x, y, z;
  cell_coords(n, p, x, y, z);
  const uint32_t bc = BC ? p.bc : 0u;
  const PopNeighbours<S::kShift> pnb(x, y, z, p, bc);
  const uint32_t self = pnb.at(0, 0, 0);
  float g[kQ];
  float4 m;
  if (kind == kFluid) {
    g[0] = load_pop<S, POL == kNtAll>(byte_offset(in, self));
    if ((flag >> 3) == 0) {
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        g[i] = load_pop<S, POL == kNtAll>(byte_offset(in + (size_t)i * P, pnb.at(-cx(i), -cy(i), -cz(i))));
    } else {
      const Neighbours nb(x, y, z, p, bc);
#pragma unroll
      for (int i = 1; i < kQ; ++i)
        g[i] = has_wall(flag, i) ? wall_value<S, BOUZIDI>(in, geo, p, nb, n, self >> S::kShift, flag, i)
                                 : S::load(byte_offset(in + (size_t)i * P, pnb.at(-cx(i), -cy(i), -cz(i))));
    }
    if (BC && (p.bc & kBcFreeSlip) && on_side(y, z, p)) slip_cell<S, POL == kNtAll>(in, p, pnb, y, z, g);
    if constexpr (OP != kOpReg) m = relax<OP, LES>(p, g, flag, LES ? buffered_stress(p, flag, n) : float4{});
  } else if (OP == kOpReg) {
    rest_state(g);  // placeholders: an EQ or wall cell's gathered values do not enter its state
  }
  if constexpr (OP == kOpReg) {
    // EQ and wall cells take their equilibrium through the fluid cells' collision, from the one
    // call site the row path has too, so that both compile the collision alike.
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (kind == kEq) {
      RowMoments up{1.0f, 0.0f, 0.0f, 0.0f};
      if (x == p.nx - 1 && x > 0) {
        if constexpr (EDGE) up = row_moments(edge_rows(in, p)[y + p.ny * z]);
        else up = stored_moments<S>(in, p, n - 1);
      }
      eq_velocity(p, x, up, ux, uy, uz);
    }
    m = relax_prescribed<OP, LES, kOutLog2<S, OP>, S::kRegroupedReg>(p, g, flag,
                                                                    LES ? buffered_stress(p, flag, n) : float4{},
                                                   kind != kFluid, ux, uy, uz);
  } else if (kind == kEq) {
    if constexpr (EDGE) {
      RowMoments up{1.0f, 0.0f, 0.0f, 0.0f};
      if (x == p.nx - 1 && x > 0) up = row_moments(edge_rows(in, p)[y + p.ny * z]);
      m = eq_from(p, x, up, g);
    } else {
      m = eq_state<S>(in, p, n, x, g);
    }
  } else if (kind != kFluid) {
    rest_state(g);
    m = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  }
  const uint32_t h = S::cell_word(n, seed);
  store_state<S, POL != kCached, (kOutLog2<S, OP> != 0)>(out, P, self, g, h);
  if constexpr (EDGE) {
    if (x + 2 == p.nx) {
      const RowMoments e = encoded_moments<S, (kOutLog2<S, OP> != 0)>(g, h);
      edge_rows(out, p)[y + p.ny * z] = make_float4(e.mass, e.mx, e.my, e.mz);
    }
  }
  if (MACRO && !is_wall(kind)) {
    macro[n] = m;
    guard_density(macro, p, m.x);
  }
}

// Raw storage bits of one element, zero-extended to 32 bits, and their decoding.
template <class S, bool NT>
__device__ __forceinline__ uint32_t load_bits(const typename S::T* p) {
  const typename S::T v = ld<NT>(p);
  if constexpr (sizeof(typename S::T) == 4) return __builtin_bit_cast(uint32_t, v);
  else return (uint32_t)__builtin_bit_cast(uint16_t, v);
}
template <class S>
__device__ __forceinline__ float decode_bits(uint32_t b) {
  using T = typename S::T;
  if constexpr (sizeof(T) == 4) return S::decode(__builtin_bit_cast(T, b));
  else return S::decode(__builtin_bit_cast(T, (uint16_t)b));
}

}  // namespace pw
