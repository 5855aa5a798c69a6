// closure_features.h -- features of the learned wall closure and its evaluation inside the step.
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

#include "mlp.h"
#include "step_kernels.h"

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

// The static part of a wall cell's features, from the packed geometry alone: computed once per
// geometry (wall_geo_kernel) so that evaluating the features reads a 48-byte record and the two
// cells' populations instead of chasing flags, link fractions and the neighbour's flag.
struct WallGeo {
  uint32_t n, flag;    // the wall cell and its flag
  uint32_t m, flag_m;  // the fluid cell one step along the rounded normal and its flag (else n)
  float ax, ay, az;    // unit wall normal, n = sum over wall links of w_i c_i (zero when |n| <= 1e-3)
  float q;             // mean wall-link fraction (1/2 when the geometry carries none)
  float exposed6;      // 6 sum over wall links of w_i
  uint32_t bits;       // kUseNeighbour | kProjected
  uint32_t pad[2];
};
static_assert(sizeof(WallGeo) == 48, "three 16-byte loads per record");
enum : uint32_t { kUseNeighbour = 1u, kProjected = 2u };

__device__ __forceinline__ WallGeo wall_geo_of(const uint32_t* __restrict__ geo, const StepParams& p, uint32_t n) {
  const size_t N = p.cells;
  WallGeo w;
  w.n = n;
  w.flag = geo[n];
  float nx = 0.0f, ny = 0.0f, nz = 0.0f, exposed = 0.0f, qsum = 0.0f;
  int links = 0;
  const uint32_t base = geo[N + n];
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    if (!has_wall(w.flag, i)) continue;
    nx += wf(i) * cx(i);
    ny += wf(i) * cy(i);
    nz += wf(i) * cz(i);
    exposed += wf(i);
    float q = 0.5f;  // the geometry's link fraction when it carries one, whatever the wall rule
    if (base != kNoQ) q = as_float(geo[2 * N + base + __builtin_popcount((w.flag >> 3) & ((1u << i) - 1u))]);
    qsum += q;
    ++links;
  }
  const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
  const bool projected = nl > 1e-3f;
  w.ax = projected ? nx / nl : 0.0f;
  w.ay = projected ? ny / nl : 0.0f;
  w.az = projected ? nz / nl : 0.0f;
  w.q = links ? qsum / links : 0.5f;
  w.exposed6 = 6.0f * exposed;
  w.bits = projected ? kProjected : 0u;
  w.m = n;
  w.flag_m = w.flag;
  // Second sample: the fluid cell one step along the rounded normal.
  const int dx = (int)rintf(w.ax), dy = (int)rintf(w.ay), dz = (int)rintf(w.az);
  if (dx | dy | dz) {
    uint32_t x, y, z;
    cell_coords(n, p, x, y, z);
    const uint32_t m = Neighbours(x, y, z, p).index(dx, dy, dz);
    const uint32_t fm = geo[m];
    if ((fm & kTypeMask) == kFluid) {
      w.m = m;
      w.flag_m = fm;
      w.bits |= kUseNeighbour;
    }
  }
  w.pad[0] = w.pad[1] = 0u;
  return w;
}

__device__ __forceinline__ WallGeo load_wall_geo(const WallGeo* __restrict__ records, uint32_t k) {
  const uint4* r = reinterpret_cast<const uint4*>(records + k);
  const uint4 a = r[0], b = r[1], c = r[2];
  WallGeo w;
  w.n = a.x;
  w.flag = a.y;
  w.m = a.z;
  w.flag_m = a.w;
  w.ax = as_float(b.x);
  w.ay = as_float(b.y);
  w.az = as_float(b.z);
  w.q = as_float(b.w);
  w.exposed6 = as_float(c.x);
  w.bits = c.y;
  w.pad[0] = w.pad[1] = 0u;
  return w;
}

// The populations of fluid cell n (flag its flag) gathered exactly as the step gathers them,
// wall links included. Every direction's load is independent of the others.
template <class S, bool BOUZIDI>
__device__ __forceinline__ void gather_cell(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                            const StepParams& p, uint32_t n, uint32_t flag, float* g) {
  const size_t P = p.run;
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

// Features of the wall cell whose static geometry is w, and the wall law's stress vector
// rho u_tau^2 t (stream_collide.wgsl:18-80) for comparison.
template <class S, bool BOUZIDI>
__device__ __forceinline__ void wall_features(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                              const StepParams& p, float u_ref, const WallGeo& w, float* f,
                                              float* law) {
  float g[kQ], h[kQ];
  gather_cell<S, BOUZIDI>(in, geo, p, w.n, w.flag, g);
  const bool second = w.bits & kUseNeighbour;
  if (second) gather_cell<S, BOUZIDI>(in, geo, p, w.m, w.flag_m, h);
  float rho, ux, uy, uz;
  velocity_of(g, rho, ux, uy, uz);
  float vx = ux, vy = uy, vz = uz;
  if (second) {
    float r2;
    velocity_of(h, r2, vx, vy, vz);
  }
  const float inv = 1.0f / u_ref;
  f[0] = ux * inv;
  f[1] = uy * inv;
  f[2] = uz * inv;
  f[3] = vx * inv;
  f[4] = vy * inv;
  f[5] = vz * inv;
  f[6] = w.q;
  f[7] = log10f(p.nu);
  f[8] = w.ax;
  f[9] = w.ay;
  f[10] = w.az;
  f[11] = w.exposed6;
  float sx = ux, sy = uy, sz = uz;
  if (w.bits & kProjected) {
    const float d = ux * w.ax + uy * w.ay + uz * w.az;
    sx -= d * w.ax;
    sy -= d * w.ay;
    sz -= d * w.az;
  }
  const float speed = sqrtf(sx * sx + sy * sy + sz * sz);
  law[0] = law[1] = law[2] = 0.0f;
  if (speed >= p.slip_floor) {
    const float ut = friction_velocity(p.nu, speed);
    const float st = rho * ut * ut / speed;
    law[0] = st * sx;
    law[1] = st * sy;
    law[2] = st * sz;
  }
}

// Inside the step: a wavefront of the row step that holds wall cells (fluid cells with wall
// links) evaluates the network for them before its own stream-collide, from the state the
// step reads. Their features come from wall_features, as the closure kernel's do; the
// wavefront packs its wall lanes into 16-row matrix-core tiles (in lane order) and runs
// mlp::forward16 per tile, so a cell's stress is the closure kernel's bit for bit (the rows of
// a product do not mix). Each lane's result (x, y, z, w > 0 valid) is left in this wavefront's
// LDS, read back by its wall collision.
//
// Per-wavefront LDS: h1 (16 x 64 bf16), then the inputs (16 x 32 bf16) which h2 (16 x 64 bf16)
// overwrites, then the tile's outputs (16 x 4 f32); the 64 lanes' results take h1's place
// after the last tile.
constexpr uint32_t kFusedH1 = 0, kFusedX = 16 * mlp::kHidden * 2, kFusedY = kFusedX + 16 * mlp::kHidden * 2;
constexpr uint32_t kFusedWaveBytes = kFusedY + 16 * 4 * 4;
static_assert(64 * sizeof(float4) <= kFusedX, "results fit in h1's place");

// The device block of an attached network: the matrix-core copy, then its normalisation.
constexpr size_t kNormOffset = (sizeof(mlp::Packed) + 255) & ~(size_t)255;
constexpr size_t kNetBlockBytes = kNormOffset + ((sizeof(Norm) + 255) & ~(size_t)255);

template <class S, bool BOUZIDI>
__device__ __forceinline__ void fused_closure(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                              const StepParams& p, uint32_t n, bool wall_lane, char* lds) {
  const mlp::Packed* P = static_cast<const mlp::Packed*>(p.closure_net);
  const Norm& norm = *reinterpret_cast<const Norm*>(static_cast<const char*>(p.closure_net) + kNormOffset);
  uint16_t* h1 = reinterpret_cast<uint16_t*>(lds + kFusedH1);
  uint16_t* xs = reinterpret_cast<uint16_t*>(lds + kFusedX);
  float* ys = reinterpret_cast<float*>(lds + kFusedY);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t mask = __ballot(wall_lane);
  const uint32_t count = __builtin_popcountll(mask);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
  float f[kFeatures];
  bool in_range = wall_lane;
  if (wall_lane) {
    float law[3];
    wall_features<S, BOUZIDI>(in, geo, p, norm.u_ref,
                              load_wall_geo(static_cast<const WallGeo*>(p.closure_geo), p.closure_slot[n]), f, law);
  }
  float4 tau = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (uint32_t t = 0; t < count; t += 16) {
    mlp::wave_sync();  // the previous tile's reads of h2 and the outputs are done
    if (wall_lane && rank - t < 16u) {
      const uint32_t r = rank - t;
      for (int j = 0; j < mlp::kInPad; ++j) {
        float v = 0.0f;
        if (j < kFeatures) {
          v = (f[j] - norm.mean[j]) * norm.scale[j];
          in_range = in_range && fabsf(v) <= norm.limit;
        }
        xs[r * mlp::kInPad + j] = mlp::bf16_bits(v);
      }
    } else if (lane < 16 && t + lane >= count) {
      for (int j = 0; j < mlp::kInPad; ++j) xs[lane * mlp::kInPad + j] = 0;
    }
    mlp::wave_sync();
    const int r = lane & 15, gq = lane >> 4;
    const mlp::f32x4_t y = mlp::forward16(P, xs, h1, xs);
    if (r < 3)
      for (int j = 0; j < 4; ++j) ys[(4 * gq + j) * 4 + r] = y[j];
    mlp::wave_sync();
    if (wall_lane && rank - t < 16u) {
      const uint32_t row = rank - t;
      const float tx = ys[row * 4] * norm.out_scale, ty = ys[row * 4 + 1] * norm.out_scale,
                  tz = ys[row * 4 + 2] * norm.out_scale;
      const bool ok = in_range && isfinite(tx) && isfinite(ty) && isfinite(tz);
      tau = make_float4(tx, ty, tz, ok ? 1.0f : 0.0f);
    }
  }
  mlp::wave_sync();
  reinterpret_cast<float4*>(lds + kFusedH1)[lane] = tau;
  mlp::wave_sync();
}

// The stress a wall collision of this wavefront uses: the fused closure's result in LDS.
__device__ __forceinline__ float4 fused_stress(const char* lds) {
  return reinterpret_cast<const float4*>(lds + kFusedH1)[threadIdx.x & 63u];
}

// Wall cells per wavefront: one 16-row matrix-core tile. The wall cells are few (thousands),
// so one tile per wavefront spreads them over many more wavefronts than four tiles would, and
// each wavefront runs one network pass instead of four in sequence.
constexpr int kCellsPerWave = 16;

// One 16-cell tile of the wall-cell list (cells 16 tile .. 16 tile + 15) by one whole
// wavefront: lanes 0-15 gather their cell's features (written raw to `records` when given),
// then, when a network is attached (P non-null), all 64 lanes run its matrix-core tile and lanes
// 0-15 store their stress into tau (w = 1 valid, 0 use the wall law). `lds` is this
// wavefront's kFusedWaveBytes of LDS (h1, the inputs that h2 overwrites, the outputs).
template <class S, bool BOUZIDI>
__device__ __forceinline__ void closure_tile(const typename S::T* __restrict__ in, const uint32_t* __restrict__ geo,
                                             const StepParams& p, const WallGeo* __restrict__ walls,
                                             uint32_t n_cells, const mlp::Packed* __restrict__ P, const Norm& norm,
                                             float4* __restrict__ tau, float* __restrict__ records, uint32_t tile,
                                             char* lds) {
  uint16_t* h1 = reinterpret_cast<uint16_t*>(lds + kFusedH1);
  uint16_t* xs = reinterpret_cast<uint16_t*>(lds + kFusedX);
  float* ys = reinterpret_cast<float*>(lds + kFusedY);
  const int l = threadIdx.x & 63;
  const uint32_t k = tile * kCellsPerWave + l;
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
  const mlp::f32x4_t y = mlp::forward16(P, xs, h1, xs);
  if (r < 3)
    for (int j = 0; j < 4; ++j) ys[(4 * gq + j) * 4 + r] = y[j];
  mlp::wave_sync();
  if (!valid) return;
  const float tx = ys[l * 4] * norm.out_scale, ty = ys[l * 4 + 1] * norm.out_scale, tz = ys[l * 4 + 2] * norm.out_scale;
  const bool ok = in_range && isfinite(tx) && isfinite(ty) && isfinite(tz);
  tau[k] = make_float4(tx, ty, tz, ok ? 1.0f : 0.0f);
}

}  // namespace closure
}  // namespace pw
