// step_row.h -- the step on row wavefronts (nx a multiple of 64: one wavefront per 64-cell
// x-run) and the step kernel that dispatches between it and the per-cell step of
// step_kernels.h. The row step evaluates the learned wall closure inside its wall wavefronts
// (closure_features.h).
#pragma once

#include "closure_features.h"

namespace pw {

// A wavefront whose 64 cells lie in one x-row (nx a multiple of 64) shares the row: its
// first cell n0, coordinates x0, y, z, the source rows of its pulls and its store bases are
// wavefront-uniform and live in scalar registers, so loads and stores address with the lane's
// byte offset alone.
struct RowWave {
  uint32_t n0, x0, y, z;
  uint32_t e0;  // population element offset of n0 (its run starts there: x0 is a multiple of 64)
};

// Gather of a row wavefront. Each plane is read with aligned, unit-stride loads of the pull's
// source row, so every cache line of the plane is fetched once. An x-shifted pull then takes
// its neighbour lane's value with a DPP wavefront shift (wave_shr:1 moves lane l - 1 to lane
// l, wave_shl:1 lane l + 1 to lane l); the lane at the wavefront's edge takes the value one
// load brings for the whole wavefront. All 64 lanes must be active.
// Pull of direction I for a row wavefront from the source row `src` of its plane.
template <class S, bool NT, int I>
__device__ __forceinline__ float pull_row(const typename S::T* src, uint32_t xs, uint32_t left, uint32_t right) {
  const uint32_t own = load_bits<S, NT>(byte_offset(src, xs));
  if constexpr (cx(I) == 0) {
    return decode_bits<S>(own);
  } else {
    const uint32_t edge = load_bits<S, NT>(byte_offset(src, cx(I) > 0 ? left : right));
    const int v = cx(I) > 0 ? __builtin_amdgcn_update_dpp((int)edge, (int)own, 0x138, 0xf, 0xf, false)
                            : __builtin_amdgcn_update_dpp((int)edge, (int)own, 0x130, 0xf, 0xf, false);
    return decode_bits<S>((uint32_t)v);
  }
}

// Sources of lane 0 for c_x = +1 and of lane 63 for c_x = -1 (the zero-gradient outlet holds
// the last wavefront's at its own column), and the lane's own column, as byte offsets.
template <int SH>
__device__ __forceinline__ void row_columns(const StepParams& p, uint32_t bc, uint32_t x0, uint32_t lane,
                                            uint32_t& left, uint32_t& right, uint32_t& xs) {
  left = pop_x(p, below(x0, p.nx, false)) << SH;
  right = pop_x(p, above(x0 + 63, p.nx, bc & kBcZeroGradient)) << SH;
  xs = (pop_x(p, x0) + lane) << SH;
}

// On the row path every run holds one wavefront's 64 cells, so direction i's run starts i * 64
// elements after the row's first run: a compile-time offset from one scalar base per source
// row (nine rows feed a D3Q19 pull), which the loads and stores carry as immediates.
constexpr uint32_t kRowRun = 64;

template <class S, bool NT>
__device__ __forceinline__ void gather_row(const typename S::T* __restrict__ in, const StepParams& p, uint32_t bc,
                                           const RowWave& w, uint32_t lane, float* g) {
  constexpr int SH = S::kShift;
  const uint32_t row = kQ * p.nx, plane = row * p.ny;
  const uint32_t y = w.y, z = w.z;
  const uint32_t yo[3] = {((y == 0 ? p.ny - 1 : y - 1) * row) << SH, (y * row) << SH,
                          ((y + 1 == p.ny ? 0 : y + 1) * row) << SH};
  const uint32_t zo[3] = {((z == 0 ? p.nz - 1 : z - 1) * plane) << SH, (z * plane) << SH,
                          ((z + 1 == p.nz ? 0 : z + 1) * plane) << SH};
  uint32_t left, right, xs;
  row_columns<SH>(p, bc, w.x0, lane, left, right, xs);
  static_for<0, kQ>([&](auto c) {
    constexpr int i = decltype(c)::value;
    g[i] = pull_row<S, NT, i>(byte_offset(in, yo[1 - cy(i)] + zo[1 - cz(i)]) + i * kRowRun, xs, left, right);
  });
}

// Free-slip side row, row path: after the plain gather, the pulls that cross a side are read
// again, from the mirrored plane of the row unshifted in the crossed axis (all wavefront-uniform).
template <class S, bool NT>
__device__ __forceinline__ void slip_row(const typename S::T* __restrict__ in, const StepParams& p,
                                         const RowWave& w, uint32_t lane, float* g) {
  constexpr int SH = S::kShift;
  const uint32_t row = kQ * p.nx, plane = row * p.ny;
  const uint32_t y = w.y, z = w.z;
  const bool ylo = y == 0, yhi = y + 1 == p.ny, zlo = z == 0, zhi = z + 1 == p.nz;
  uint32_t left, right, xs;
  row_columns<SH>(p, p.bc, w.x0, lane, left, right, xs);
  static_for<1, kQ>([&](auto c) {
    constexpr int i = decltype(c)::value;
    if constexpr (cy(i) != 0 || cz(i) != 0) {
      const bool my = cy(i) > 0 ? ylo : (cy(i) < 0 ? yhi : false);
      const bool mz = cz(i) > 0 ? zlo : (cz(i) < 0 ? zhi : false);
      if (my || mz) {  // a pull that does not cross stays in range in that axis
        const uint32_t sy = my ? y : y - cy(i), sz = mz ? z : z - cz(i);
        const int j = slip_plane<i>(ylo, yhi, zlo, zhi);
        g[i] = pull_row<S, NT, i>(byte_offset(in + (size_t)j * p.run, (sy * row + sz * plane) << SH), xs, left,
                                  right);
      }
    }
  });
}

// Stores of a row wavefront: run i of the row at e0 (the wavefront's first element), lane
// `lane`. 16-bit formats round two directions into one register and store its halves.
// With `moments` (wavefront-uniform) the stored values' moments accumulate into m.
template <class S, bool NT, bool SCALED = false>
__device__ __forceinline__ void store_row(typename S::T* __restrict__ out, uint32_t e0, uint32_t lane,
                                          const float* g, uint32_t h, bool moments, RowMoments& m) {
  using T = typename S::T;
  if constexpr (sizeof(T) == 4) {
    T* b = out + e0 + lane;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const float v = encode_out<S, SCALED>(g[i], S::dir_word(h, i));
      st<NT>(b + i * kRowRun, v);
      if (moments) add_moment(m, i, v);
    }
  } else {
    // Every store first; the outlet wavefront's moments afterwards from the packed words.
    uint16_t* b = reinterpret_cast<uint16_t*>(out) + e0 + lane;
    const uint16_t v0 = __builtin_bit_cast(uint16_t, encode_out<S, SCALED>(g[0], S::dir_word(h, 0)));
    st<NT>(b, v0);
    uint32_t words[kPairs];
#pragma unroll
    for (int q = 0; q < kPairs; ++q) {
      const int i = 2 * q + 1;
      const float scale = SCALED ? 1.0f : kOut<S::kScaleLog2>;
      // The pair's first random word is dead after its conversion, so its register takes the pair.
      const uint32_t ra = S::dir_word(h, i);
      words[q] = S::encode_pair_scaled(ra, g[i] * scale, ra, g[i + 1] * scale, S::dir_word(h, i + 1));
      st<NT>(b + i * kRowRun, (uint16_t)(words[q] & 0xffffu));
      st<NT>(b + (i + 1) * kRowRun, (uint16_t)(words[q] >> 16));
    }
    if (moments) {
      add_moment(m, 0, decode_bits<S>(v0));
#pragma unroll
      for (int q = 0; q < kPairs; ++q) {
        add_moment(m, 2 * q + 1, decode_bits<S>(words[q] & 0xffffu));
        add_moment(m, 2 * q + 2, decode_bits<S>(words[q] >> 16));
      }
    }
  }
}

// The closure buffer's stress for the wall lanes (fluid cells with wall links) of a row
// wavefront, the only lanes whose collision reads it. The wall lanes of a run are consecutive in
// the wall-cell list from the base its descriptor carries (a row lattice always has descriptors
// when a buffer is attached), so a wall lane's entry is that base plus the wall lanes below it.
// A wavefront with wall lanes loads on every lane (the others take the base's entry, which
// exists), one load after the flags with no per-cell index in between. All 64 lanes must be
// active.
__device__ __forceinline__ float4 row_stress(const StepParams& p, uint32_t desc, uint32_t flag) {
  float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (!p.closure_tau) return t;  // the wall law: one scalar test, no lane work
  const bool wall_lane = (flag & kTypeMask) == kFluid && (flag >> 3) != 0;
  const uint64_t walls = __ballot(wall_lane);
  if (walls) {
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(walls >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)walls, 0u));
    t = p.closure_tau[(desc >> kWaveBaseShift) + (wall_lane ? rank : 0u)];
  }
  return t;
}

// stream_collide_cell for a row wavefront: the whole wavefront gathers (gather_row), then fluid
// lanes replace wall-link directions and collide, EQ and wall lanes take their fixed states,
// and every lane stores. Every load that does not depend on another (the pulls, the flags, the
// outlet's upstream moments) is issued before any of them is waited for, so a wavefront waits
// for memory once; a wavefront of walls only stores their rest state again like any other.
// CLOSURE: the instance that evaluates the learned closure in its wall wavefronts (REG + LES,
// p.closure_net set, `lds` this wavefront's closure::kFusedWaveBytes); other instances read
// the closure kernel's buffer, if any, and carry none of its code.
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, int POL, bool BC, bool CLOSURE = false>
__device__ __forceinline__ void stream_collide_row_wave(const typename S::T* __restrict__ in,
                                                        typename S::T* __restrict__ out,
                                                        const uint32_t* __restrict__ geo,
                                                        float4* __restrict__ macro, const StepParams& p,
                                                        uint32_t n, uint32_t seed, char* lds) {
  const uint32_t lane = threadIdx.x & 63u;
  RowWave w;
  w.n0 = __builtin_amdgcn_readfirstlane(n - lane);
  cell_coords(w.n0, p, w.x0, w.y, w.z);
  w.e0 = pop_x(p, w.x0) + pop_row(p, w.y + p.ny * w.z);
  const uint32_t row = w.y + p.ny * w.z;
  // The last wavefront of a row holds the outlet plane (lane 63) and its upstream cell (lane 62).
  const bool last_wave = w.x0 + 64 == p.nx;
  RowMoments up{1.0f, 0.0f, 0.0f, 0.0f};
  if (last_wave) up = row_moments(edge_rows(in, p)[row]);
  // The run's descriptor (a scalar load) tells plain runs, whose flags follow from the lane,
  // from the few that touch a body and read every cell's flag.
  const uint32_t desc = p.wave_desc ? p.wave_desc[w.n0 >> 6] : kWaveGeneral;
  // The learned closure of this wavefront's wall cells (only runs that touch a body hold any).
  static_assert(!CLOSURE || LES, "the closure belongs to the LES wall model");
  const bool fused = CLOSURE && (desc & kWaveGeneral);
  if (fused) {
    const uint32_t f0 = geo[n];
    closure::fused_closure<S, BOUZIDI>(in, geo, p, n, (f0 & kTypeMask) == kFluid && (f0 >> 3) != 0, lds);
  }
  // The flag load goes before the pulls for f32 storage and after them for 16-bit storage: on
  // the 256x128x128 body (same-box comparisons, 2000 steps) the f32 step measured 1-1.5 %
  // faster with the flag first and the f16 step 0.2-0.4 % slower.
  constexpr bool kFlagFirst = sizeof(typename S::T) == 4;
  uint32_t flag = 0;
  auto load_flag = [&] {
    if (desc & kWaveGeneral)
      flag = geo[n];
    else
      flag = ((lane == 0 && (desc & kWaveEqFirst)) || (lane == 63 && (desc & kWaveEqLast))) ? (uint32_t)kEq : 0u;
  };
  if constexpr (kFlagFirst) load_flag();
  float g[kQ];
  const uint32_t bc = BC ? p.bc : 0u;
  gather_row<S, POL == kNtAll>(in, p, bc, w, lane, g);
  if constexpr (!kFlagFirst) load_flag();
  const uint32_t kind = flag & kTypeMask;
  float4 learned = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  // Only runs that touch a body hold wall cells; the others skip the lookup on a scalar test.
  if constexpr (!CLOSURE && LES)
    if (desc & kWaveGeneral) learned = row_stress(p, desc, flag);
  const bool outlet_wave = last_wave && (__builtin_amdgcn_readlane(kind, 63) == kEq);
  if (BC && (p.bc & kBcFreeSlip) && on_side(w.y, w.z, p)) slip_row<S, POL == kNtAll>(in, p, w, lane, g);
  const uint32_t x = w.x0 + lane;
  if (kind == kFluid && (flag >> 3) != 0) {
    const Neighbours nb(x, w.y, w.z, p, bc);
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      if (has_wall(flag, i)) g[i] = wall_value<S, BOUZIDI>(in, geo, p, nb, n, w.e0 + lane, flag, i);
  }
  if constexpr (CLOSURE) {
    if (fused && kind == kFluid && (flag >> 3) != 0) learned = closure::fused_stress(lds);
  }
  float4 m;
  if constexpr (OP == kOpReg) {
    // Every lane through one collision: EQ and wall lanes with prescribed moments.
    const bool prescribed = kind != kFluid;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (kind == kEq) eq_velocity(p, x, up, ux, uy, uz);
    m = relax_prescribed<OP, LES, kOutLog2<S, OP>, S::kRegroupedReg>(p, g, flag, learned, prescribed, ux, uy, uz);
  } else if (kind == kFluid) {
    m = relax<OP, LES>(p, g, flag, learned);
  } else if (kind == kEq) {
    m = eq_from(p, x, up, g);
  } else {
    rest_state(g);
    m = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  }
  RowMoments stored{1.0f, 0.0f, 0.0f, 0.0f};
  store_row<S, POL != kCached, (kOutLog2<S, OP> != 0)>(out, w.e0, lane, g, S::cell_word(n, seed), outlet_wave,
                                                          stored);
  if (outlet_wave && lane == 62) edge_rows(out, p)[row] = make_float4(stored.mass, stored.mx, stored.my, stored.mz);
  if (MACRO && !is_wall(kind)) {
    macro[n] = m;
    guard_density(macro, p, m.x);
  }
}

// Blocks are renumbered so that each of the 8 XCDs (which receive blocks round-robin) steps
// contiguous chunks of `chunk` blocks: XCD k takes chunks k, k + 8, k + 16, ... The cache lines
// that neighbouring rows and wavefronts share stay in one XCD's L2 within a chunk. One chunk per
// XCD (an eighth of the lattice each) suits uniform lattices; smaller chunks spread a body's
// wall cells, and their slower wavefronts, over all XCDs. The grid holds a multiple of 8 chunks.
// A row-wavefront pull copy with no arithmetic (f16, 256x128x128, 128 threads per block) took
// 58.8 us per step with 1 chunk per XCD, 55.5 us with 16, 53.8 us with 64 and 53.5 us with 256
// (16 blocks per chunk); the autotuner times the chunkings in kXcdChunkings.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdChunkings[] = {1, 16, 64, 256};
constexpr uint32_t kMaxXcdChunks = 1024;
// Chunks per XCD a step launch accepts: a power of two up to kMaxXcdChunks.
__host__ __device__ constexpr bool valid_xcd_chunks(uint32_t c) {
  return c >= 1 && c <= kMaxXcdChunks && (c & (c - 1)) == 0;
}

// Waves per SIMD the step kernel is compiled for (a register budget of 512 / waves VGPRs).
// Seven leaves REG + LES a few spills and measured fastest; TRT + LES keeps the populations,
// their pair sums and the equilibrium live through the wall function and spills heavily
// below about 106 registers, so it runs at four.
template <int OP, bool LES>
constexpr int kStepWaves = (LES && OP == kOpTrt) ? 4 : 7;

// BC: the instance for boundary options (p.bc != 0); the reference's boundaries take BC = false,
// which compiles the option code out of the step. ALONG: the instance whose first
// p.closure_blocks blocks evaluate the learned closure over the wall-cell list (one 16-cell tile
// per wavefront, closure::closure_tile) into p.closure_out while the rest of the grid steps; the
// launch dispatches them first, so their dependent gathers overlap the step's streaming instead
// of preceding it. p.closure_blocks is a multiple of the XCD count, so the step blocks keep
// their XCD.
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, int POL, bool BC, bool CLOSURE = false,
          bool ALONG = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kStepWaves<OP, LES>))) void stream_collide_kernel(const typename S::T* __restrict__ in,
                                                             typename S::T* __restrict__ out,
                                                             const uint32_t* __restrict__ geo,
                                                             float4* __restrict__ macro, StepParams p,
                                                             uint32_t seed, uint32_t chunk) {
  // Dynamic LDS of the closure instances: closure::kFusedWaveBytes per wavefront.
  extern __shared__ __attribute__((aligned(16))) char step_lds[];
  uint32_t b = blockIdx.x;
  if constexpr (ALONG) {
    if (b < p.closure_blocks) {
      const uint32_t tile = b * (blockDim.x >> 6) + (threadIdx.x >> 6);
      if (tile * closure::kCellsPerWave < p.closure_cells)  // wavefront-uniform: whole tiles only
        closure::closure_tile<S, BOUZIDI>(
            in, geo, p, static_cast<const closure::WallGeo*>(p.closure_geo), p.closure_cells,
            static_cast<const mlp::Packed*>(p.closure_net),
            *reinterpret_cast<const closure::Norm*>(static_cast<const char*>(p.closure_net) + closure::kNormOffset),
            p.closure_out, nullptr, tile, step_lds + (threadIdx.x >> 6) * closure::kFusedWaveBytes);
      return;
    }
    b -= p.closure_blocks;
  }
  const uint32_t local = b / kXcds;
  uint32_t logical = ((local / chunk) * kXcds + b % kXcds) * chunk + local % chunk;
  if constexpr (CLOSURE) logical = p.tile_order[logical];
  const uint32_t n = logical * blockDim.x + threadIdx.x;
  if (n >= p.cells) return;  // whole wavefronts when nx is a multiple of 64
  if constexpr (CLOSURE) {
    // The launch guarantees p.rows.
    stream_collide_row_wave<S, OP, LES, BOUZIDI, MACRO, POL, BC, true>(
        in, out, geo, macro, p, n, seed, step_lds + (threadIdx.x >> 6) * closure::kFusedWaveBytes);
  } else if (p.rows)
    stream_collide_row_wave<S, OP, LES, BOUZIDI, MACRO, POL, BC>(in, out, geo, macro, p, n, seed, nullptr);
  else
    stream_collide_cell<S, OP, LES, BOUZIDI, MACRO, POL, true, BC>(in, out, geo, macro, p, n, seed);
}

}  // namespace pw
