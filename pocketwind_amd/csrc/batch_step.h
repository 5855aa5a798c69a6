// batch_step.h -- the step kernel of batched tunnels and its launch, instantiated
// once per storage type (batch_f32.hip, batch_f16.hip, batch_bf16.hip).
#pragma once

#include "force_kernels.h"
#include "step_row.h"

namespace pw {
namespace batch {

constexpr int kBlock = 256;  // largest block of a batched step (the kernels' launch bound)
constexpr uint32_t kBatchBlocks[] = {64, 128, 256};  // the block sizes the tuner times

struct Tunnel {
  StepParams p;  // p.closure_tau / p.closure_slot (closure kernel) or p.closure_net (in the step) once attached
  void* pops[2];
  const uint32_t* geo;
  float4* macro;
  const uint32_t* boundary;
  uint32_t n_boundary;
  uint32_t partial_base;  // first force partial of this tunnel
  uint32_t n_partials;
  const uint32_t* wall_cells;  // learned closure: fluid cells with wall links, in cell order
  uint32_t n_wall;
  const void* wall_geo;  // their closure::WallGeo records
  float4* closure_buf;   // the closure kernel's output for them (the step reads it as p.closure_tau)
};

struct BatchLaunch {
  unsigned grid;
  hipStream_t stream;
  const Tunnel* tunnels;
  const uint32_t* tile_tunnel;
  const uint32_t* tile_first;
  uint32_t n_tiles;
  uint32_t* seeds;  // device: seeds[parity] is the step count a launch of that parity runs
  pw_params params;
  uint32_t policy;  // population cache policy (kCached, kNtAll, kNtStores)
  uint32_t chunk;   // consecutive tiles one XCD takes (the grid holds a multiple of 8 chunks)
  uint32_t lds;     // dynamic LDS bytes per block: the closure evaluated in the step, else 0
  // Persistent form (steps without macro writes, the closure not in the step): this many
  // resident blocks (a multiple of 8) walk the `grid` block indices of the one-per-tile form;
  // 0 launches one block per tile.
  uint32_t persistent;
  uint32_t block;  // threads per block = cells per tile (64, 128 or 256)
};

// One step of every tunnel. The rounding seed is the step count, read on the device so a
// replayed graph advances it: a launch of parity k reads seeds[k] and writes seeds[k ^ 1] for
// the next launch, which in stream order starts after this one ends.
// A tunnel whose nx is a multiple of 64 steps its cells as row wavefronts (the single solver's
// fast path; its sets carry the upstream-outlet moments after the 19 N elements), others per
// cell. Either way a tile is one block of one tunnel.
// Block b of the one-per-tile grid steps one tile; blocks are renumbered as in the single
// solver's step so that each XCD takes runs of `chunk` consecutive tiles (neighbouring rows of
// one tunnel share its L2).
// CLOSURE: row tunnels evaluate the learned closure in their wall wavefronts (REG + LES).
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, int POL, bool CLOSURE>
__device__ __forceinline__ void batch_block(const Tunnel* __restrict__ tunnels, const uint32_t* __restrict__ tile_tunnel,
                                            const uint32_t* __restrict__ tile_first, uint32_t n_tiles, int parity,
                                            uint32_t seed, uint32_t chunk, uint32_t b) {
  const uint32_t local = b / kXcds;
  const uint32_t tile = ((local / chunk) * kXcds + b % kXcds) * chunk + local % chunk;
  if (tile >= n_tiles) return;
  const uint32_t t = tile_tunnel[tile];
  const Tunnel& tu = tunnels[t];
  const uint32_t n = (tile - tile_first[t]) * blockDim.x + threadIdx.x;
  if (n >= tu.p.cells) return;  // whole wavefronts on row tunnels (their cell count is a multiple of 64)
  const auto* in = (const typename S::T*)tu.pops[parity];
  auto* out = (typename S::T*)tu.pops[parity ^ 1];
  if constexpr (CLOSURE) {
    extern __shared__ __attribute__((aligned(16))) char step_lds[];
    if (tu.p.rows) {
      stream_collide_row_wave<S, OP, LES, BOUZIDI, MACRO, POL, false, true>(
          in, out, tu.geo, tu.macro, tu.p, n, seed, step_lds + (threadIdx.x >> 6) * closure::kFusedWaveBytes);
      return;
    }
  } else if (tu.p.rows) {
    stream_collide_row_wave<S, OP, LES, BOUZIDI, MACRO, POL, false>(in, out, tu.geo, tu.macro, tu.p, n, seed, nullptr);
    return;
  }
  stream_collide_cell<S, OP, LES, BOUZIDI, MACRO, POL>(in, out, tu.geo, tu.macro, tu.p, n, seed);
}

// PERSIST: gridDim.x resident blocks walk the `blocks` block indices of the one-per-tile grid
// (block b takes b, b + gridDim.x, ...; gridDim.x is a multiple of 8, so every index stays on the
// XCD that the one-per-tile launch gives it). No block waits for another, so no launch shape can
// deadlock; the tuner times it against one block per tile.
template <class S, int OP, bool LES, bool BOUZIDI, bool MACRO, int POL, bool CLOSURE = false, bool PERSIST = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kStepWaves<OP, LES>))) void batch_step_kernel(const Tunnel* __restrict__ tunnels,
                                                            const uint32_t* __restrict__ tile_tunnel,
                                                            const uint32_t* __restrict__ tile_first,
                                                            uint32_t n_tiles, int parity, uint32_t* seeds,
                                                            uint32_t chunk, uint32_t blocks) {
  const uint32_t seed = seeds[parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) seeds[parity ^ 1] = seed + 1;
  if constexpr (PERSIST) {
    for (uint32_t b = blockIdx.x; b < blocks; b += gridDim.x)
      batch_block<S, OP, LES, BOUZIDI, MACRO, POL, CLOSURE>(tunnels, tile_tunnel, tile_first, n_tiles, parity, seed,
                                                            chunk, b);
  } else {
    batch_block<S, OP, LES, BOUZIDI, MACRO, POL, CLOSURE>(tunnels, tile_tunnel, tile_first, n_tiles, parity, seed,
                                                          chunk, blockIdx.x);
  }
}

// Grid of a batched step: the tiles rounded up to whole chunks on each of the 8 XCDs.
inline unsigned batch_grid(uint32_t n_tiles, uint32_t chunks_per_xcd, uint32_t* chunk) {
  const uint32_t per_xcd = ((n_tiles + kXcds - 1) / kXcds + chunks_per_xcd - 1) / chunks_per_xcd * chunks_per_xcd;
  *chunk = per_xcd / chunks_per_xcd;
  return kXcds * per_xcd;
}

template <class S, int OP, bool LES, bool BZ, bool MACRO>
void batch_step_t(const BatchLaunch& L, int parity) {
  auto k = L.policy == kNtAll ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtAll>
           : L.policy == kNtStores ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtStores>
                                   : batch_step_kernel<S, OP, LES, BZ, MACRO, kCached>;
  unsigned grid = L.grid;
  if constexpr (!MACRO) {
    if (L.persistent && !L.lds && L.persistent < L.grid) {
      k = L.policy == kNtAll ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtAll, false, true>
          : L.policy == kNtStores ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtStores, false, true>
                                  : batch_step_kernel<S, OP, LES, BZ, MACRO, kCached, false, true>;
      grid = L.persistent;
    }
  }
  if constexpr (OP == kOpReg && LES) {
    if (L.lds)
      k = L.policy == kNtAll ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtAll, true>
          : L.policy == kNtStores ? batch_step_kernel<S, OP, LES, BZ, MACRO, kNtStores, true>
                                  : batch_step_kernel<S, OP, LES, BZ, MACRO, kCached, true>;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(L.block), L.lds, L.stream, L.tunnels, L.tile_tunnel, L.tile_first, L.n_tiles,
                     parity, L.seeds, L.chunk, L.grid);
}

// Resident blocks per CU of the persistent form for a collision: the kernel's waves per SIMD
// (its register budget) on the CU's 4 SIMDs over the waves of a block.
template <int OP, bool LES>
constexpr uint32_t persistent_blocks_per_cu(uint32_t block) {
  return (uint32_t)(kStepWaves<OP, LES> * 4) / (block / 64);
}

template <class S, bool MACRO>
void batch_step_m(const BatchLaunch& L, int parity) {
  const bool reg = L.params.op == PW_OP_REG, les = L.params.les != 0, bz = L.params.walls == PW_WALLS_BOUZIDI;
  if (reg) {
    if (les) bz ? batch_step_t<S, kOpReg, true, true, MACRO>(L, parity) : batch_step_t<S, kOpReg, true, false, MACRO>(L, parity);
    else bz ? batch_step_t<S, kOpReg, false, true, MACRO>(L, parity) : batch_step_t<S, kOpReg, false, false, MACRO>(L, parity);
  } else {
    if (les) bz ? batch_step_t<S, kOpTrt, true, true, MACRO>(L, parity) : batch_step_t<S, kOpTrt, true, false, MACRO>(L, parity);
    else bz ? batch_step_t<S, kOpTrt, false, true, MACRO>(L, parity) : batch_step_t<S, kOpTrt, false, false, MACRO>(L, parity);
  }
}

template <class S>
void batch_step_s(const BatchLaunch& L, int parity, bool macro) {
  macro ? batch_step_m<S, true>(L, parity) : batch_step_m<S, false>(L, parity);
}

}  // namespace batch
}  // namespace pw
