// step_launch.h -- launch of the step-kernel instance selected by a solver's options and
// step variant. Instantiated once per storage type in its own translation unit
// (step_f32.hip, step_f16.hip, step_bf16.hip) so the builds run in parallel.
#pragma once

#include <type_traits>

#include "step_row.h"

namespace pw {

constexpr int kStepBlock = 256;  // largest block of a step launch (the kernels' launch bound)

template <class S, int OP, bool LES, bool BB>
void launch_step_t(pw_solver* s, const void* in, void* out, bool macro, const StepParams& p, uint32_t seed) {
  constexpr bool BOUZIDI = !BB;
  auto pick = [&](auto with_macro, auto with_bc) {
    constexpr bool M = decltype(with_macro)::value, BC = decltype(with_bc)::value;
    switch (s->kernel) {
      case PW_KERNEL_CELL_NT: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtAll, BC>;
      case PW_KERNEL_CELL_NT_STORES: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtStores, BC>;
      default: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kCached, BC>;
    }
  };
  // Boundary options select their own instances, so the reference's boundaries run a step
  // without their code.
  auto k = p.bc ? (macro ? pick(std::true_type{}, std::true_type{}) : pick(std::false_type{}, std::true_type{}))
                : (macro ? pick(std::true_type{}, std::false_type{}) : pick(std::false_type{}, std::false_type{}));
  // The learned closure evaluated in the step: its own instances (REG + LES, the reference's
  // boundaries, row path; launch_closure sets p.closure_net only for those).
  if constexpr (OP == kOpReg && LES) {
    if (p.closure_net) {
      auto pick_c = [&](auto with_macro) {
        constexpr bool M = decltype(with_macro)::value;
        switch (s->kernel) {
          case PW_KERNEL_CELL_NT: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtAll, false, true>;
          case PW_KERNEL_CELL_NT_STORES: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtStores, false, true>;
          default: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kCached, false, true>;
        }
      };
      k = macro ? pick_c(std::true_type{}) : pick_c(std::false_type{});
    }
    // The learned closure evaluated beside the step (launch_closure sets p.closure_out only for
    // the reference's boundaries).
    if (p.closure_out) {
      auto pick_a = [&](auto with_macro) {
        constexpr bool M = decltype(with_macro)::value;
        switch (s->kernel) {
          case PW_KERNEL_CELL_NT: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtAll, false, false, true>;
          case PW_KERNEL_CELL_NT_STORES:
            return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kNtStores, false, false, true>;
          default: return stream_collide_kernel<S, OP, LES, BOUZIDI, M, kCached, false, false, true>;
        }
      };
      k = macro ? pick_a(std::true_type{}) : pick_a(std::false_type{});
    }
  }
  const uint32_t block = s->block;
  const uint64_t blocks = (s->cells + block - 1) / block;
  // Each XCD takes s->xcd_chunks chunks; the grid is rounded up to whole chunks (the extra
  // blocks fall beyond the lattice and return at once).
  const uint32_t parts = s->xcd_chunks;
  const uint32_t per_xcd = (uint32_t)(((blocks + kXcds - 1) / kXcds + parts - 1) / parts * parts);
  const uint32_t chunk = per_xcd / parts;
  const uint32_t lds = p.closure_net ? block / 64 * closure::kFusedWaveBytes : 0;
  hipLaunchKernelGGL(k, dim3(kXcds * per_xcd + p.closure_blocks), dim3(block), lds, s->stream, (const typename S::T*)in,
                     (typename S::T*)out, s->geometry, s->macro, p, seed, chunk);
}

template <class S>
void launch_step_s(pw_solver* s, const void* in, void* out, bool macro, const StepParams& p, uint32_t seed) {
  const bool reg = s->params.op == PW_OP_REG, les = s->params.les != 0, bb = s->params.walls == PW_WALLS_BB;
  if (reg) {
    if (les) bb ? launch_step_t<S, kOpReg, true, true>(s, in, out, macro, p, seed) : launch_step_t<S, kOpReg, true, false>(s, in, out, macro, p, seed);
    else bb ? launch_step_t<S, kOpReg, false, true>(s, in, out, macro, p, seed) : launch_step_t<S, kOpReg, false, false>(s, in, out, macro, p, seed);
  } else {
    if (les) bb ? launch_step_t<S, kOpTrt, true, true>(s, in, out, macro, p, seed) : launch_step_t<S, kOpTrt, true, false>(s, in, out, macro, p, seed);
    else bb ? launch_step_t<S, kOpTrt, false, true>(s, in, out, macro, p, seed) : launch_step_t<S, kOpTrt, false, false>(s, in, out, macro, p, seed);
  }
}

}  // namespace pw
