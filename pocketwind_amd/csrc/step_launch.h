// step_launch.h -- launch of the step-kernel instance selected by a solver's options and
// step variant. Instantiated once per storage type in its own translation unit
// (step_f32.hip, step_f16.hip, step_bf16.hip) so the builds run in parallel.
#pragma once

#include "step_kernels.h"

namespace pw {

constexpr int kStepBlock = 256;

inline unsigned step_blocks_for(size_t n) { return (unsigned)((n + kStepBlock - 1) / kStepBlock); }

template <class S, int OP, bool LES, bool BB>
void launch_step_t(pw_solver* s, const void* in, void* out, bool macro, const StepParams& p) {
  constexpr bool BOUZIDI = !BB;
  const bool nt = s->kernel == PW_KERNEL_CELL_NT || s->kernel == PW_KERNEL_PAIR_NT;
  if (s->kernel == PW_KERNEL_PAIR || s->kernel == PW_KERNEL_PAIR_NT) {
    auto k = macro ? (nt ? stream_collide_pair_kernel<S, OP, LES, BOUZIDI, true, true>
                         : stream_collide_pair_kernel<S, OP, LES, BOUZIDI, true, false>)
                   : (nt ? stream_collide_pair_kernel<S, OP, LES, BOUZIDI, false, true>
                         : stream_collide_pair_kernel<S, OP, LES, BOUZIDI, false, false>);
    hipLaunchKernelGGL(k, dim3(step_blocks_for(s->cells / 2)), dim3(kStepBlock), 0, s->stream,
                       (const typename S::T*)in, (typename S::T*)out, s->geometry, s->macro, p);
    return;
  }
  auto k = macro ? (nt ? stream_collide_kernel<S, OP, LES, BOUZIDI, true, true> : stream_collide_kernel<S, OP, LES, BOUZIDI, true, false>)
                 : (nt ? stream_collide_kernel<S, OP, LES, BOUZIDI, false, true> : stream_collide_kernel<S, OP, LES, BOUZIDI, false, false>);
  hipLaunchKernelGGL(k, dim3(step_blocks_for(s->cells)), dim3(kStepBlock), 0, s->stream,
                     (const typename S::T*)in, (typename S::T*)out, s->geometry, s->macro, p);
}

template <class S>
void launch_step_s(pw_solver* s, const void* in, void* out, bool macro, const StepParams& p) {
  const bool reg = s->params.op == PW_OP_REG, les = s->params.les != 0, bb = s->params.walls == PW_WALLS_BB;
  if (reg) {
    if (les) bb ? launch_step_t<S, kOpReg, true, true>(s, in, out, macro, p) : launch_step_t<S, kOpReg, true, false>(s, in, out, macro, p);
    else bb ? launch_step_t<S, kOpReg, false, true>(s, in, out, macro, p) : launch_step_t<S, kOpReg, false, false>(s, in, out, macro, p);
  } else {
    if (les) bb ? launch_step_t<S, kOpTrt, true, true>(s, in, out, macro, p) : launch_step_t<S, kOpTrt, true, false>(s, in, out, macro, p);
    else bb ? launch_step_t<S, kOpTrt, false, true>(s, in, out, macro, p) : launch_step_t<S, kOpTrt, false, false>(s, in, out, macro, p);
  }
}


}  // namespace pw
