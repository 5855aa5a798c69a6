// step_f16_reference.hip -- step kernels with f16 population storage rounded by the reference's
// rule (prelude-f16.wgsl:7-22), selected per solver.
#include "step_launch.h"

namespace pw {
template void launch_step_s<StoreF16Reference>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
}  // namespace pw
