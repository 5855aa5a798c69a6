// step_f32.hip -- step kernels with f32 population storage.
#include "step_launch.h"

namespace pw {
template void launch_step_s<StoreF32>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
}  // namespace pw
