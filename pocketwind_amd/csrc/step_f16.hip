// step_f16.hip -- step kernels with f16 population storage.
#include "step_launch.h"

namespace pw {
template void launch_step_s<StoreF16>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
}  // namespace pw
