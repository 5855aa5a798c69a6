// step_bf16.hip -- step kernels with bf16 population storage.
#include "step_launch.h"

namespace pw {
template void launch_step_s<StoreBF16>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
}  // namespace pw
