// batch_f32.hip -- batched step kernels with f32 population storage.
#include "batch_step.h"

namespace pw {
namespace batch {
template void batch_step_s<StoreF32>(const BatchLaunch&, int, bool);
}  // namespace batch
}  // namespace pw
