// batch_bf16.hip -- batched step kernels with bf16 population storage.
#include "batch_step.h"

namespace pw {
namespace batch {
template void batch_step_s<StoreBF16>(const BatchLaunch&, int, bool);
}  // namespace batch
}  // namespace pw
