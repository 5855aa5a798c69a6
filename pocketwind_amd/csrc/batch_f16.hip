// batch_f16.hip -- batched step kernels with f16 population storage.
#include "batch_step.h"

namespace pw {
namespace batch {
template void batch_step_s<StoreF16>(const BatchLaunch&, int, bool);
}  // namespace batch
}  // namespace pw
