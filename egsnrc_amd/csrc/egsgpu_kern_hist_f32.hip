// Explicit instantiations of the history-schedule kernels (plain and HOWFARLESS) in float;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_HISTORY_KERNELS(EGSGPU_INSTANTIATE, float)
}  // namespace egsgpu
