// Explicit instantiations of the kerma-mode history kernels in float;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_KERMA_KERNELS(EGSGPU_INSTANTIATE, float)
}  // namespace egsgpu
