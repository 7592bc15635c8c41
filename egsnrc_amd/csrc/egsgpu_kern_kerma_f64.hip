// Explicit instantiations of the kerma-mode history kernels in double;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_KERMA_KERNELS(EGSGPU_INSTANTIATE, double)
}  // namespace egsgpu
