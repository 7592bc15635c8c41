// Explicit instantiations of the history kernels with photon splitting, radiative splitting or range rejection in float;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_SPLIT_KERNELS(EGSGPU_INSTANTIATE, float)
}  // namespace egsgpu
