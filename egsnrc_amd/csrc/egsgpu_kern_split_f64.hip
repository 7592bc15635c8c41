// Explicit instantiations of the history kernels with photon splitting, radiative splitting or range rejection in double;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_SPLIT_KERNELS(EGSGPU_INSTANTIATE, double)
}  // namespace egsgpu
