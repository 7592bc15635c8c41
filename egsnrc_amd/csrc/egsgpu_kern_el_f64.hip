// Explicit instantiations of the event-schedule charged-particle kernels (2 waves per SIMD) in double;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_ELECTRON_KERNELS(EGSGPU_INSTANTIATE, double)
}  // namespace egsgpu
