// Explicit instantiations of the event-schedule source and photon kernels in float;
// the kernel lists and templates are in egsgpu_kernels.h.
#include "egsgpu_kernels.h"

namespace egsgpu {
EGSGPU_EVENT_KERNELS(EGSGPU_INSTANTIATE, float)
}  // namespace egsgpu
