// egsgpu_common.h -- shared host-side helpers of the engine library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <string>

namespace egsgpu {

void set_error(const char *fmt, ...);
const std::string &get_error();

#define EGS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ::egsgpu::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__,  \
                                #call, hipGetErrorString(e_));               \
            return EGSGPU_E_HIP;                                             \
        }                                                                    \
    } while (0)

#define EGS_ARG(cond, msg)                                                   \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::egsgpu::set_error("invalid argument: %s", msg);                \
            return EGSGPU_E_ARG;                                             \
        }                                                                    \
    } while (0)

// Check that a gfx950 device is current; sets the error text otherwise.
int require_device();

}  // namespace egsgpu
