// geneig_err.hpp -- error reporting shared by the host-side sources: the
// message is kept per thread for geneig_last_error() and echoed to stderr.
#pragma once

#include <hip/hip_runtime.h>

namespace geneig_host {

int fail(int code, const char* fmt, ...);

}  // namespace geneig_host

#define GENEIG_HIP_TRY(expr)                                                                          \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return geneig_host::fail(GENEIG_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));   \
    } while (0)
