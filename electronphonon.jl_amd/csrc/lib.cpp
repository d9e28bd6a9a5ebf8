// Host-side C-ABI entries: error reporting, version, backend primitives.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/ephip.h"

namespace ephip {

static thread_local char g_last_error[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

}  // namespace ephip

extern "C" const char *ephip_last_error(void) { return ephip::g_last_error; }

extern "C" int ephip_version(void) { return 100; }

extern "C" int ephip_free_bytes(int64_t *free_bytes, int64_t *total_bytes) {
    size_t f = 0, t = 0;
    hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) {
        ephip::set_error("hipMemGetInfo failed: %s", hipGetErrorString(e));
        return EPHIP_ERR_HIP;
    }
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return EPHIP_OK;
}

extern "C" int ephip_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        ephip::set_error("stream synchronize failed: %s", hipGetErrorString(e));
        return EPHIP_ERR_HIP;
    }
    return EPHIP_OK;
}
