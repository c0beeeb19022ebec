// Error reporting and version of the gmaster_amd C-ABI (host code).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace gm {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GM_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
    return GM_OK;
}

}  // namespace gm

extern "C" const char* gm_last_error(void) { return gm::g_last_error; }

extern "C" int gm_abi_version(void) { return 2; }
