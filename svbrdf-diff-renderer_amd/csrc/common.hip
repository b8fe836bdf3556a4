// common.hip -- error state shared by the translation units of libsvbrdf_hip.
#include "common.hpp"

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/svbrdf_hip.h"

namespace svbrdf_internal {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SVBRDF_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
    return SVBRDF_OK;
}

const char *last_error() { return g_last_error.c_str(); }

}  // namespace svbrdf_internal
