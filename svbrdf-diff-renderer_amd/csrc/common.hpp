// common.hpp -- error state shared by the translation units of libsvbrdf_hip.
#pragma once

#include <hip/hip_runtime.h>

namespace svbrdf_internal {

// Records a printf-style message for svbrdf_last_error() (per host thread)
// and returns `code`.
__attribute__((visibility("hidden"), format(printf, 2, 3))) int fail(int code, const char *fmt, ...);

// SVBRDF_EHIP with the launch error when the last launch failed, else 0.
__attribute__((visibility("hidden"))) int check_launch(const char *what);

__attribute__((visibility("hidden"))) const char *last_error();

}  // namespace svbrdf_internal
