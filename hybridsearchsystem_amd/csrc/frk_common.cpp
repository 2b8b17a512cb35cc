// frk_common.cpp -- error reporting shared by the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "frk_common.h"

namespace frk {
namespace {
thread_local char g_last_error[512] = "";
}

void SetError(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}
}  // namespace frk

extern "C" const char* frk_last_error(void) { return frk::g_last_error; }
