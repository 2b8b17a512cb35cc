// frk_common.cpp -- error reporting shared by the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include <atomic>

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

int CurrentDeviceCUs(int* cus) {
    constexpr int kMaxDevices = 64;
    static std::atomic<int> cached[kMaxDevices] = {};  // 0 = not queried yet
    int dev = 0;
    FRK_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        SetError("device %d out of range", dev);
        return FRK_ERR_UNSUPPORTED;
    }
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        FRK_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        if (n <= 0) n = 1;
        cached[dev].store(n, std::memory_order_relaxed);  // every racing thread stores the same value
    }
    *cus = n;
    return FRK_OK;
}
}  // namespace frk

extern "C" const char* frk_last_error(void) { return frk::g_last_error; }
