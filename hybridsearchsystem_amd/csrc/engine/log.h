// log.h -- stage logging for the host engine: errors go to stderr and become the thread's
// frk_last_error() text; NPUR_LOG_LEVEL (ERROR/WARN/INFO/DEBUG) selects what else is printed, as the
// reference's logger stub does (platform/src/utils/logger.h).
#pragma once

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

namespace frk {
void SetError(const char* fmt, ...);
}

namespace NpuRetrieval {
inline int LogLevel() {
    static const int level = [] {
        const char* v = std::getenv("NPUR_LOG_LEVEL");
        if (!v) return 1;  // WARN and above
        if (!std::strcmp(v, "ERROR")) return 0;
        if (!std::strcmp(v, "WARN")) return 1;
        if (!std::strcmp(v, "INFO")) return 2;
        return 3;
    }();
    return level;
}
}  // namespace NpuRetrieval

#define FRK_LOG_ERROR(msg)                                               \
    do {                                                                 \
        std::ostringstream frk_os_;                                      \
        frk_os_ << msg;                                                  \
        ::frk::SetError("%s", frk_os_.str().c_str());                    \
        std::cerr << "[frk][ERROR] " << frk_os_.str() << std::endl;      \
    } while (0)

#define FRK_LOG_AT(lvl, tag, msg)                                                                   \
    do {                                                                                            \
        if (::NpuRetrieval::LogLevel() >= (lvl)) std::cerr << "[frk][" tag "] " << msg << std::endl; \
    } while (0)

#define FRK_LOG_WARN(msg) FRK_LOG_AT(1, "WARN", msg)
#define FRK_LOG_INFO(msg) FRK_LOG_AT(2, "INFO", msg)
#define FRK_LOG_DEBUG(msg) FRK_LOG_AT(3, "DEBUG", msg)
