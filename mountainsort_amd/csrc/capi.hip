// Library identity and error reporting of the C ABI (include/msamd.h).
#include <cstring>

#include "common.h"

namespace msamd {

namespace {
thread_local char g_last_error[1024] = {0};
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_last_error[0] = 0; }

}  // namespace msamd

extern "C" int msamd_version(void) { return 10000; }

extern "C" const char* msamd_last_error(void) { return msamd::g_last_error; }
