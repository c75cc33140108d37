// Error reporting and version of the libwotk C ABI.
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace {
thread_local char g_last_error[1024] = "";
}

namespace wotk {
int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace wotk

extern "C" const char *wotk_last_error(void) { return g_last_error; }
extern "C" int wotk_abi_version(void) { return 1; }
