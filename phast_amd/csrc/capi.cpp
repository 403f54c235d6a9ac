// Library-wide C entry points: version string and the thread-local error message.
#include <cstdarg>
#include <cstdio>

#include "../../include/phast_mi.h"

namespace phast {

static thread_local char g_error[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

}  // namespace phast

extern "C" const char *phast_version(void) { return "phast-mi355x 0.1.0 (gfx950, fp64)"; }

extern "C" const char *phast_last_error(void) { return phast::g_error; }
