// The calling thread's last error message (gz_last_error), shared by every translation unit of
// the library. Host-only, so the host C++ sources (BVH builder, PLY reader, CPU traversal and
// shade, Sobol table) build on their own, for instance with the sanitizers.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gonzales_amd.h"

namespace gz {

static thread_local std::string t_error;

void set_error(const char *fmt, ...) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        t_error = buf;
}

void clear_error() { t_error.clear(); }

// Allocation-failure injection for the robustness tests. With GZ_FAIL_ALLOC_AFTER=k in the
// environment, the k-th allocation (counting from 0) after an entry point armed the countdown
// fails as if the allocator had, on the calling thread; unset, nothing fails.
static thread_local long t_faultCountdown = -1;

void alloc_fault_arm() {
        const char *e = std::getenv("GZ_FAIL_ALLOC_AFTER");
        t_faultCountdown = (e && *e) ? std::strtol(e, nullptr, 10) : -1;
}

bool alloc_fault(const char *what) {
        if (t_faultCountdown < 0) return false;
        if (t_faultCountdown-- > 0) return false;
        set_error("injected allocation failure (GZ_FAIL_ALLOC_AFTER): %s", what);
        return true;
}

}  // namespace gz

extern "C" const char *gz_last_error(void) { return gz::t_error.c_str(); }
