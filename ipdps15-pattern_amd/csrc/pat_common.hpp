// Shared helpers of libpattern: error state, status macros, host threading.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "pattern.h"

namespace pat {

// Records the calling thread's last error message (returned by pat_last_error).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// Number of host threads for parallel host loops (generators).
int host_threads();

// Runs body(begin, end) over [0, n) split into contiguous chunks, one per
// thread.  Each chunk's result must depend only on the indices it covers so
// that the output does not depend on the thread count.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body);

// Library's current device (set by pat_init, default 0).
int current_device();

}  // namespace pat

#define PAT_HIP(call)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess) {                                                                            \
            pat::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));     \
            return PAT_ERR_HIP;                                                                            \
        }                                                                                                  \
    } while (0)

#define PAT_REQUIRE(cond, code, ...)                                                                       \
    do {                                                                                                   \
        if (!(cond)) {                                                                                     \
            pat::set_error(__VA_ARGS__);                                                                   \
            return code;                                                                                   \
        }                                                                                                  \
    } while (0)

#define PAT_TRY(expr)                                                                                      \
    do {                                                                                                   \
        int r_ = (expr);                                                                                   \
        if (r_ != PAT_OK) return r_;                                                                       \
    } while (0)

// Launch-error check after a kernel launch.
#define PAT_LAUNCH_CHECK() PAT_HIP(hipGetLastError())
