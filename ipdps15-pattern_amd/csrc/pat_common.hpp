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

// Device memory for every buffer the library owns.  By default these are
// hipMalloc/hipFree; on a thread that has an arena bound (the serving loop's
// workers) allocations are carved from that arena instead, so jobs neither
// allocate from nor free to the runtime (hipFree waits for the whole
// device).  dev_free finds the owner by address, so it may run on any thread.
struct Arena;
hipError_t dev_malloc(void **p, size_t bytes);
template <typename T>
inline hipError_t dev_malloc(T **p, size_t bytes)
{
    return dev_malloc((void **)p, bytes);
}
hipError_t dev_free(void *p);
Arena *arena_create(size_t bytes);  // one hipMalloc of 'bytes'; nullptr on failure (error set)
void arena_destroy(Arena *a);       // the caller guarantees nothing in it is still in use
void arena_bind(Arena *a);          // this thread's allocations come from a (nullptr: the runtime)
size_t arena_in_use(const Arena *a);
size_t arena_largest_free(const Arena *a);

// Whether kernel autotuning may run on this thread (default yes; off with
// PAT_NO_TUNE in the environment, or per thread, as the serving loop's
// workers do: timings taken beside other jobs would be noise).
bool tuning_enabled();
// PAT_TUNE_LOG in the environment: the autotuners print their candidates'
// times to stderr.
bool tune_log();
void set_thread_tuning(bool on);

// Two HIP events destroyed on every path out of the scope that owns them.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    hipError_t create()
    {
        const hipError_t e = hipEventCreate(&a);
        return e != hipSuccess ? e : hipEventCreate(&b);
    }
};

// HIP events destroyed on every path out of the scope that owns them.
struct EventVec {
    std::vector<hipEvent_t> ev;
    EventVec() = default;
    EventVec(const EventVec &) = delete;
    EventVec &operator=(const EventVec &) = delete;
    ~EventVec()
    {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    hipError_t create(size_t n)
    {
        ev.assign(n, nullptr);
        for (hipEvent_t &e : ev) {
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                e = nullptr;
                return r;
            }
        }
        return hipSuccess;
    }
    hipEvent_t &operator[](size_t i) { return ev[i]; }
    size_t size() const { return ev.size(); }
};

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
