// Runtime environment of libpattern: error state, device selection, host
// threading helpers.  RuntimeInit/RuntimeFinalize of the reference
// (kmeans/include/runtime.cpp:4-13) initialise MPI; with one GPU holding the
// whole domain they select the device instead.
#include <cstring>
#include <mutex>

#include "pat_common.hpp"

namespace pat {

static thread_local std::string g_error;
static int g_device = 0;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

void clear_error() { g_error.clear(); }

int host_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (const char *s = getenv("OMP_NUM_THREADS")) {
        int v = atoi(s);
        if (v > 0) n = (unsigned)v;
    }
    if (n == 0) n = 1;
    if (n > 64) n = 64;
    return (int)n;
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body)
{
    if (n <= 0) return;
    int nt = host_threads();
    if (n < 4096 || nt == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        int64_t b = t * per, e = b + per < n ? b + per : n;
        if (b >= e) break;
        th.emplace_back([&body, b, e] { body(b, e); });
    }
    for (auto &t : th) t.join();
}

int current_device() { return g_device; }

}  // namespace pat

extern "C" {

const char *pat_last_error(void) { return pat::g_error.c_str(); }

int pat_init(int device)
{
    int n = 0;
    PAT_HIP(hipGetDeviceCount(&n));
    PAT_REQUIRE(device >= 0 && device < n, PAT_ERR_INVALID, "pat_init: device %d out of range (%d devices)", device, n);
    PAT_HIP(hipSetDevice(device));
    pat::g_device = device;
    return PAT_OK;
}

int pat_finalize(void)
{
    PAT_HIP(hipDeviceSynchronize());
    return PAT_OK;
}

int pat_device_sync(void)
{
    PAT_HIP(hipDeviceSynchronize());
    return PAT_OK;
}

int pat_info(char *buf, int64_t buf_len)
{
    PAT_REQUIRE(buf && buf_len > 0, PAT_ERR_INVALID, "pat_info: null buffer");
    char name[600] = "none";
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, pat::g_device) == hipSuccess) snprintf(name, sizeof(name), "%s (%s)", p.name, p.gcnArchName);
    }
    snprintf(buf, (size_t)buf_len, "libpattern 0.1 target=gfx950 device=%s", name);
    return PAT_OK;
}

}  // extern "C"
