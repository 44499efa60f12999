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

static thread_local bool t_tuning = true;

bool tuning_enabled() { return t_tuning && !getenv("PAT_NO_TUNE"); }
bool tune_log() { return getenv("PAT_TUNE_LOG") != nullptr; }

void set_thread_tuning(bool on) { t_tuning = on; }

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

int pat_dev_malloc(void **ptr, int64_t bytes)
{
    PAT_REQUIRE(ptr && bytes >= 0, PAT_ERR_INVALID, "pat_dev_malloc: bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return PAT_OK;
    hipError_t e = hipMalloc(ptr, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pat::set_error("pat_dev_malloc: %lld bytes: %s", (long long)bytes, hipGetErrorString(e));
        return PAT_ERR_NOMEM;
    }
    return PAT_OK;
}

int pat_dev_free(void *ptr)
{
    if (ptr) PAT_HIP(hipFree(ptr));
    return PAT_OK;
}

int pat_memcpy_h2d(void *dst, const void *src, int64_t bytes, void *stream)
{
    PAT_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), PAT_ERR_INVALID, "pat_memcpy_h2d: bad arguments");
    if (bytes == 0) return PAT_OK;
    PAT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    PAT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PAT_OK;
}

int pat_memcpy_d2h(void *dst, const void *src, int64_t bytes, void *stream)
{
    PAT_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), PAT_ERR_INVALID, "pat_memcpy_d2h: bad arguments");
    if (bytes == 0) return PAT_OK;
    PAT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PAT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PAT_OK;
}

int pat_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *stream)
{
    PAT_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), PAT_ERR_INVALID, "pat_memcpy_d2d: bad arguments");
    if (bytes == 0) return PAT_OK;
    PAT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return PAT_OK;
}

int pat_memset_dev(void *ptr, int value, int64_t bytes, void *stream)
{
    PAT_REQUIRE(bytes >= 0 && (bytes == 0 || ptr), PAT_ERR_INVALID, "pat_memset_dev: bad arguments");
    if (bytes == 0) return PAT_OK;
    PAT_HIP(hipMemsetAsync(ptr, value, (size_t)bytes, (hipStream_t)stream));
    return PAT_OK;
}

int pat_stream_create(void **stream)
{
    PAT_REQUIRE(stream, PAT_ERR_INVALID, "pat_stream_create: null argument");
    hipStream_t s;
    PAT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return PAT_OK;
}

int pat_stream_destroy(void *stream)
{
    if (stream) PAT_HIP(hipStreamDestroy((hipStream_t)stream));
    return PAT_OK;
}

int pat_stream_sync(void *stream)
{
    PAT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PAT_OK;
}

int pat_event_create(void **event)
{
    PAT_REQUIRE(event, PAT_ERR_INVALID, "pat_event_create: null argument");
    hipEvent_t e;
    PAT_HIP(hipEventCreate(&e));
    *event = e;
    return PAT_OK;
}

int pat_event_destroy(void *event)
{
    if (event) PAT_HIP(hipEventDestroy((hipEvent_t)event));
    return PAT_OK;
}

int pat_event_record(void *event, void *stream)
{
    PAT_REQUIRE(event, PAT_ERR_INVALID, "pat_event_record: null event");
    PAT_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return PAT_OK;
}

int pat_event_elapsed_ms(void *start, void *end, float *ms)
{
    PAT_REQUIRE(start && end && ms, PAT_ERR_INVALID, "pat_event_elapsed_ms: null argument");
    PAT_HIP(hipEventSynchronize((hipEvent_t)end));
    PAT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return PAT_OK;
}

int pat_device_query(int *num_cus, int64_t *total_bytes, int64_t *free_bytes)
{
    hipDeviceProp_t p;
    PAT_HIP(hipGetDeviceProperties(&p, pat::g_device));
    if (num_cus) *num_cus = p.multiProcessorCount;
    size_t fr = 0, tot = 0;
    PAT_HIP(hipMemGetInfo(&fr, &tot));
    if (total_bytes) *total_bytes = (int64_t)tot;
    if (free_bytes) *free_bytes = (int64_t)fr;
    return PAT_OK;
}

int pat_info(char *buf, int64_t buf_len)
{
    PAT_REQUIRE(buf && buf_len > 0, PAT_ERR_INVALID, "pat_info: null buffer");
    char name[800] = "none";
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, pat::g_device) == hipSuccess)
            snprintf(name, sizeof(name), "%s (%s) cus=%d lds_per_block=%zu clock_khz=%d", p.name, p.gcnArchName,
                     p.multiProcessorCount, (size_t)p.sharedMemPerBlock, p.clockRate);
    }
    snprintf(buf, (size_t)buf_len, "libpattern 0.1 target=gfx950 device=%s", name);
    return PAT_OK;
}

}  // extern "C"
