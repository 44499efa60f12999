// Handle-level generalized-reduction runtime for kmeans: the device-resident
// form of RegularRuntime (kmeans/regular/regular_runtime.h:11-106) running
// the kmeans functors (map1/reduce1, kmeans/regular.h:75-136) plus the Lloyd
// update the reference leaves out (km.cu runs one map pass and
// RegularGetOutput is empty, regular_runtime.cu:288-291).
//
// The reference splits units over MPI ranks and over CPU threads and GPUs
// with a mutex-guarded chunk scheduler, and its GPU kernel reads the input
// from mapped host memory over PCIe.  Here the points stay resident in HBM,
// the centres start as the first K points, and each iteration is one
// assignment+accumulation kernel and one update kernel on the handle's
// stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "pat_common.hpp"

#include "kmeans_dev.hpp"

struct pat_kmeans {
    pat::KmDev d{};
    int mode = PAT_KM_MFMA;
    bool own_x = false;
    float *xbuf = nullptr;
    void *rowbuf = nullptr;  // BF16 rows
    unsigned long long *maxbits = nullptr;
    unsigned long long *snap = nullptr;  // last pass's reduction object (K*D + 2K)
    hipStream_t stream = nullptr;
    float maxabs = 0.0f;
    int iters_done = 0;
    std::vector<unsigned long long> rescored_hist;
    std::map<int, hipGraphExec_t> graphs;

    void drop_graphs()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }

    // One Lloyd iteration: assignment + accumulation, snapshot of the
    // reduction object, centre update (which also clears the object).
    int enqueue_iter(hipStream_t s)
    {
        if (mode == PAT_KM_MFMA) PAT_TRY(pat::kmeans_assign_mfma(d, s));
        else PAT_TRY(pat::kmeans_assign_plain(d, s));
        const int64_t KD = (int64_t)d.K * d.D;
        PAT_HIP(hipMemcpyAsync(snap, d.gsum, 8 * KD, hipMemcpyDeviceToDevice, s));
        PAT_HIP(hipMemcpyAsync(snap + KD, d.gcnt, 8 * (size_t)d.K, hipMemcpyDeviceToDevice, s));
        PAT_HIP(hipMemcpyAsync(snap + KD + d.K, d.gdsum, 8 * (size_t)d.K, hipMemcpyDeviceToDevice, s));
        PAT_TRY(pat::kmeans_finalize(d, 1, s));
        return PAT_OK;
    }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (void *p : {(void *)xbuf, (void *)d.c, (void *)d.assign, (void *)d.mindist, (void *)d.gsum, (void *)d.gcnt,
                        (void *)d.gdsum, (void *)d.csnap, rowbuf, (void *)d.cnorm, (void *)d.rescored,
                        (void *)maxbits, (void *)snap, (void *)d.bidx, (void *)d.bcnt, (void *)d.boff,
                        (void *)d.bcur})
            if (p) (void)pat::dev_free(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int amalloc(void **p, size_t b)
{
    if (b == 0) b = 16;
    if (pat::dev_malloc(p, b) != hipSuccess) {
        (void)hipGetLastError();
        pat::set_error("kmeans: hipMalloc of %zu bytes failed", b);
        return PAT_ERR_NOMEM;
    }
    return PAT_OK;
}

// Terms of the MFMA path (1 or 3) chosen per (log2 N, D, log2 K) class for
// the process; both give the same assignments bit for bit, so the choice
// only changes speed.
std::map<std::tuple<int, int, int>, int> g_terms_cache;
std::mutex g_terms_mutex;

// When the three-term rows fit in LDS the one-term form fits too: it runs
// fewer than half the MFMAs and fragment reads but has a wider margin, so
// more points may need rescoring -- which wins depends on the data.  Times
// one assignment pass of each (interleaved rounds, best of each) on the
// current centres, keeps the faster, and clears what the passes
// accumulated.
int tune_terms(pat_kmeans *h)
{
    auto &D = h->d;
    const auto key = std::make_tuple(63 - __builtin_clzll((unsigned long long)D.N | 1), D.D,
                                     31 - __builtin_clz((unsigned)D.K | 1));
    int nt = 0;
    {
        std::lock_guard<std::mutex> g(g_terms_mutex);
        auto it = g_terms_cache.find(key);
        if (it != g_terms_cache.end()) nt = it->second;
    }
    if (!nt) {
        pat::EventPair ev;
        PAT_HIP(ev.create());
        hipEvent_t e0 = ev.a, e1 = ev.b;
        const int cand[2] = {3, 1};
        float best[2] = {1e30f, 1e30f};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 2; c++) {
                D.NT = cand[c];
                PAT_TRY(pat::kmeans_finalize(D, 0, h->stream));
                PAT_HIP(hipEventRecord(e0, h->stream));
                PAT_TRY(pat::kmeans_assign_mfma(D, h->stream));
                PAT_HIP(hipEventRecord(e1, h->stream));
                PAT_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                PAT_HIP(hipEventElapsedTime(&ms, e0, e1));
                best[c] = std::min(best[c], ms);
            }
        nt = best[1] < best[0] ? 1 : 3;
        if (pat::tune_log())
            fprintf(stderr, "[tune kmeans N=%lld D=%d K=%d] three terms %.3f ms, one term %.3f ms -> %d\n",
                    (long long)D.N, D.D, D.K, best[0], best[1], nt);
        std::lock_guard<std::mutex> g(g_terms_mutex);
        g_terms_cache[key] = nt;
    }
    D.NT = nt;
    PAT_TRY(pat::kmeans_zero(D.gsum, (int64_t)D.K * D.D, h->stream));
    PAT_TRY(pat::kmeans_zero(D.gcnt, D.K, h->stream));
    PAT_TRY(pat::kmeans_zero(D.gdsum, D.K, h->stream));
    PAT_TRY(pat::kmeans_zero(D.rescored, 1, h->stream));
    PAT_TRY(pat::kmeans_finalize(D, 0, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    return PAT_OK;
}

int setup(pat_kmeans *h, const float *x, int64_t n, int d, int k, int mode, int x_on_device)
{
    PAT_TRY(pat::kmeans_prepare_kernels());
    PAT_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->mode = mode;
    auto &D = h->d;
    D.N = n;
    D.D = d;
    D.K = k;
    D.K_pad = (k + 31) / 32 * 32;
    D.KS = pat::kmeans_ks_for(d);
    D.gc = pat::kmeans_gc_for(k, d) ? 1 : 0;
    D.NT = D.gc ? 3 : std::max(1, pat::kmeans_nt_for(k, d));  // rows are written even when only the plain path runs
    int dev = 0;
    PAT_HIP(hipGetDevice(&dev));
    PAT_HIP(hipDeviceGetAttribute(&D.cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (x_on_device) {
        D.x = x;
    } else {
        PAT_TRY(amalloc((void **)&h->xbuf, sizeof(float) * (size_t)n * d));
        PAT_HIP(hipMemcpyAsync(h->xbuf, x, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice, h->stream));
        D.x = h->xbuf;
        h->own_x = true;
    }
    PAT_TRY(amalloc((void **)&D.c, sizeof(float) * (size_t)k * d));
    PAT_TRY(amalloc((void **)&D.assign, sizeof(int32_t) * (size_t)n));
    PAT_TRY(amalloc((void **)&D.mindist, sizeof(float) * (size_t)n));
    PAT_TRY(amalloc((void **)&D.gsum, 8 * (size_t)k * d));
    PAT_TRY(amalloc((void **)&D.gcnt, 8 * (size_t)k));
    PAT_TRY(amalloc((void **)&D.gdsum, 8 * (size_t)k));
    PAT_TRY(amalloc((void **)&D.csnap, 8 * (size_t)k));
    PAT_TRY(amalloc(&h->rowbuf, 2 * (size_t)D.K_pad * 16 * (D.NT * D.KS + 1)));
    D.cb = h->rowbuf;
    PAT_TRY(amalloc((void **)&D.cnorm, sizeof(float) * (size_t)k));
    // a reduction object too large for LDS is summed bucket by bucket
    D.bucketed = !pat::kmeans_object_in_lds(k, d) && n < (1LL << 31);
    if (D.bucketed) {
        PAT_TRY(amalloc((void **)&D.bidx, sizeof(int) * (size_t)n));
        PAT_TRY(amalloc((void **)&D.bcnt, sizeof(int) * (size_t)k));
        PAT_TRY(amalloc((void **)&D.boff, sizeof(int) * ((size_t)k + 1)));
        PAT_TRY(amalloc((void **)&D.bcur, sizeof(int) * (size_t)k));
    }
    PAT_TRY(amalloc((void **)&D.rescored, 8));
    PAT_TRY(amalloc((void **)&h->maxbits, 8));
    PAT_TRY(amalloc((void **)&h->snap, 8 * ((size_t)k * d + 2 * (size_t)k)));
    // centres start as the first k points
    PAT_HIP(hipMemcpyAsync(D.c, D.x, sizeof(float) * (size_t)k * d, hipMemcpyDeviceToDevice, h->stream));
    // fixed-point scales from max |x|
    PAT_TRY(pat::kmeans_maxabs(D.x, n * d, (unsigned int *)h->maxbits, h->stream));
    unsigned long long mb = 0;
    PAT_HIP(hipMemcpyAsync(&mb, h->maxbits, 8, hipMemcpyDeviceToHost, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    unsigned int bits = (unsigned int)mb;
    memcpy(&h->maxabs, &bits, 4);
    D.shift = pat::kmeans_fixed_shift(n, h->maxabs);
    D.dshift = pat::kmeans_fixed_shift(n, (float)(2.0 * (double)h->maxabs * std::sqrt((double)d)));
    // the kernels scale by 2^shift in FP32
    PAT_REQUIRE(std::isfinite(h->maxabs) && D.shift >= -120 && D.dshift >= -120, PAT_ERR_INVALID,
                "kmeans: point coordinates out of range for the fixed-point reduction object (max |x| = %g)",
                (double)h->maxabs);
    PAT_TRY(pat::kmeans_zero(D.gsum, (int64_t)k * d, h->stream));
    PAT_TRY(pat::kmeans_zero(D.gcnt, k, h->stream));
    PAT_TRY(pat::kmeans_zero(D.gdsum, k, h->stream));
    PAT_TRY(pat::kmeans_zero(D.rescored, 1, h->stream));
    PAT_TRY(pat::kmeans_finalize(D, 0, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    // timed only where a pass is long enough to pay for six trial passes
    if (mode == PAT_KM_MFMA && !D.gc && D.NT == 3 && n >= (1 << 20) && pat::tuning_enabled()) PAT_TRY(tune_terms(h));
    return PAT_OK;
}

}  // namespace

extern "C" {

int pat_kmeans_fixed_shift(int64_t n, float maxabs, int *shift)
{
    PAT_REQUIRE(shift && n > 0, PAT_ERR_INVALID, "pat_kmeans_fixed_shift: bad arguments");
    *shift = pat::kmeans_fixed_shift(n, maxabs);
    return PAT_OK;
}

int pat_kmeans_create(const float *x, int64_t n, int d, int k, int mode, int x_on_device, pat_kmeans **out)
{
    PAT_REQUIRE(out && x, PAT_ERR_INVALID, "pat_kmeans_create: null argument");
    PAT_TRY(pat::kmeans_prepare_kernels());
    PAT_REQUIRE(n > 0 && n < (1LL << 40), PAT_ERR_INVALID, "pat_kmeans_create: bad point count");
    PAT_REQUIRE(d > 0 && d <= 128, PAT_ERR_UNSUPPORTED, "pat_kmeans_create: D = %d outside 1..128", d);
    PAT_REQUIRE(k > 0 && k <= n && k <= 65536, PAT_ERR_INVALID, "pat_kmeans_create: K = %d outside 1..min(N, 65536)",
                k);
    PAT_REQUIRE(mode == PAT_KM_PLAIN || mode == PAT_KM_MFMA, PAT_ERR_UNSUPPORTED, "pat_kmeans_create: unknown mode %d",
                mode);
    if (mode == PAT_KM_MFMA)
        PAT_REQUIRE(pat::kmeans_mfma_fits(k, d), PAT_ERR_UNSUPPORTED,
                    "pat_kmeans_create: K = %d, D = %d exceed the MFMA path's LDS staging (use PAT_KM_PLAIN)", k, d);
    auto *h = new pat_kmeans();
    int rc = setup(h, x, n, d, k, mode, x_on_device);
    if (rc != PAT_OK) {
        h->release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

int pat_kmeans_set_mode(pat_kmeans *h, int mode)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_set_mode: null handle");
    PAT_REQUIRE(mode == PAT_KM_PLAIN || mode == PAT_KM_MFMA, PAT_ERR_UNSUPPORTED, "pat_kmeans_set_mode: unknown mode");
    if (mode == PAT_KM_MFMA)
        PAT_REQUIRE(pat::kmeans_mfma_fits(h->d.K, h->d.D), PAT_ERR_UNSUPPORTED,
                    "pat_kmeans_set_mode: shape too large for the MFMA path");
    h->mode = mode;
    h->drop_graphs();
    return PAT_OK;
}

int pat_kmeans_set_terms(pat_kmeans *h, int terms)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_set_terms: null handle");
    PAT_REQUIRE(terms == 1 || terms == 3, PAT_ERR_INVALID, "pat_kmeans_set_terms: terms %d not 1 or 3", terms);
    auto &D = h->d;
    const int fits = pat::kmeans_nt_for(D.K, D.D);  // 3: both forms fit, 1: only one term, 0: neither
    PAT_REQUIRE(!D.gc && (fits == 3 || (fits == 1 && terms == 1)), PAT_ERR_UNSUPPORTED,
                "pat_kmeans_set_terms: %d-term rows for K = %d, D = %d do not fit in LDS", terms, D.K, D.D);
    if (D.NT != terms) {
        D.NT = terms;
        PAT_TRY(pat::kmeans_finalize(D, 0, h->stream));
        PAT_HIP(hipStreamSynchronize(h->stream));
        h->drop_graphs();
    }
    return PAT_OK;
}

int pat_kmeans_get_terms(pat_kmeans *h, int *terms)
{
    PAT_REQUIRE(h && terms, PAT_ERR_INVALID, "pat_kmeans_get_terms: null argument");
    *terms = h->d.NT;
    return PAT_OK;
}

int pat_kmeans_set_centroids(pat_kmeans *h, const float *c)
{
    PAT_REQUIRE(h && c, PAT_ERR_INVALID, "pat_kmeans_set_centroids: null argument");
    PAT_HIP(hipMemcpyAsync(h->d.c, c, sizeof(float) * (size_t)h->d.K * h->d.D, hipMemcpyHostToDevice, h->stream));
    PAT_TRY(pat::kmeans_finalize(h->d, 0, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    return PAT_OK;
}

int pat_kmeans_run(pat_kmeans *h, int iters)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_run: null handle");
    PAT_REQUIRE(iters >= 0, PAT_ERR_INVALID, "pat_kmeans_run: negative iteration count");
    if (iters == 0) return PAT_OK;
    auto it = h->graphs.find(iters);
    if (it == h->graphs.end()) {
        hipGraph_t g;
        PAT_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int rc = PAT_OK;
        for (int i = 0; i < iters && rc == PAT_OK; i++) rc = h->enqueue_iter(h->stream);
        hipError_t ec = hipStreamEndCapture(h->stream, &g);
        if (rc != PAT_OK) return rc;
        PAT_HIP(ec);
        hipGraphExec_t ge;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PAT_HIP(ei);
        it = h->graphs.emplace(iters, ge).first;
    }
    PAT_HIP(hipGraphLaunch(it->second, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    h->iters_done += iters;
    return PAT_OK;
}

}  // extern "C"

namespace pat {

// pat_kmeans_run's iterations launched one by one on the handle's stream
// (no graph): for a handle that runs once, where instantiating the graph
// would cost more than it saves.
int kmeans_run_eager(pat_kmeans *h, int iters)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "kmeans_run_eager: null handle");
    PAT_REQUIRE(iters >= 0, PAT_ERR_INVALID, "kmeans_run_eager: negative iteration count");
    for (int i = 0; i < iters; i++) PAT_TRY(h->enqueue_iter(h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    h->iters_done += iters;
    return PAT_OK;
}

}  // namespace pat

extern "C" {

int pat_kmeans_run_timed(pat_kmeans *h, int iters, float *total_ms, float *assign_ms)
{
    PAT_REQUIRE(h && total_ms, PAT_ERR_INVALID, "pat_kmeans_run_timed: null argument");
    PAT_REQUIRE(iters > 0, PAT_ERR_INVALID, "pat_kmeans_run_timed: need at least one iteration");
    pat::EventVec ev;
    PAT_HIP(ev.create(2 * (size_t)iters + 2));
    int rc = PAT_OK;
    (void)hipEventRecord(ev[0], h->stream);
    for (int i = 0; i < iters && rc == PAT_OK; i++) {
        (void)hipEventRecord(ev[1 + 2 * i], h->stream);
        rc = h->mode == PAT_KM_MFMA ? pat::kmeans_assign_mfma(h->d, h->stream) : pat::kmeans_assign_plain(h->d, h->stream);
        (void)hipEventRecord(ev[2 + 2 * i], h->stream);
        if (rc != PAT_OK) break;
        const int64_t KD = (int64_t)h->d.K * h->d.D;
        if (hipMemcpyAsync(h->snap, h->d.gsum, 8 * KD, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(h->snap + KD, h->d.gcnt, 8 * (size_t)h->d.K, hipMemcpyDeviceToDevice, h->stream) !=
                hipSuccess ||
            hipMemcpyAsync(h->snap + KD + h->d.K, h->d.gdsum, 8 * (size_t)h->d.K, hipMemcpyDeviceToDevice,
                           h->stream) != hipSuccess)
            rc = PAT_ERR_HIP;
        if (rc == PAT_OK) rc = pat::kmeans_finalize(h->d, 1, h->stream);
    }
    (void)hipEventRecord(ev[2 * (size_t)iters + 1], h->stream);
    if (rc == PAT_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = PAT_ERR_HIP;
    if (rc == PAT_OK) {
        (void)hipEventElapsedTime(total_ms, ev[0], ev[2 * (size_t)iters + 1]);
        if (assign_ms)
            for (int i = 0; i < iters; i++) (void)hipEventElapsedTime(&assign_ms[i], ev[1 + 2 * i], ev[2 + 2 * i]);
        h->iters_done += iters;
    } else if (rc == PAT_ERR_HIP) {
        pat::set_error("pat_kmeans_run_timed: HIP call failed");
    }
    return rc;
}

int pat_kmeans_get(pat_kmeans *h, float *centroids, int32_t *assign, float *mindist)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_get: null handle");
    auto &D = h->d;
    if (centroids)
        PAT_HIP(hipMemcpyAsync(centroids, D.c, sizeof(float) * (size_t)D.K * D.D, hipMemcpyDeviceToHost, h->stream));
    if (assign) PAT_HIP(hipMemcpyAsync(assign, D.assign, sizeof(int32_t) * (size_t)D.N, hipMemcpyDeviceToHost, h->stream));
    if (mindist) PAT_HIP(hipMemcpyAsync(mindist, D.mindist, sizeof(float) * (size_t)D.N, hipMemcpyDeviceToHost, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    return PAT_OK;
}

int pat_kmeans_get_object(pat_kmeans *h, int64_t *sums, int64_t *counts, int64_t *dsums, int *shift, int *dshift)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_get_object: null handle");
    PAT_REQUIRE(h->iters_done > 0, PAT_ERR_STATE, "pat_kmeans_get_object: no pass has run");
    const int64_t KD = (int64_t)h->d.K * h->d.D;
    if (sums) PAT_HIP(hipMemcpyAsync(sums, h->snap, 8 * KD, hipMemcpyDeviceToHost, h->stream));
    if (counts) PAT_HIP(hipMemcpyAsync(counts, h->snap + KD, 8 * (size_t)h->d.K, hipMemcpyDeviceToHost, h->stream));
    if (dsums) PAT_HIP(hipMemcpyAsync(dsums, h->snap + KD + h->d.K, 8 * (size_t)h->d.K, hipMemcpyDeviceToHost, h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    if (shift) *shift = h->d.shift;
    if (dshift) *dshift = h->d.dshift;
    return PAT_OK;
}

int pat_kmeans_stats(pat_kmeans *h, int64_t *rescored_total, float *maxabs)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_kmeans_stats: null handle");
    if (rescored_total) {
        unsigned long long r = 0;
        PAT_HIP(hipMemcpyAsync(&r, h->d.rescored, 8, hipMemcpyDeviceToHost, h->stream));
        PAT_HIP(hipStreamSynchronize(h->stream));
        *rescored_total = (int64_t)r;
    }
    if (maxabs) *maxabs = h->maxabs;
    return PAT_OK;
}

int pat_kmeans_destroy(pat_kmeans *h)
{
    if (!h) return PAT_OK;
    h->release();
    delete h;
    return PAT_OK;
}

}  // extern "C"
