// Handle-level stencil runtime: the device-resident form of StencilRuntime
// (kmeans/lib/stencil_runtime.h:13-93, heat_3D/lib/stencil_runtime.h:16).
//
// The reference splits the per-node grid into a CPU slab and GPU slabs,
// exchanges halos through MPI and mapped host buffers every iteration and
// never swaps its double buffer (kmeans/lib/grid.h:61-65).  With one GPU
// holding the whole domain, the runtime keeps two halo-inclusive device grids,
// ping-pongs them each step, and replays repeated step counts from a captured
// hipGraph so that small grids are not bound by launch latency.
#include <cstring>
#include <map>

#include "pat_common.hpp"

namespace pat {
int heat3d_sweep(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int variant, int bx,
                 int by, int ry, int zc, hipStream_t stream);
void heat3d_default_tile(int64_t nx, int64_t ny, int64_t nz, int *bx, int *by, int *ry, int *zc);
int heat3d_shapes(const int **triples);
int heat3d_tuned_tile(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, hipStream_t st,
                      int *bx, int *by, int *ry, int *zc);
}  // namespace pat

struct pat_stencil {
    int ndims = 3;
    int64_t n[3] = {0, 0, 0};  // interior x, y, z
    int width = 1;
    int unit = 4;
    int idx = PAT_STENCIL_HEAT;
    std::vector<unsigned char> param;
    int variant = PAT_STENCIL_AUTO;
    int bx = 0, by = 0, ry = 0, zc = 0;
    float *buf[2] = {nullptr, nullptr};
    int cur = 0;  // buffer holding the current field
    bool loaded = false;
    hipStream_t stream = nullptr;
    // captured step sequences keyed by (nsteps, starting buffer)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;

    int64_t cells() const { return (n[0] + 2 * width) * (n[1] + 2 * width) * (n[2] + 2 * width); }

    void drop_graphs()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }

    float alpha() const
    {
        float a = 0.1f;
        if (param.size() >= sizeof(float)) memcpy(&a, param.data(), sizeof(float));
        return a;
    }

    int resolve_tile(int *tbx, int *tby, int *tr, int *tzc)
    {
        int dbx, dby, dry, dzc;
        if (variant == PAT_STENCIL_AUTO && bx == 0 && by == 0 && zc == 0 && loaded)
            PAT_TRY(pat::heat3d_tuned_tile(buf[cur], buf[1 - cur], n[0], n[1], n[2], alpha(), stream, &dbx, &dby, &dry,
                                           &dzc));
        else
            pat::heat3d_default_tile(n[0], n[1], n[2], &dbx, &dby, &dry, &dzc);
        *tbx = bx > 0 ? bx : dbx;
        *tby = by > 0 ? by : dby;
        *tr = ry > 0 ? ry : dry;
        *tzc = zc > 0 ? zc : dzc;
        return PAT_OK;
    }

    int enqueue(int nsteps, int start)
    {
        int tbx, tby, tr, tzc;
        PAT_TRY(resolve_tile(&tbx, &tby, &tr, &tzc));
        int v = variant == PAT_STENCIL_GENERIC ? PAT_STENCIL_GENERIC : PAT_STENCIL_LDS25D;
        int s = start;
        for (int k = 0; k < nsteps; k++) {
            PAT_TRY(pat::heat3d_sweep(buf[s], buf[1 - s], n[0], n[1], n[2], alpha(), v, tbx, tby, tr, tzc, stream));
            s = 1 - s;
        }
        return PAT_OK;
    }
};

extern "C" {

int pat_stencil_create(int ndims, const int64_t *interior, int width, int unit_size, const void *param,
                       int64_t param_size, pat_stencil **out)
{
    PAT_REQUIRE(out && interior, PAT_ERR_INVALID, "pat_stencil_create: null argument");
    PAT_REQUIRE(ndims == 3, PAT_ERR_UNSUPPORTED, "pat_stencil_create: only 3-D grids are supported (got %d)", ndims);
    PAT_REQUIRE(width == 1, PAT_ERR_UNSUPPORTED, "pat_stencil_create: only halo width 1 is supported (got %d)", width);
    PAT_REQUIRE(unit_size == 4, PAT_ERR_UNSUPPORTED, "pat_stencil_create: only FP32 grids are supported (got %d B)",
                unit_size);
    PAT_REQUIRE(param_size >= 0 && (param_size == 0 || param), PAT_ERR_INVALID, "pat_stencil_create: bad parameter");
    for (int d = 0; d < 3; d++)
        PAT_REQUIRE(interior[d] > 0 && interior[d] < (1 << 20), PAT_ERR_INVALID, "pat_stencil_create: bad extent");
    auto *s = new pat_stencil();
    s->ndims = ndims;
    for (int d = 0; d < 3; d++) s->n[d] = interior[d];
    s->width = width;
    s->unit = unit_size;
    if (param_size > 0) s->param.assign((const unsigned char *)param, (const unsigned char *)param + param_size);
    const size_t bytes = sizeof(float) * (size_t)s->cells();
    for (int b = 0; b < 2; b++) {
        if (pat::dev_malloc(&s->buf[b], bytes) != hipSuccess) {
            if (s->buf[0]) (void)pat::dev_free(s->buf[0]);
            delete s;
            pat::set_error("pat_stencil_create: hipMalloc of %zu bytes failed", bytes);
            return PAT_ERR_NOMEM;
        }
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)pat::dev_free(s->buf[0]);
        (void)pat::dev_free(s->buf[1]);
        delete s;
        pat::set_error("pat_stencil_create: stream creation failed");
        return PAT_ERR_HIP;
    }
    *out = s;
    return PAT_OK;
}

int pat_stencil_set_idx(pat_stencil *s, int idx)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_set_idx: null handle");
    PAT_REQUIRE(idx == PAT_STENCIL_HEAT, PAT_ERR_UNSUPPORTED, "pat_stencil_set_idx: functor index %d not in the table",
                idx);
    s->idx = idx;
    return PAT_OK;
}

int pat_stencil_set_variant(pat_stencil *s, int variant, int bx, int by, int zc)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_set_variant: null handle");
    PAT_REQUIRE(variant >= PAT_STENCIL_AUTO && variant <= PAT_STENCIL_GENERIC, PAT_ERR_UNSUPPORTED,
                "pat_stencil_set_variant: unknown variant %d", variant);
    int tby = 0, tr = 0;
    if (by > 0) {
        const int *sh;
        int ns = pat::heat3d_shapes(&sh);
        int wbx = bx > 0 ? bx : 64;
        bool ok = false;
        for (int i = 0; i < ns; i++)
            if (sh[3 * i] == wbx && sh[3 * i + 1] * sh[3 * i + 2] == by) {
                tby = sh[3 * i + 1];
                tr = sh[3 * i + 2];
                ok = true;
                break;
            }
        PAT_REQUIRE(ok, PAT_ERR_UNSUPPORTED, "pat_stencil_set_variant: tile %dx%d not instantiated", wbx, by);
    }
    s->variant = variant;
    s->bx = bx;
    s->by = tby;
    s->ry = tr;
    s->zc = zc;
    s->drop_graphs();
    return PAT_OK;
}

int pat_stencil_upload(pat_stencil *s, const void *host)
{
    PAT_REQUIRE(s && host, PAT_ERR_INVALID, "pat_stencil_upload: null argument");
    const size_t bytes = sizeof(float) * (size_t)s->cells();
    // Both buffers get the same field, so both carry the Dirichlet halo.
    PAT_HIP(hipMemcpyAsync(s->buf[0], host, bytes, hipMemcpyHostToDevice, s->stream));
    PAT_HIP(hipMemcpyAsync(s->buf[1], host, bytes, hipMemcpyHostToDevice, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    s->cur = 0;
    s->loaded = true;
    return PAT_OK;
}

int pat_stencil_run(pat_stencil *s, int nsteps)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_run: null handle");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_run: no grid uploaded");
    PAT_REQUIRE(nsteps >= 0, PAT_ERR_INVALID, "pat_stencil_run: negative step count");
    if (nsteps == 0) return PAT_OK;
    auto key = std::make_pair(nsteps, s->cur);
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        int tbx, tby, tr, tzc;
        PAT_TRY(s->resolve_tile(&tbx, &tby, &tr, &tzc));  // tunes outside the capture
        hipGraph_t g;
        PAT_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = s->enqueue(nsteps, s->cur);
        hipError_t ec = hipStreamEndCapture(s->stream, &g);
        if (rc != PAT_OK) return rc;
        PAT_HIP(ec);
        hipGraphExec_t ge;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PAT_HIP(ei);
        it = s->graphs.emplace(key, ge).first;
    }
    PAT_HIP(hipGraphLaunch(it->second, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    if (nsteps % 2) s->cur = 1 - s->cur;
    return PAT_OK;
}

int pat_stencil_run_timed(pat_stencil *s, int nsteps, float *total_ms, float *kernel_ms)
{
    PAT_REQUIRE(s && total_ms, PAT_ERR_INVALID, "pat_stencil_run_timed: null argument");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_run_timed: no grid uploaded");
    PAT_REQUIRE(nsteps > 0, PAT_ERR_INVALID, "pat_stencil_run_timed: need at least one step");
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
    for (auto &e : ev) PAT_HIP(hipEventCreate(&e));
    int tbx, tby, tr, tzc;
    PAT_TRY(s->resolve_tile(&tbx, &tby, &tr, &tzc));
    const int v = s->variant == PAT_STENCIL_GENERIC ? PAT_STENCIL_GENERIC : PAT_STENCIL_LDS25D;
    int rc = PAT_OK;
    int c = s->cur;
    for (int k = 0; k < nsteps && rc == PAT_OK; k++) {
        (void)hipEventRecord(ev[2 * k], s->stream);
        rc = pat::heat3d_sweep(s->buf[c], s->buf[1 - c], s->n[0], s->n[1], s->n[2], s->alpha(), v, tbx, tby, tr, tzc,
                               s->stream);
        (void)hipEventRecord(ev[2 * k + 1], s->stream);
        c = 1 - c;
    }
    if (rc == PAT_OK && hipStreamSynchronize(s->stream) != hipSuccess) {
        pat::set_error("pat_stencil_run_timed: stream synchronisation failed");
        rc = PAT_ERR_HIP;
    }
    if (rc == PAT_OK) {
        (void)hipEventElapsedTime(total_ms, ev[0], ev[2 * (size_t)nsteps - 1]);
        if (kernel_ms)
            for (int k = 0; k < nsteps; k++) (void)hipEventElapsedTime(&kernel_ms[k], ev[2 * k], ev[2 * k + 1]);
        s->cur = c;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int pat_stencil_download(pat_stencil *s, void *host)
{
    PAT_REQUIRE(s && host, PAT_ERR_INVALID, "pat_stencil_download: null argument");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_download: no grid uploaded");
    PAT_HIP(hipMemcpyAsync(host, s->buf[s->cur], sizeof(float) * (size_t)s->cells(), hipMemcpyDeviceToHost, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    return PAT_OK;
}

int pat_stencil_tile(pat_stencil *s, int *bx, int *by, int *zc)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_tile: null handle");
    int tbx, tby, tr, tzc;
    PAT_TRY(s->resolve_tile(&tbx, &tby, &tr, &tzc));
    if (bx) *bx = tbx;
    if (by) *by = tby * tr;
    if (zc) *zc = tzc;
    return PAT_OK;
}

int pat_stencil_device_ptr(pat_stencil *s, void **dptr)
{
    PAT_REQUIRE(s && dptr, PAT_ERR_INVALID, "pat_stencil_device_ptr: null argument");
    *dptr = s->buf[s->cur];
    return PAT_OK;
}

int pat_stencil_destroy(pat_stencil *s)
{
    if (!s) return PAT_OK;
    (void)hipStreamSynchronize(s->stream);
    s->drop_graphs();
    (void)pat::dev_free(s->buf[0]);
    (void)pat::dev_free(s->buf[1]);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return PAT_OK;
}

}  // extern "C"
