// Handle-level stencil runtime: the device-resident form of StencilRuntime
// (kmeans/lib/stencil_runtime.h:13-93, heat_3D/lib/stencil_runtime.h:16).
//
// The reference splits the per-node grid into a CPU slab and GPU slabs,
// exchanges halos through MPI and mapped host buffers every iteration and
// never swaps its double buffer (kmeans/lib/grid.h:61-65).  With one GPU
// holding the whole domain, the runtime keeps two halo-inclusive device grids,
// ping-pongs them each step, and replays repeated step counts from a captured
// hipGraph so that small grids are not bound by launch latency.
//
// Functors: the library's tuned heat engine runs PAT_STENCIL_HEAT on 3-D FP32
// grids of halo width 1.  An application that compiles its own functor table
// for the device (api/gpu/stencil_table.hpp) registers it with
// pat_stencil_set_table; then any entry runs, on grids of any halo width and
// element size the entry declares, through the table's one-thread-per-cell
// kernels -- except entries declared equal to the tuned functor, which keep
// the tuned engine unless PAT_STENCIL_GENERIC is selected.
#include <cstring>
#include <map>

#include "pat_common.hpp"

#include "stencil_dev.hpp"

struct pat_stencil {
    int ndims = 3;
    int64_t n[3] = {0, 0, 0};  // interior x, y, z
    int width = 1;
    int unit = 4;
    int idx = PAT_STENCIL_HEAT;
    std::vector<unsigned char> param;
    int variant = PAT_STENCIL_AUTO;
    int bx = 0, by = 0, ry = 0, zc = 0;
    float *buf[2] = {nullptr, nullptr};  // element type: 'unit' bytes (the tuned engine: float)
    int cur = 0;  // buffer holding the current field
    const pat_stencil_table *table = nullptr;  // application functors compiled for the device
    void *param_dev = nullptr;                  // the parameter bytes, for table functors
    bool loaded = false;
    hipStream_t stream = nullptr;
    // captured step sequences keyed by (nsteps, starting buffer), with the
    // buffer each leaves the result in
    std::map<std::pair<int, int>, std::pair<hipGraphExec_t, int>> graphs;

    // halo-inclusive extent of dimension d (1 beyond ndims)
    int64_t ext(int d) const { return d < ndims ? n[d] + 2 * width : 1; }
    int64_t cells() const { return ext(0) * ext(1) * ext(2); }
    size_t bytes() const { return (size_t)unit * (size_t)cells(); }

    // The tuned heat engine runs this handle's functor.
    bool tuned() const
    {
        const bool heat = table ? (table->same_as && table->same_as[idx] == PAT_STENCIL_HEAT) : idx == PAT_STENCIL_HEAT;
        return heat && ndims == 3 && width == 1 && unit == 4 && !(table && variant == PAT_STENCIL_GENERIC);
    }

    // Whether the current functor can run on this grid at all.
    int check_runnable() const
    {
        if (tuned()) return PAT_OK;
        PAT_REQUIRE(table, PAT_ERR_UNSUPPORTED,
                    "stencil: functor %d on a %d-D grid of %d-byte elements with halo width %d needs an application "
                    "functor table (pat_stencil_set_table); the library's own engine runs the heat functor on 3-D "
                    "FP32 grids of halo width 1",
                    idx, ndims, unit, width);
        PAT_REQUIRE(idx >= 0 && idx < table->count, PAT_ERR_UNSUPPORTED, "stencil: functor index %d not in the table",
                    idx);
        PAT_REQUIRE(!table->unit_size || table->unit_size[idx] == unit, PAT_ERR_UNSUPPORTED,
                    "stencil: functor %d reads %d-byte elements, the grid holds %d-byte ones", idx,
                    table->unit_size ? table->unit_size[idx] : 0, unit);
        return PAT_OK;
    }

    // One sweep of the table functor from buffer c into 1 - c.
    int table_sweep(int c, hipStream_t st) const
    {
        const int32_t size[3] = {(int32_t)ext(0), (int32_t)ext(1), (int32_t)ext(2)};
        const int rc = table->sweep(idx, buf[c], buf[1 - c], size, ndims, width, param_dev, st);
        PAT_REQUIRE(rc == PAT_OK, rc, "stencil: the table's sweep of functor %d failed (%d)", idx, rc);
        return PAT_OK;
    }

    void drop_graphs()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second.first);
        graphs.clear();
    }

    float alpha() const
    {
        float a = 0.1f;
        if (param.size() >= sizeof(float)) memcpy(&a, param.data(), sizeof(float));
        return a;
    }

    // The tile the runs use: tuned (plain or fused) for AUTO without an
    // explicit tile, else the variant's default with the set fields.
    int resolve_tile(pat::HeatTile *t)
    {
        if (variant == PAT_STENCIL_AUTO && bx == 0 && by == 0 && zc == 0 && loaded) {
            PAT_TRY(pat::heat3d_tuned(buf[cur], buf[1 - cur], n[0], n[1], n[2], alpha(), stream, true, t));
            return PAT_OK;
        }
        if (variant == PAT_STENCIL_CUBES) return pat::heat3d_cube_tile(bx, t);
        if (variant == PAT_STENCIL_TEMPORAL2) {
            *t = pat::HeatTile{1, 64, 4, ry > 0 ? ry : 8, zc > 0 ? zc : (int)std::min<int64_t>(64, n[2])};
            return PAT_OK;
        }
        t->fused = 0;
        pat::heat3d_default_tile(n[0], n[1], n[2], &t->bx, &t->by, &t->ry, &t->zc);
        if (bx > 0) t->bx = bx;
        if (by > 0) t->by = by;
        if (ry > 0) t->ry = ry;
        if (zc > 0) t->zc = zc;
        return PAT_OK;
    }

    int kernel_variant() const { return variant == PAT_STENCIL_GENERIC ? PAT_STENCIL_GENERIC : PAT_STENCIL_LDS25D; }

    // *end = the buffer holding the result (fused launches advance two
    // steps per buffer swap)
    int enqueue(int nsteps, int start, int *end)
    {
        if (!tuned()) {
            int c = start;
            for (int k = 0; k < nsteps; k++, c = 1 - c) PAT_TRY(table_sweep(c, stream));
            *end = c;
            return PAT_OK;
        }
        pat::HeatTile t;
        PAT_TRY(resolve_tile(&t));
        return pat::heat3d_steps(buf, start, n[0], n[1], n[2], alpha(), t, kernel_variant(), nsteps, stream, end);
    }
};

extern "C" {

int pat_stencil_create(int ndims, const int64_t *interior, int width, int unit_size, const void *param,
                       int64_t param_size, pat_stencil **out)
{
    PAT_REQUIRE(out && interior, PAT_ERR_INVALID, "pat_stencil_create: null argument");
    PAT_REQUIRE(ndims == 2 || ndims == 3, PAT_ERR_UNSUPPORTED, "pat_stencil_create: 2-D or 3-D grids (got %d)", ndims);
    PAT_REQUIRE(width >= 1 && width <= 16, PAT_ERR_UNSUPPORTED, "pat_stencil_create: halo width %d not in 1..16",
                width);
    PAT_REQUIRE(unit_size == 4 || unit_size == 8, PAT_ERR_UNSUPPORTED,
                "pat_stencil_create: element size %d B (4 = float, 8 = double)", unit_size);
    PAT_REQUIRE(param_size >= 0 && (param_size == 0 || param), PAT_ERR_INVALID, "pat_stencil_create: bad parameter");
    for (int d = 0; d < ndims; d++)
        PAT_REQUIRE(interior[d] > 0 && interior[d] + 2 * width < (1 << 20), PAT_ERR_INVALID,
                    "pat_stencil_create: bad extent");
    auto *s = new pat_stencil();
    s->ndims = ndims;
    for (int d = 0; d < 3; d++) s->n[d] = d < ndims ? interior[d] : 1;
    s->width = width;
    s->unit = unit_size;
    if (ndims == 3)
        PAT_REQUIRE(s->n[2] <= 65535, PAT_ERR_UNSUPPORTED, "pat_stencil_create: at most 65,535 interior planes");
    if (param_size > 0) s->param.assign((const unsigned char *)param, (const unsigned char *)param + param_size);
    const size_t bytes = s->bytes();
    for (int b = 0; b < 2; b++) {
        if (pat::dev_malloc(&s->buf[b], bytes) != hipSuccess) {
            (void)hipGetLastError();
            if (s->buf[0]) (void)pat::dev_free(s->buf[0]);
            delete s;
            pat::set_error("pat_stencil_create: hipMalloc of %zu bytes failed", bytes);
            return PAT_ERR_NOMEM;
        }
    }
    if (param_size > 0) {
        if (pat::dev_malloc(&s->param_dev, (size_t)param_size) != hipSuccess ||
            hipMemcpy(s->param_dev, param, (size_t)param_size, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            if (s->param_dev) (void)pat::dev_free(s->param_dev);
            (void)pat::dev_free(s->buf[0]);
            (void)pat::dev_free(s->buf[1]);
            delete s;
            pat::set_error("pat_stencil_create: copying %lld parameter bytes to the device failed",
                           (long long)param_size);
            return PAT_ERR_NOMEM;
        }
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        if (s->param_dev) (void)pat::dev_free(s->param_dev);
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
    if (s->table)
        PAT_REQUIRE(idx >= 0 && idx < s->table->count, PAT_ERR_UNSUPPORTED,
                    "pat_stencil_set_idx: functor index %d not in the table", idx);
    else
        PAT_REQUIRE(idx == PAT_STENCIL_HEAT, PAT_ERR_UNSUPPORTED,
                    "pat_stencil_set_idx: functor index %d not in the library's table (register the application's "
                    "with pat_stencil_set_table)", idx);
    s->idx = idx;
    s->drop_graphs();
    return PAT_OK;
}

int pat_stencil_set_table(pat_stencil *s, const pat_stencil_table *table)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_set_table: null handle");
    PAT_REQUIRE(!table || (table->count > 0 && table->sweep), PAT_ERR_INVALID, "pat_stencil_set_table: bad table");
    if (table)
        PAT_REQUIRE(s->idx < table->count, PAT_ERR_UNSUPPORTED,
                    "pat_stencil_set_table: the handle's functor index %d is not in the table", s->idx);
    s->table = table;
    s->drop_graphs();
    return PAT_OK;
}

int pat_stencil_set_variant(pat_stencil *s, int variant, int bx, int by, int zc)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_set_variant: null handle");
    PAT_REQUIRE(variant >= PAT_STENCIL_AUTO && variant <= PAT_STENCIL_CUBES, PAT_ERR_UNSUPPORTED,
                "pat_stencil_set_variant: unknown variant %d", variant);
    int tby = 0, tr = 0;
    if (variant == PAT_STENCIL_CUBES) {
        pat::HeatTile t;
        PAT_TRY(pat::heat3d_cube_tile(bx, &t));
    } else if (variant == PAT_STENCIL_TEMPORAL2) {
        PAT_REQUIRE((bx == 0 || bx == 64) && (by == 0 || by == 8 || by == 16 || by == 32) && zc >= 0,
                    PAT_ERR_UNSUPPORTED, "pat_stencil_set_variant: fused windows are 64 wide and 8, 16 or 32 rows");
        tby = by > 0 ? 4 : 0;
        tr = by == 8 ? 2 : (by == 16 ? 4 : (by == 32 ? 8 : 0));
    } else if (by > 0) {
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
    const size_t bytes = s->bytes();
    // Both buffers get the same field, so both carry the Dirichlet halo.
    PAT_HIP(hipMemcpyAsync(s->buf[0], host, bytes, hipMemcpyHostToDevice, s->stream));
    PAT_HIP(hipMemcpyAsync(s->buf[1], host, bytes, hipMemcpyHostToDevice, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    s->cur = 0;
    s->loaded = true;
    return PAT_OK;
}

// nsteps sweeps replayed from the handle's captured graph for that step
// count and start buffer (captured on first use, after the tile is
// resolved); device_ms, when non-NULL, receives the replay's time between
// HIP events recorded on the handle's stream around it.  Blocking.
static int stencil_run_graph(pat_stencil *s, int nsteps, float *device_ms)
{
    auto key = std::make_pair(nsteps, s->cur);
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        PAT_TRY(s->check_runnable());
        if (s->tuned()) {
            pat::HeatTile t;
            PAT_TRY(s->resolve_tile(&t));  // tunes outside the capture
        }
        hipGraph_t g;
        int end = s->cur;
        PAT_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = s->enqueue(nsteps, s->cur, &end);
        hipError_t ec = hipStreamEndCapture(s->stream, &g);
        if (rc != PAT_OK) return rc;
        PAT_HIP(ec);
        hipGraphExec_t ge;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PAT_HIP(ei);
        it = s->graphs.emplace(key, std::make_pair(ge, end)).first;
    }
    pat::EventPair ev;
    if (device_ms) {
        PAT_HIP(ev.create());
        PAT_HIP(hipEventRecord(ev.a, s->stream));
    }
    PAT_HIP(hipGraphLaunch(it->second.first, s->stream));
    if (device_ms) PAT_HIP(hipEventRecord(ev.b, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    if (device_ms) PAT_HIP(hipEventElapsedTime(device_ms, ev.a, ev.b));
    s->cur = it->second.second;
    return PAT_OK;
}

int pat_stencil_run(pat_stencil *s, int nsteps)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_run: null handle");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_run: no grid uploaded");
    PAT_REQUIRE(nsteps >= 0, PAT_ERR_INVALID, "pat_stencil_run: negative step count");
    if (nsteps == 0) return PAT_OK;
    return stencil_run_graph(s, nsteps, nullptr);
}

int pat_stencil_run_graph_timed(pat_stencil *s, int nsteps, float *device_ms)
{
    PAT_REQUIRE(s && device_ms, PAT_ERR_INVALID, "pat_stencil_run_graph_timed: null argument");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_run_graph_timed: no grid uploaded");
    PAT_REQUIRE(nsteps > 0, PAT_ERR_INVALID, "pat_stencil_run_graph_timed: need at least one step");
    return stencil_run_graph(s, nsteps, device_ms);
}

int pat_stencil_run_timed(pat_stencil *s, int nsteps, float *total_ms, float *kernel_ms)
{
    PAT_REQUIRE(s && total_ms, PAT_ERR_INVALID, "pat_stencil_run_timed: null argument");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_run_timed: no grid uploaded");
    PAT_REQUIRE(nsteps > 0, PAT_ERR_INVALID, "pat_stencil_run_timed: need at least one step");
    PAT_TRY(s->check_runnable());
    pat::EventVec ev;
    PAT_HIP(ev.create(2 * (size_t)nsteps));
    if (!s->tuned()) {
        int c = s->cur, rc = PAT_OK;
        for (int k = 0; k < nsteps && rc == PAT_OK; k++, c = 1 - c) {
            (void)hipEventRecord(ev[2 * k], s->stream);
            rc = s->table_sweep(c, s->stream);
            (void)hipEventRecord(ev[2 * k + 1], s->stream);
        }
        if (rc == PAT_OK && hipStreamSynchronize(s->stream) != hipSuccess) {
            pat::set_error("pat_stencil_run_timed: stream synchronisation failed");
            rc = PAT_ERR_HIP;
        }
        if (rc != PAT_OK) return rc;
        (void)hipEventElapsedTime(total_ms, ev[0], ev[2 * (size_t)nsteps - 1]);
        if (kernel_ms)
            for (int k = 0; k < nsteps; k++) (void)hipEventElapsedTime(&kernel_ms[k], ev[2 * k], ev[2 * k + 1]);
        s->cur = c;
        return PAT_OK;
    }
    pat::HeatTile t;
    PAT_TRY(s->resolve_tile(&t));
    const int v = s->kernel_variant();
    // one launch per step, or per group of steps for a multi-sweep tile
    // (then each step of the group is reported as its share of the
    // launch); steps left over after the last group use the plain kernel
    pat::HeatTile plain = t;
    const int spl = pat::heat3d_steps_per_launch(t);
    if (t.fused) {
        plain.fused = 0;
        pat::heat3d_default_tile(s->n[0], s->n[1], s->n[2], &plain.bx, &plain.by, &plain.ry, &plain.zc);
    }
    std::vector<int> steps_of;
    int rc = PAT_OK;
    int c = s->cur;
    int launches = 0;
    for (int k = 0; k < nsteps && rc == PAT_OK; launches++) {
        const bool group = spl > 1 && k + spl <= nsteps;
        (void)hipEventRecord(ev[2 * launches], s->stream);
        rc = pat::heat3d_launch(s->buf[c], s->buf[1 - c], s->n[0], s->n[1], s->n[2], s->alpha(), group ? t : plain,
                                v, s->stream);
        (void)hipEventRecord(ev[2 * launches + 1], s->stream);
        steps_of.push_back(group ? spl : 1);
        k += group ? spl : 1;
        c = 1 - c;
    }
    if (rc == PAT_OK && hipStreamSynchronize(s->stream) != hipSuccess) {
        pat::set_error("pat_stencil_run_timed: stream synchronisation failed");
        rc = PAT_ERR_HIP;
    }
    if (rc == PAT_OK) {
        (void)hipEventElapsedTime(total_ms, ev[0], ev[2 * (size_t)launches - 1]);
        if (kernel_ms) {
            int k = 0;
            for (int l = 0; l < launches; l++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, ev[2 * l], ev[2 * l + 1]);
                for (int j = 0; j < steps_of[l]; j++) kernel_ms[k++] = ms / steps_of[l];
            }
        }
        s->cur = c;
    }
    return rc;
}

int pat_stencil_download(pat_stencil *s, void *host)
{
    PAT_REQUIRE(s && host, PAT_ERR_INVALID, "pat_stencil_download: null argument");
    PAT_REQUIRE(s->loaded, PAT_ERR_STATE, "pat_stencil_download: no grid uploaded");
    PAT_HIP(hipMemcpyAsync(host, s->buf[s->cur], s->bytes(), hipMemcpyDeviceToHost, s->stream));
    PAT_HIP(hipStreamSynchronize(s->stream));
    return PAT_OK;
}

int pat_stencil_tile(pat_stencil *s, int *bx, int *by, int *zc)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_stencil_tile: null handle");
    if (!s->tuned()) {  // the table kernels' fixed shape: 64 x 4 cells per workgroup, one plane deep
        if (bx) *bx = 64;
        if (by) *by = 4;
        if (zc) *zc = 1;
        return PAT_OK;
    }
    pat::HeatTile t;
    PAT_TRY(s->resolve_tile(&t));
    if (bx) *bx = t.bx;
    if (by) *by = t.by * t.ry;
    if (zc) *zc = t.zc;
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
    if (s->param_dev) (void)pat::dev_free(s->param_dev);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return PAT_OK;
}

}  // extern "C"
