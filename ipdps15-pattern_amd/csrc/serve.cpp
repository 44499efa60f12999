// Serving loop: heat3D, LJ and kmeans jobs of different sizes arriving over
// time, admitted by device-memory budget and run concurrently.
//
// The reference's runtimes run one application per process (its mains call
// RuntimeInit, one runtime, RuntimeFinalize).  Here one process keeps the
// GPU busy with many jobs: a pool of worker threads ("slots"), each with its
// own HIP stream, takes jobs from an arrival queue.  A job is admitted when
// a slot is free and its device-memory estimate fits in what the budget has
// left, so jobs that arrive while others run start as soon as room frees up
// and their kernels overlap on the device.  Every job allocates from one
// arena reserved at start (no hipMalloc/hipFree per job: hipFree waits for
// the whole device), generates its input on the device from its seed, runs
// the engine, and hands back a checksum (and, if asked, the full result).
//
// Admission is first-fit over the queue in arrival order; after 64
// consecutive admissions that passed over the oldest job, only that job may
// start, so a large job is never starved.  A job whose allocation fails
// while others hold memory, or a fused batch whose allocation fails, is
// queued again and then runs alone (each job of the batch on its own).
//
// Uneven instances are fused: a slot that admits an LJ job or a small
// heat3D job also admits the jobs of that kind waiting with the same step
// count (within the budget and the batch limits) and runs them as one
// batch.  LJ jobs become one MD batch (pat_mdb): every step is one force +
// velocity-Verlet launch over all their systems, and the steps between
// neighbour rebuilds replay a captured hipGraph; each system's sums keep
// their own order.  Heat3D grids are swept two steps per launch over every
// grid, the launches captured as one hipGraph; any tiling gives the serial
// sweep's values.  So every job gets the bits it gets alone.  kmeans jobs
// run alone (their D and K select different kernels).
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "kmeans_dev.hpp"
#include "pat_common.hpp"
#include "stencil_dev.hpp"

namespace pat {
int sum_blocks();
int sum_f32_dev(const float *p, int64_t n, double *part, double *out, hipStream_t st);
int sum_xyz_dev(const float *p, int64_t n, double *part, double *out, hipStream_t st);
int set_w_dev(float *p, int64_t n, float w, hipStream_t st);
int mdb_create_dev(int nsys, const int64_t *offsets, const float *pos4, const double *box, double cutoff,
                   double skin, int every, double dt, pat_mdb **out);
int mdb_positions(pat_mdb *h, int sys, const float **pos4, int64_t *n);
int mdb_lanes(pat_mdb *h);
int kmeans_run_eager(pat_kmeans *h, int iters);
int irregular_preload();
int gen_preload();
}  // namespace pat

namespace {

constexpr int kMaxBypass = 64;
constexpr float kHeatAlpha = 0.1f;  // heat3D update u' = u + 0.1 (sum6 - 6u)
constexpr double kRho = 0.8442, kJitter = 0.05, kCut = 2.5, kSkin = 0.3, kDt = 0.005;
constexpr int kEvery = 20;
// Jobs of one kind advanced as one batch: at most this many, and LJ atoms
// or heat3D cells in all; heat3D grids up to this side (larger ones fill
// the GPU alone, and their tuned tiles beat the batch's fixed one)
constexpr int kFuseMax = 16;
constexpr int64_t kFuseAtoms = int64_t(1) << 21, kFuseCells = int64_t(1) << 23;
constexpr int64_t kFuseHeatSide = 128;

int64_t lj_atoms(const pat_job &j) { return 4 * j.size * j.size * j.size; }

// An LJ job whose box holds three neighbour bins of cutoff + skin per side
// (the batch's binned build needs them; smaller boxes test every pair), or
// a small heat3D grid.
bool fusable(const pat_job &j)
{
    if (j.kind == PAT_JOB_HEAT3D) return j.size <= kFuseHeatSide;
    const double box = (double)j.size * std::pow(4.0 / kRho, 1.0 / 3.0);  // the FCC generator's side
    return j.kind == PAT_JOB_LJ && box >= 3 * (kCut + kSkin) * 1.0001;
}

// What a batch's limit counts: atoms (LJ) or interior cells (heat3D).
int64_t fuse_units(const pat_job &j) { return j.kind == PAT_JOB_LJ ? lj_atoms(j) : j.size * j.size * j.size; }

struct Job {
    int64_t id;
    pat_job d;
    int64_t est;
    double t_submit;
    int attempts = 0;
};

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct pat_serve {
    pat::Arena *arena = nullptr;
    int64_t budget = 0;
    int slots = 0;
    int flags = 0;
    int device = 0;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    int64_t reserved = 0;
    int running = 0;
    int64_t next_id = 1;
    int bypass = 0;
    bool stopping = false;
    std::map<int64_t, pat_job_result> done;
    int64_t submitted = 0, finished = 0;
    std::vector<std::thread> workers;

    // Device bytes a job will hold (its buffers, the handle's work space).
    static int64_t estimate(const pat_job &j)
    {
        const int64_t MB = 1 << 20;
        switch (j.kind) {
        case PAT_JOB_HEAT3D: {
            const int64_t s = j.size + 2;
            return 2 * s * s * s * 4 + MB;
        }
        case PAT_JOB_LJ: {
            // per atom: positions, velocity, force (~80 B), CSR and its
            // sliced and halo copies (~1.1 KB at ~80 entries per row), the
            // rebuild's per-atom slots (512 B) or second column array, sort
            // keys of the spatial slice order
            const int64_t n = 4 * j.size * j.size * j.size;
            return 2400 * n + 2 * MB;
        }
        case PAT_JOB_KMEANS: {
            const int64_t kd = (int64_t)j.k * j.dim;
            // points, assignments and distances, bucket indices (bucketed
            // reduction), rows and the reduction object, per-centre arrays
            return j.size * j.dim * 4 + 12 * j.size + 40 * kd + 80 * (int64_t)j.k + MB;
        }
        default: return 0;
        }
    }

    // Oldest eligible job that fits, under mu.
    bool pick(Job *out)
    {
        for (size_t i = 0; i < pending.size(); i++) {
            const Job &j = pending[i];
            const bool fits = j.attempts > 0 ? running == 0 : (reserved + j.est <= budget || running == 0);
            if (fits) {
                *out = j;
                pending.erase(pending.begin() + (long)i);
                bypass = i > 0 ? bypass + 1 : 0;
                return true;
            }
            if (bypass >= kMaxBypass) return false;  // the oldest job waits for room
        }
        return false;
    }

    // The jobs of g[0]'s kind waiting with the same step count, admitted
    // with it while the budget, the batch limits and the bypass guard allow
    // (under mu).
    void take_partners(std::vector<Job> &g)
    {
        const int kind = g[0].d.kind;
        const int64_t limit = kind == PAT_JOB_LJ ? kFuseAtoms : kFuseCells;
        int64_t units = fuse_units(g[0].d);
        for (size_t i = 0; i < pending.size() && (int)g.size() < kFuseMax && bypass < kMaxBypass;) {
            const Job &q = pending[i];
            if (q.attempts == 0 && q.d.kind == kind && q.d.steps == g[0].d.steps && fusable(q.d) &&
                units + fuse_units(q.d) <= limit && reserved + q.est <= budget) {
                units += fuse_units(q.d);
                reserved += q.est;
                g.push_back(q);
                pending.erase(pending.begin() + (long)i);
                bypass = i > 0 ? bypass + 1 : 0;
            } else {
                i++;
            }
        }
    }

    int run_heat(const pat_job &j, hipStream_t st, double *part, pat_job_result *r)
    {
        const int64_t n = j.size, cells = (n + 2) * (n + 2) * (n + 2);
        float *a = nullptr, *b = nullptr;
        int rc = PAT_OK;
        if (pat::dev_malloc(&a, sizeof(float) * cells) != hipSuccess ||
            pat::dev_malloc(&b, sizeof(float) * cells) != hipSuccess) {
            pat::set_error("serve: heat3D job %lld: no device memory for 2 x %lld cells", (long long)j.size,
                           (long long)cells);
            rc = PAT_ERR_NOMEM;
        }
        int in_b = 0;
        if (rc == PAT_OK) rc = pat_gen_heat3d_dev(n, n, n, j.seed, a, st);
        if (rc == PAT_OK && hipMemsetAsync(b, 0, sizeof(float) * cells, st) != hipSuccess) rc = PAT_ERR_HIP;
        if (rc == PAT_OK)
            rc = pat_heat3d_run_dev(a, b, n, n, n, j.steps, kHeatAlpha, PAT_STENCIL_AUTO, 0, 0, 0, &in_b, st);
        if (rc == PAT_OK) rc = finish_heat(j, in_b ? b : a, st, part, r);
        (void)hipStreamSynchronize(st);
        (void)pat::dev_free(a);
        (void)pat::dev_free(b);
        return rc;
    }

    // A heat3D job's checksum (every cell of the final grid) and, if asked,
    // its interior on the host.
    int finish_heat(const pat_job &j, const float *res, hipStream_t st, double *part, pat_job_result *r)
    {
        const int64_t n = j.size, cells = (n + 2) * (n + 2) * (n + 2);
        int rc = pat::sum_f32_dev(res, cells, part, &r->checksum, st);
        if (rc == PAT_OK && j.out) {
            if (j.out_len < n * n * n) {
                pat::set_error("serve: heat3D job needs %lld output floats", (long long)(n * n * n));
                rc = PAT_ERR_INVALID;
            } else {
                std::vector<float> h((size_t)cells);
                if (hipMemcpyAsync(h.data(), res, sizeof(float) * cells, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess) {
                    rc = PAT_ERR_HIP;
                } else {
                    const int64_t s = n + 2;
                    for (int64_t z = 0; z < n; z++)
                        for (int64_t y = 0; y < n; y++)
                            memcpy(j.out + (z * n + y) * n, h.data() + ((z + 1) * s + (y + 1)) * s + 1,
                                   sizeof(float) * n);
                }
            }
        }
        return rc;
    }

    // Several heat3D jobs as one batch: their sweeps run two per launch over
    // every grid (the fused-window kernel, each workgroup finding its grid
    // in a device table), captured as one hipGraph.  Same bits as each job
    // alone: every tiling gives the serial sweep's results.
    int run_heat_batch(const std::vector<Job> &g, hipStream_t st, double *part, std::vector<pat_job_result> &res)
    {
        const int ng = (int)g.size();
        std::vector<float *> a(ng, nullptr), b(ng, nullptr);
        std::vector<int64_t> side(ng);
        pat::HeatBatchGrid *table = nullptr;
        int rc = PAT_OK;
        for (int i = 0; i < ng && rc == PAT_OK; i++) {
            side[i] = g[i].d.size;
            const int64_t cells = (side[i] + 2) * (side[i] + 2) * (side[i] + 2);
            if (pat::dev_malloc(&a[i], sizeof(float) * cells) != hipSuccess ||
                pat::dev_malloc(&b[i], sizeof(float) * cells) != hipSuccess) {
                pat::set_error("serve: heat3D batch: no device memory for 2 x %lld cells", (long long)cells);
                rc = PAT_ERR_NOMEM;
                break;
            }
            rc = pat_gen_heat3d_dev(side[i], side[i], side[i], g[i].d.seed, a[i], st);
            if (rc == PAT_OK && hipMemsetAsync(b[i], 0, sizeof(float) * cells, st) != hipSuccess) rc = PAT_ERR_HIP;
        }
        if (rc == PAT_OK && pat::dev_malloc(&table, sizeof(pat::HeatBatchGrid) * 2 * ng) != hipSuccess) {
            pat::set_error("serve: heat3D batch: no device memory for the grid table");
            rc = PAT_ERR_NOMEM;
        }
        int nblocks = 0, in_b = 0;
        if (rc == PAT_OK)
            rc = pat::heat3d_batch_table(ng, a.data(), b.data(), side.data(), side.data(), side.data(), table, st,
                                         &nblocks);
        if (rc == PAT_OK && g[0].d.steps > 0) {  // zero steps: the generated grids are the results
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = PAT_ERR_HIP;
            } else {
                rc = pat::heat3d_batch_enqueue(ng, a.data(), b.data(), side.data(), side.data(), side.data(),
                                               kHeatAlpha, g[0].d.steps, table, nblocks, st, &in_b);
                const hipError_t ec = hipStreamEndCapture(st, &graph);
                if (rc == PAT_OK && ec != hipSuccess) rc = PAT_ERR_HIP;
            }
            if (rc == PAT_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = PAT_ERR_HIP;
            if (graph) (void)hipGraphDestroy(graph);
            if (rc == PAT_OK && (hipGraphLaunch(exec, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess))
                rc = PAT_ERR_HIP;
            if (exec) (void)hipGraphExecDestroy(exec);
            if (rc == PAT_ERR_HIP) pat::set_error("serve: heat3D batch: graph capture or replay failed");
        }
        for (int i = 0; i < ng; i++) {
            res[i].batch = ng;
            res[i].status = rc == PAT_OK ? finish_heat(g[i].d, in_b ? b[i] : a[i], st, part, &res[i]) : rc;
        }
        (void)hipStreamSynchronize(st);
        for (int i = 0; i < ng; i++) {
            if (a[i]) (void)pat::dev_free(a[i]);
            if (b[i]) (void)pat::dev_free(b[i]);
        }
        if (table) (void)pat::dev_free(table);
        return rc;
    }

    int run_lj(const pat_job &j, hipStream_t st, double *part, pat_job_result *r)
    {
        const int64_t c = j.size, n = 4 * c * c * c;
        double box[3];
        float *pos4 = nullptr;
        int rc = PAT_OK;
        if (pat::dev_malloc(&pos4, sizeof(float) * 4 * n) != hipSuccess) {
            pat::set_error("serve: LJ job: no device memory for %lld atoms", (long long)n);
            return PAT_ERR_NOMEM;
        }
        rc = pat_gen_fcc_dev(c, c, c, kRho, kJitter, j.seed, pos4, box, st);
        if (rc == PAT_OK && hipStreamSynchronize(st) != hipSuccess) rc = PAT_ERR_HIP;
        pat_md *md = nullptr;
        // neighbour CSR rebuilt on the device every kEvery steps (binned, or
        // every pair tested in boxes of two bins per side)
        if (rc == PAT_OK) rc = pat_md_create_binned_dev(n, pos4, nullptr, box[0], kCut, kSkin, kEvery, kDt, &md);
        if (rc == PAT_OK) {
            int v = 0, l = 0, o = 0;
            if (pat_md_get_variant(md, &v, &l, &o) == PAT_OK) r->kernel = l;
        }
        if (rc == PAT_OK) rc = pat_md_run(md, j.steps);
        if (rc == PAT_OK) {
            void *p = nullptr;
            rc = pat_md_device_buffers(md, &p, nullptr, nullptr);
            if (rc == PAT_OK) rc = pat::sum_xyz_dev((const float *)p, n, part, &r->checksum, st);
        }
        if (rc == PAT_OK && j.out) {
            if (j.out_len < 3 * n) {
                pat::set_error("serve: LJ job needs %lld output floats", (long long)(3 * n));
                rc = PAT_ERR_INVALID;
            } else {
                rc = pat_md_get(md, j.out, nullptr, nullptr);
            }
        }
        if (md) (void)pat_md_destroy(md);
        (void)pat::dev_free(pos4);
        return rc;
    }

    // Several LJ jobs as one pat_mdb batch: each system generated on the
    // device at its offset (its box side in w), one fused launch per step
    // over all of them.  Sets every result's status (the batch's own error,
    // or the job's) and returns the batch's.
    int run_lj_batch(const std::vector<Job> &g, hipStream_t st, double *part, std::vector<pat_job_result> &res)
    {
        const int ns = (int)g.size();
        std::vector<int64_t> off(ns + 1, 0);
        std::vector<double> box(ns, 0.0);
        for (int s = 0; s < ns; s++) off[s + 1] = off[s] + lj_atoms(g[s].d);
        const int64_t N = off[ns];
        float *pos4 = nullptr;
        int rc = PAT_OK;
        if (pat::dev_malloc(&pos4, sizeof(float) * 4 * N) != hipSuccess) {
            pat::set_error("serve: LJ batch: no device memory for %lld atoms", (long long)N);
            rc = PAT_ERR_NOMEM;
        }
        for (int s = 0; s < ns && rc == PAT_OK; s++) {
            const int64_t c = g[s].d.size;
            double b3[3];
            rc = pat_gen_fcc_dev(c, c, c, kRho, kJitter, g[s].d.seed, pos4 + 4 * off[s], b3, st);
            box[s] = b3[0];
            if (rc == PAT_OK) rc = pat::set_w_dev(pos4 + 4 * off[s], off[s + 1] - off[s], (float)b3[0], st);
        }
        if (rc == PAT_OK && hipStreamSynchronize(st) != hipSuccess) rc = PAT_ERR_HIP;
        pat_mdb *h = nullptr;
        if (rc == PAT_OK) rc = pat::mdb_create_dev(ns, off.data(), pos4, box.data(), kCut, kSkin, kEvery, kDt, &h);
        if (pos4) (void)pat::dev_free(pos4);  // the batch holds its own copy
        if (rc == PAT_OK) rc = pat_mdb_run(h, g[0].d.steps, nullptr, nullptr);
        const int lanes = pat::mdb_lanes(h);
        for (int s = 0; s < ns; s++) {
            pat_job_result &r = res[s];
            r.kernel = lanes;
            r.batch = ns;
            int rs = rc;
            const float *p = nullptr;
            int64_t n = 0;
            if (rs == PAT_OK) rs = pat::mdb_positions(h, s, &p, &n);
            if (rs == PAT_OK) rs = pat::sum_xyz_dev(p, n, part, &r.checksum, st);
            if (rs == PAT_OK && g[s].d.out) {
                if (g[s].d.out_len < 3 * n) {
                    pat::set_error("serve: LJ job needs %lld output floats", (long long)(3 * n));
                    rs = PAT_ERR_INVALID;
                } else {
                    rs = pat_mdb_get(h, s, g[s].d.out, nullptr, nullptr);
                }
            }
            r.status = rs;
        }
        if (h) (void)pat_mdb_destroy(h);
        return rc;
    }

    int run_kmeans(const pat_job &j, hipStream_t st, pat_job_result *r)
    {
        float *x = nullptr;
        if (pat::dev_malloc(&x, sizeof(float) * j.size * j.dim) != hipSuccess) {
            pat::set_error("serve: kmeans job: no device memory for %lld x %d points", (long long)j.size, j.dim);
            return PAT_ERR_NOMEM;
        }
        int rc = pat_gen_gmm_dev(j.size, j.dim, j.k, 1.0, -10.0, 10.0, j.seed, x, st);
        if (rc == PAT_OK && hipStreamSynchronize(st) != hipSuccess) rc = PAT_ERR_HIP;
        const int mode = (!(flags & PAT_SERVE_PLAIN) && pat::kmeans_mfma_fits(j.k, j.dim)) ? PAT_KM_MFMA : PAT_KM_PLAIN;
        r->kernel = mode;
        pat_kmeans *km = nullptr;
        if (rc == PAT_OK) rc = pat_kmeans_create(x, j.size, j.dim, j.k, mode, 1, &km);
        // launched one by one: a graph of the iterations, used once, costs more
        // to instantiate than it saves (kmeans p50 at 60 jobs/s: 2.4 vs 2.0 ms)
        if (rc == PAT_OK) rc = pat::kmeans_run_eager(km, j.steps);
        std::vector<float> c((size_t)j.k * j.dim);
        if (rc == PAT_OK) rc = pat_kmeans_get(km, c.data(), nullptr, nullptr);
        if (rc == PAT_OK) {
            double s = 0.0;
            for (float v : c) s += (double)v;
            r->checksum = s;
            if (j.out) {
                if (j.out_len < (int64_t)c.size()) {
                    pat::set_error("serve: kmeans job needs %lld output floats", (long long)c.size());
                    rc = PAT_ERR_INVALID;
                } else {
                    memcpy(j.out, c.data(), sizeof(float) * c.size());
                }
            }
        }
        if (km) (void)pat_kmeans_destroy(km);
        (void)pat::dev_free(x);
        return rc;
    }

    void worker()
    {
        (void)hipSetDevice(device);
        pat::arena_bind(arena);
        pat::set_thread_tuning(false);
        hipStream_t st = nullptr;
        double *part = nullptr;
        const bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
                        pat::dev_malloc(&part, sizeof(double) * pat::sum_blocks()) == hipSuccess;
        for (;;) {
            std::vector<Job> g;  // the job, and the jobs of its kind fused with it
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return pick_ready() || (stopping && pending.empty()); });
                Job j;
                if (!pick(&j)) break;  // stopping with nothing left
                reserved += j.est;
                running++;
                g.push_back(j);
                if (!(flags & PAT_SERVE_NO_FUSE) && j.attempts == 0 && fusable(j.d)) take_partners(g);
            }
            std::vector<pat_job_result> res(g.size());
            const double t_start = ms_since(t0);
            for (size_t s = 0; s < g.size(); s++) {
                memset(&res[s], 0, sizeof(pat_job_result));
                res[s].t_submit_ms = g[s].t_submit;
                res[s].t_start_ms = t_start;
                res[s].device_bytes = g[s].est;
                res[s].batch = 1;
                res[s].status = PAT_ERR_STATE;
            }
            int rc = PAT_ERR_STATE;
            if (ok && g.size() > 1) {
                rc = g[0].d.kind == PAT_JOB_LJ ? run_lj_batch(g, st, part, res) : run_heat_batch(g, st, part, res);
            } else if (ok) {
                switch (g[0].d.kind) {
                case PAT_JOB_HEAT3D: rc = run_heat(g[0].d, st, part, &res[0]); break;
                case PAT_JOB_LJ: rc = run_lj(g[0].d, st, part, &res[0]); break;
                case PAT_JOB_KMEANS: rc = run_kmeans(g[0].d, st, &res[0]); break;
                default: rc = PAT_ERR_INVALID;
                }
                res[0].status = rc;
            }
            const double t_end = ms_since(t0);
            for (auto &r : res) r.t_end_ms = t_end;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (const Job &q : g) reserved -= q.est;
                running--;
                if (rc == PAT_ERR_NOMEM && g[0].attempts == 0 && (running > 0 || g.size() > 1)) {
                    // run again once each can have the arena alone (a batch's
                    // jobs one at a time, whatever else was running)
                    for (size_t s = g.size(); s-- > 0;) {
                        g[s].attempts = 1;
                        pending.push_front(g[s]);
                    }
                } else {
                    for (size_t s = 0; s < g.size(); s++) done[g[s].id] = res[s];
                    finished += (int64_t)g.size();
                    cv_done.notify_all();
                }
            }
            cv_work.notify_all();
        }
        if (part) (void)pat::dev_free(part);
        if (st) (void)hipStreamDestroy(st);
        pat::arena_bind(nullptr);
    }

    // Whether pick() would succeed now (for the condition variable), under mu.
    bool pick_ready()
    {
        for (size_t i = 0; i < pending.size(); i++) {
            const Job &j = pending[i];
            if (j.attempts > 0 ? running == 0 : (reserved + j.est <= budget || running == 0)) return true;
            if (bypass >= kMaxBypass) return false;
        }
        return false;
    }
};

extern "C" {

int pat_serve_start(int64_t hbm_budget, int slots, int flags, pat_serve **out)
{
    PAT_REQUIRE(out && hbm_budget > 0 && slots > 0 && slots <= 64, PAT_ERR_INVALID,
                "pat_serve_start: need a positive budget and 1..64 slots");
    auto *s = new pat_serve();
    s->budget = hbm_budget;
    s->slots = slots;
    s->flags = flags;
    s->device = pat::current_device();
    PAT_HIP(hipSetDevice(s->device));
    PAT_TRY(pat::kmeans_prepare_kernels());  // attribute calls stay off the workers' path
    // every engine's code object loaded now: at a first launch the loading
    // (~55 ms for the LJ kernels) would land on that job's latency
    PAT_TRY(pat::stencil_preload());
    PAT_TRY(pat::irregular_preload());
    PAT_TRY(pat::gen_preload());
    s->arena = pat::arena_create((size_t)hbm_budget);
    if (!s->arena) {
        delete s;
        return PAT_ERR_NOMEM;
    }
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < slots; i++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return PAT_OK;
}

int pat_serve_submit(pat_serve *s, const pat_job *job, int64_t *id)
{
    PAT_REQUIRE(s && job, PAT_ERR_INVALID, "pat_serve_submit: null argument");
    const pat_job &j = *job;
    switch (j.kind) {
    case PAT_JOB_HEAT3D:
        PAT_REQUIRE(j.size >= 1 && j.size <= 2048 && j.steps >= 0, PAT_ERR_INVALID,
                    "pat_serve_submit: heat3D side %lld / steps %d out of range", (long long)j.size, j.steps);
        break;
    case PAT_JOB_LJ:
        // 4 size^3 atoms with ~78 neighbour-list entries each (cutoff + skin
        // 2.8 at density 0.8442) must stay below 2^31 entries: 160^3 cells
        // list ~1.3e9.
        // at least 4 cells: the box must hold two neighbour bins of 2.8
        PAT_REQUIRE(j.size >= 4 && j.size <= 160 && j.steps >= 0, PAT_ERR_INVALID,
                    "pat_serve_submit: LJ cells %lld / steps %d out of range (4..160 cells per side)",
                    (long long)j.size, j.steps);
        break;
    case PAT_JOB_KMEANS:
        PAT_REQUIRE(j.dim >= 1 && j.dim <= 128 && j.k >= 1 && j.k <= j.size && j.steps >= 1, PAT_ERR_INVALID,
                    "pat_serve_submit: kmeans N=%lld D=%d K=%d iters=%d out of range", (long long)j.size, j.dim, j.k,
                    j.steps);
        break;
    default: PAT_REQUIRE(false, PAT_ERR_INVALID, "pat_serve_submit: unknown job kind %d", j.kind);
    }
    Job q;
    q.d = j;
    q.est = pat_serve::estimate(j);
    PAT_REQUIRE(q.est <= s->budget, PAT_ERR_BUDGET, "pat_serve_submit: job needs about %lld bytes, budget %lld",
                (long long)q.est, (long long)s->budget);
    {
        std::lock_guard<std::mutex> g(s->mu);
        PAT_REQUIRE(!s->stopping, PAT_ERR_STATE, "pat_serve_submit: the loop is stopping");
        q.id = s->next_id++;
        q.t_submit = ms_since(s->t0);
        s->pending.push_back(q);
        s->submitted++;
    }
    s->cv_work.notify_one();
    if (id) *id = q.id;
    return PAT_OK;
}

int pat_serve_poll(pat_serve *s, int64_t id, int *done, pat_job_result *res)
{
    PAT_REQUIRE(s && done, PAT_ERR_INVALID, "pat_serve_poll: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    PAT_REQUIRE(id > 0 && id < s->next_id, PAT_ERR_INVALID, "pat_serve_poll: unknown job %lld", (long long)id);
    auto it = s->done.find(id);
    *done = it != s->done.end();
    if (*done && res) *res = it->second;
    return PAT_OK;
}

int pat_serve_wait(pat_serve *s, int64_t id, pat_job_result *res)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_serve_wait: null handle");
    std::unique_lock<std::mutex> g(s->mu);
    PAT_REQUIRE(id > 0 && id < s->next_id, PAT_ERR_INVALID, "pat_serve_wait: unknown job %lld", (long long)id);
    s->cv_done.wait(g, [&] { return s->done.count(id) > 0; });
    if (res) *res = s->done[id];
    return PAT_OK;
}

int pat_serve_drain(pat_serve *s)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_serve_drain: null handle");
    std::unique_lock<std::mutex> g(s->mu);
    s->cv_done.wait(g, [&] { return s->finished == s->submitted; });
    return PAT_OK;
}

int pat_serve_stats(pat_serve *s, int64_t *submitted, int64_t *finished, int64_t *queued, int64_t *reserved_bytes,
                    int64_t *arena_in_use)
{
    PAT_REQUIRE(s, PAT_ERR_INVALID, "pat_serve_stats: null handle");
    std::lock_guard<std::mutex> g(s->mu);
    if (submitted) *submitted = s->submitted;
    if (finished) *finished = s->finished;
    if (queued) *queued = (int64_t)s->pending.size();
    if (reserved_bytes) *reserved_bytes = s->reserved;
    if (arena_in_use) *arena_in_use = (int64_t)pat::arena_in_use(s->arena);
    return PAT_OK;
}

int pat_serve_stop(pat_serve *s)
{
    if (!s) return PAT_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    (void)hipDeviceSynchronize();
    pat::arena_destroy(s->arena);
    delete s;
    return PAT_OK;
}

}  // extern "C"
