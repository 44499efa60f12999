// Handle-level irregular runtime: the device-resident form of
// IrregularRuntime (kmeans/irregular/irregular_runtime.h:12-112) plus the
// fused molecular-dynamics loop the moldyn/miniMD drivers run around it.
//
// The reference's IrregularInit permutes the caller's coordinates, edges and
// node data in place (irregular_runtime.cu:57-69), re-partitions on the host
// at iterations 0 and 1, exchanges halo nodes through MPI every step and
// copies the device hash back for the host to integrate.  Here the caller's
// arrays are never modified: edges are uploaded once and turned into a
// by-destination CSR on the device, node data stays resident, and results
// come back in the caller's node order.
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "pat_common.hpp"

namespace pat {
int64_t csr_workspace_bytes(int64_t N, int64_t E);
int csr_edge_ids(const int32_t *edges, int64_t E, int64_t N, int32_t *rowptr, int32_t *eids, void *work,
                 int64_t work_bytes, hipStream_t st);
int csr_build(const int32_t *edges, int64_t E, int64_t N, int order, int32_t *rowptr, int32_t *cols, void *work,
              int64_t work_bytes, hipStream_t st);
int sell_build(const int32_t *rowptr, const int32_t *cols, const int32_t *perm, int64_t N, double max_pad,
               int32_t *width_d, int32_t *sptr, int32_t **scol, int64_t *cap, int64_t *total, int pad, hipStream_t st);
int sell_perm_build(const float4 *pos, int64_t N, float box, int pbc, float rc2, int32_t *perm, hipStream_t st);
struct SortWork;
SortWork *sort_create();
void sort_destroy(SortWork *w);
int sell_perm_spatial(const float4 *pos, int64_t N, float box, int pbc, float h, const int32_t *sysoff, int nsys,
                      int32_t *perm, SortWork *w, hipStream_t st);
int halo_build(const int32_t *rowptr, const int32_t *cols, const int32_t *perm, const int32_t *sptr, int64_t N,
               int64_t stotal, int32_t *hptr, int32_t **hidx, int64_t *hcap, uint16_t **scol16, int64_t *scap16,
               int *max_nh, hipStream_t st);
int lj_force_halo(const float4 *pos, const int32_t *sptr, const uint16_t *scol16, const int32_t *hptr,
                  const int32_t *hidx, const int32_t *perm, int64_t N, int max_nh, float box, int pbc, float rc2,
                  float dt, int mode, float4 *force, float4 *vel, float4 *pos_out, int U, int *inbox, hipStream_t st);
int lj_force_sell(const float4 *pos, const int32_t *rowptr, const int32_t *sptr, const int32_t *scol,
                  const int32_t *perm, int64_t N, float box, int pbc, float rc2, float dt, int mode, float4 *force,
                  float4 *vel, float4 *pos_out, int kind, int *inbox, hipStream_t st);
int lj_force_csr(const float4 *pos, const int32_t *rowptr, const int32_t *cols, int64_t N, float box, int pbc,
                 float rc2, float dt, int lpr, int mode, float4 *force, float4 *vel, float4 *pos_out, int *inbox,
                 hipStream_t st);
int box_check_dev(const float4 *pos, int64_t N, float box, int pbc, int *flag, hipStream_t st);
int lj_energy_csr(const float4 *pos, const int32_t *rowptr, const int32_t *cols, int64_t N, float box, int pbc,
                  float rc2, double *node_e, hipStream_t st);
int lj_force_atomic(const float4 *pos, const int32_t *edges, int64_t E, int64_t N, float box, int pbc, float rc2,
                    float4 *force, hipStream_t st);
int vv_kick_drift_dev(float4 *pos, float4 *vel, const float4 *force, int64_t N, float box, int pbc, float dt,
                      int *inbox, hipStream_t st);
int vv_kick_dev(float4 *vel, const float4 *force, int64_t N, float dt, hipStream_t st);
int rcp_selftest(uint32_t lo, uint32_t hi, int64_t *bad_one, int64_t *bad_two);
int sell_build_into(const int32_t *rowptr, const int32_t *cols, int64_t N, int32_t *sptr, int32_t *scol, int64_t cap,
                    int64_t *total, hipStream_t st);
struct NbWork;
NbWork *nb_create();
void nb_destroy(NbWork *w);
int neighbor_build_multi(const float4 *pos, int nsys, const double *Ls, const int64_t *off, double cut, NbWork *w,
                         int32_t *rowptr, int32_t **cols, int64_t *cap, int64_t *total, hipStream_t st);
int neighbor_build(const float4 *pos, int64_t N, double L, double cut, NbWork *w, int32_t *rowptr, int32_t **cols,
                   int64_t *cap, int64_t *total, hipStream_t st);
}  // namespace pat

namespace {

constexpr int kForceOnly = 0, kKick = 1, kKickDrift = 2;

// Default lanes per CSR row before tuning: the smallest power of two
// covering about a quarter of the mean degree, within [4, 32].
int default_lpr(int64_t N, int64_t E)
{
    double mean = N > 0 ? 2.0 * (double)E / (double)N : 0.0;
    int l = 4;
    while (l < 32 && l < mean / 4.0) l *= 2;
    return l;
}

// Padded entries the sliced layout may hold, as a multiple of the CSR's.
constexpr double kSellMaxPad = 1.5;

// Tuned lanes per row, cached per (node count, mean degree, row order) class
// for the process: the in-order variants give bit-identical sums, so the
// choice only changes speed.
struct KernelChoice {
    int lanes;      // lanes per row, or PAT_IR_SLICED
    int kind;       // sliced kernel form (lj_force_sell kind)
    int slice_order;  // sliced rows in node (0), row-class (1) or spatial (2) order
};
std::map<std::tuple<int64_t, int, int>, KernelChoice> g_lpr_cache;
// choices made on positions that have moved away from their initial lattice
// (the re-tune of a binned handle at its first rebuild)
std::map<std::tuple<int64_t, int, int>, KernelChoice> g_lpr_cache_moved;
std::mutex g_lpr_mutex;

// Device state shared by the IrregularRuntime handle and the MD handle.
struct IrState {
    int64_t N = 0, E = 0;
    int32_t *edges = nullptr;   // half list (plain kernel)
    int32_t *rowptr = nullptr;  // CSR by destination
    int32_t *cols = nullptr;
    float *pos[2] = {nullptr, nullptr};  // float4 per node
    float *vel = nullptr;
    float *force = nullptr;
    double *node_e = nullptr;
    hipStream_t stream = nullptr;
    int variant = PAT_IR_CSR;
    int lpr = 0;
    int order = PAT_IR_ORDER_NODE;  // row order of a CSR built from the edge list
    // sliced copy of the CSR (see sell_build); sell_ok when it exists
    int32_t *sptr = nullptr, *scol = nullptr, *swidth = nullptr;
    int64_t scap = 0, stotal = 0;
    bool sell_ok = false;
    int sell_kind = 3;  // sliced kernel form (lj_force_sell kind), chosen by tune()
    // device flag, 1 while every position lies inside the box (set by
    // check_box, cleared by any drift that leaves it); the sentinel kinds
    // skip the minimum image per wavefront only while it is set
    int32_t *inbox = nullptr;
    // slice order of the sentinel layout (perm): 0 node order (identity);
    // 1 row class: rows of each 256-row window grouped by row class (nodes
    // within a cutoff of a face, per axis) so most wavefronts take the
    // minimum image on no axis; 2 spatial: rows sorted by the Morton code
    // of a half-cutoff bin, so a slice's rows share one small gather
    // neighbourhood.  Rows keep their entries, so every order gives the
    // same sums; the tuner picks the fastest.
    int32_t *perm = nullptr;
    int order_on = 0;    // order of the current sliced layout
    int want_order = 0;  // order wanted for the next build
    pat::SortWork *sortw = nullptr;
    std::vector<int32_t> sysoff_h;  // batched systems' node offsets (empty: one system)
    int32_t *sysoff_d = nullptr;
    // halo copy of the sentinel layout (kinds 6, 7): per 256-row block the
    // distinct neighbours (hptr/hidx) and the entries as LDS offsets
    int32_t *hptr = nullptr, *hidx = nullptr;
    uint16_t *scol16 = nullptr;
    int64_t hcap = 0, scap16 = 0;
    int halo_max = 0;
    bool halo_ok = false;
    bool want_halo = false;  // built with every sliced layout while set
    int rebuild_every = 0;   // steps between neighbour rebuilds (0: fixed list); the tuner charges
                             // the halo layout's build time over that many force evaluations
    // kernel choice made by tune() (not fixed by the caller); with
    // rebuilds, tuned once more (fresh) at the first rebuild after creation,
    // when the positions have left their initial lattice
    bool choice_tuned = false, retuned = false;

    // Sliced kernels padded with a sentinel node: positions hold N+1
    // entries, the last far beyond any cutoff (set at allocation, never
    // written by a step), which needs N < 2^28 - 1.
    bool sentinel() const { return N < (1LL << 28) - 1; }
    // bumped whenever a buffer or kernel choice the force launches use
    // changes, so captured graphs can tell they are stale
    int64_t layout_gen = 0;

    // csr: build the CSR from host_edges now (false for handles whose CSR
    // comes from a neighbour build)
    int alloc(int64_t n, int64_t e, const int32_t *host_edges, bool md, bool csr = true)
    {
        N = n;
        E = e;
        sell_kind = sentinel() ? 3 : 0;
        PAT_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        auto am = [](void **p, size_t b) -> int {
            if (b == 0) b = 16;
            if (pat::dev_malloc(p, b) != hipSuccess) {
                (void)hipGetLastError();
                pat::set_error("irregular: hipMalloc of %zu bytes failed", b);
                return PAT_ERR_NOMEM;
            }
            return PAT_OK;
        };
        PAT_TRY(am((void **)&edges, sizeof(int32_t) * 2 * (size_t)E));
        PAT_TRY(am((void **)&rowptr, sizeof(int32_t) * (size_t)(N + 1)));
        PAT_TRY(am((void **)&cols, sizeof(int32_t) * 2 * (size_t)E));
        PAT_TRY(am((void **)&pos[0], sizeof(float) * 4 * (size_t)(N + 1)));
        PAT_TRY(am((void **)&force, sizeof(float) * 4 * (size_t)N));
        PAT_TRY(am((void **)&node_e, sizeof(double) * (size_t)N));
        PAT_TRY(am((void **)&inbox, sizeof(int32_t)));
        PAT_HIP(hipMemsetAsync(inbox, 0, sizeof(int32_t), stream));
        if (md) {
            PAT_TRY(am((void **)&pos[1], sizeof(float) * 4 * (size_t)(N + 1)));
            PAT_TRY(am((void **)&vel, sizeof(float) * 4 * (size_t)N));
        }
        // sentinel node N in both position buffers: beyond every cutoff
        const float far[4] = {3.0e38f, 3.0e38f, 3.0e38f, 0.0f};
        for (float *pb : {pos[0], pos[1]})
            if (pb) PAT_HIP(hipMemcpyAsync(pb + 4 * N, far, sizeof(far), hipMemcpyHostToDevice, stream));
        PAT_HIP(hipStreamSynchronize(stream));
        const int64_t nsl = (N + 63) / 64;
        PAT_TRY(am((void **)&sptr, sizeof(int32_t) * (size_t)(nsl + 1)));
        PAT_TRY(am((void **)&perm, sizeof(int32_t) * (size_t)(64 * nsl)));
        PAT_TRY(am((void **)&hptr, sizeof(int32_t) * (size_t)((nsl + 3) / 4 + 1)));
        PAT_TRY(am((void **)&swidth, sizeof(int32_t) * (size_t)nsl));
        if (E > 0) PAT_HIP(hipMemcpyAsync(edges, host_edges, sizeof(int32_t) * 2 * E, hipMemcpyHostToDevice, stream));
        lpr = 1;
        return csr ? build_csr(order) : PAT_OK;
    }

    // (Re)builds the CSR from the uploaded edge list with rows in order o,
    // then its sliced copy.
    int build_csr(int o)
    {
        const int64_t wb = pat::csr_workspace_bytes(N, E);
        void *work = nullptr;
        if (pat::dev_malloc(&work, (size_t)wb) != hipSuccess) {
            (void)hipGetLastError();
            pat::set_error("irregular: hipMalloc of %lld workspace bytes failed", (long long)wb);
            return PAT_ERR_NOMEM;
        }
        int rc = pat::csr_build(edges, E, N, o, rowptr, cols, work, wb, stream);
        hipError_t se = hipStreamSynchronize(stream);
        (void)pat::dev_free(work);
        if (rc != PAT_OK) return rc;
        PAT_HIP(se);
        order = o;
        return build_sell();
    }

    // Sliced copy of the current CSR; left unavailable (in-order kernels on
    // the plain CSR only) when its padding would exceed kSellMaxPad.  With
    // positions p4 (periodic box, sentinel layout) the slices follow the
    // row-class order computed from them.
    // What the captured force launches take as arguments: buffers, kernel
    // choice and launch shape.  A rebuild that leaves all of it unchanged
    // keeps captured graphs valid: their kernels read the rebuilt contents
    // in place.
    std::array<int64_t, 20> launch_sig() const
    {
        return {(int64_t)(intptr_t)rowptr, (int64_t)(intptr_t)cols, (int64_t)(intptr_t)sptr, (int64_t)(intptr_t)scol,
                (int64_t)(intptr_t)scol16, (int64_t)(intptr_t)hptr, (int64_t)(intptr_t)hidx, (int64_t)(intptr_t)perm,
                (int64_t)(intptr_t)force, (int64_t)(intptr_t)vel, (int64_t)(intptr_t)pos[0], (int64_t)(intptr_t)pos[1],
                lpr, variant, sell_kind, order_on, (int64_t)sell_ok, (int64_t)halo_ok, halo_max, N};
    }

    int build_sell(const float *p4 = nullptr, float box = 0.f, int pbc = 0, float rc2 = 0.f)
    {
        const auto before = launch_sig();
        const int rc = build_sell_body(p4, box, pbc, rc2);
        if (launch_sig() != before) layout_gen++;
        return rc;
    }

    int build_sell_body(const float *p4, float box, int pbc, float rc2)
    {
        order_on = 0;
        halo_ok = false;
        if (want_order && p4 && pbc && sentinel()) {
            if (want_order == 1) {
                PAT_TRY(pat::sell_perm_build((const float4 *)p4, N, box, pbc, rc2, perm, stream));
            } else {
                const int nsys = sysoff_h.size() > 2 ? (int)sysoff_h.size() - 1 : 1;
                if (!sortw) sortw = pat::sort_create();
                if (nsys > 1 && !sysoff_d) {
                    if (pat::dev_malloc((void **)&sysoff_d, sizeof(int32_t) * sysoff_h.size()) != hipSuccess) {
                        (void)hipGetLastError();
                        pat::set_error("irregular: hipMalloc of %zu system offsets failed", sysoff_h.size());
                        return PAT_ERR_NOMEM;
                    }
                    PAT_HIP(hipMemcpyAsync(sysoff_d, sysoff_h.data(), sizeof(int32_t) * sysoff_h.size(),
                                           hipMemcpyHostToDevice, stream));
                    PAT_HIP(hipStreamSynchronize(stream));
                }
                PAT_TRY(pat::sell_perm_spatial((const float4 *)p4, N, box, pbc, 0.5f * std::sqrt(rc2),
                                               nsys > 1 ? sysoff_d : nullptr, nsys, perm, sortw, stream));
            }
            order_on = want_order;
        }
        // padded with the sentinel node N when the handle can hold it
        int rc = pat::sell_build(rowptr, cols, order_on ? perm : nullptr, N, kSellMaxPad, swidth, sptr, &scol, &scap,
                                 &stotal, sentinel() ? (int)N : -1, stream);
        if (rc == PAT_ERR_UNSUPPORTED) {
            pat::clear_error();
            sell_ok = false;
            if (lpr == PAT_IR_SLICED && variant == PAT_IR_CSR) lpr = 1;  // same sums, bit for bit
            return PAT_OK;
        }
        if (rc != PAT_OK) return rc;
        sell_ok = true;
        return want_halo ? build_halo_body() : PAT_OK;
    }

    // Halo copy of the current sentinel layout; halo_ok false when some
    // block does not fit it (the global-gather kinds remain).
    int build_halo()
    {
        const auto before = launch_sig();
        const int rc = build_halo_body();
        if (launch_sig() != before) layout_gen++;
        return rc;
    }

    int build_halo_body()
    {
        halo_ok = false;
        if (!sell_ok || !sentinel()) return PAT_OK;
        int rc = pat::halo_build(rowptr, cols, order_on ? perm : nullptr, sptr, N, stotal, hptr, &hidx, &hcap,
                                 &scol16, &scap16, &halo_max, stream);
        if (rc == PAT_ERR_UNSUPPORTED) {
            pat::clear_error();
            if (lpr == PAT_IR_SLICED && sell_kind >= 6) sell_kind = 3;  // same sums, bit for bit
            return PAT_OK;
        }
        if (rc != PAT_OK) return rc;
        halo_ok = true;
        return PAT_OK;
    }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {(void *)edges, (void *)rowptr, (void *)cols, (void *)pos[0], (void *)pos[1], (void *)vel,
                        (void *)force, (void *)node_e, (void *)sptr, (void *)scol, (void *)swidth, (void *)inbox,
                        (void *)perm, (void *)hptr, (void *)hidx, (void *)scol16, (void *)sysoff_d})
            if (p) (void)pat::dev_free(p);
        pat::sort_destroy(sortw);
        sortw = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
    }

    int upload3(float *dst4, const void *src, int elm)
    {
        std::vector<float> h(4 * (size_t)N);
        if (elm == 12) {
            const float *s = (const float *)src;
            for (int64_t i = 0; i < N; i++) {
                h[4 * i] = s[3 * i];
                h[4 * i + 1] = s[3 * i + 1];
                h[4 * i + 2] = s[3 * i + 2];
                h[4 * i + 3] = 0.0f;
            }
        } else if (elm == 24) {
            const double *s = (const double *)src;
            for (int64_t i = 0; i < N; i++) {
                h[4 * i] = (float)s[3 * i];
                h[4 * i + 1] = (float)s[3 * i + 1];
                h[4 * i + 2] = (float)s[3 * i + 2];
                h[4 * i + 3] = 0.0f;
            }
        } else {
            pat::set_error("irregular: node element size %d not supported (12 = float3, 24 = double3)", elm);
            return PAT_ERR_UNSUPPORTED;
        }
        PAT_HIP(hipMemcpyAsync(dst4, h.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice, stream));
        PAT_HIP(hipStreamSynchronize(stream));
        return PAT_OK;
    }

    int download3(float *dst, const float *src4)
    {
        std::vector<float> h(4 * (size_t)N);
        PAT_HIP(hipMemcpyAsync(h.data(), src4, sizeof(float) * 4 * N, hipMemcpyDeviceToHost, stream));
        PAT_HIP(hipStreamSynchronize(stream));
        for (int64_t i = 0; i < N; i++) {
            dst[3 * i] = h[4 * i];
            dst[3 * i + 1] = h[4 * i + 1];
            dst[3 * i + 2] = h[4 * i + 2];
        }
        return PAT_OK;
    }

    // lanes-per-row argument of the force kernels (negative = tree order,
    // PAT_IR_SLICED = the sliced layout)
    int klpr() const { return variant == PAT_IR_CSR_TREE ? -lpr : lpr; }

    // One CSR force launch with kernel choice l (klpr() form) reading
    // positions p4; mode kKick/kKickDrift also update vel (and write pos_out).
    int csr_force_l(int l, const float *p4, float box, int pbc, float rc2, float dt, int mode, float *pos_out,
                    hipStream_t s)
    {
        if (l == PAT_IR_SLICED) {
            PAT_REQUIRE(sell_ok, PAT_ERR_STATE, "irregular: the sliced layout is not built");
            // kinds 3..7 read the sentinel layout (6, 7 its halo copy), 0..2 indices
            PAT_REQUIRE((sell_kind >= 3) == sentinel(), PAT_ERR_STATE, "irregular: sliced kind %d does not fit the layout",
                        sell_kind);
            if (sell_kind >= 6) {
                PAT_REQUIRE(halo_ok, PAT_ERR_STATE, "irregular: the halo layout is not built");
                return pat::lj_force_halo((const float4 *)p4, sptr, scol16, hptr, hidx, order_on ? perm : nullptr, N,
                                          halo_max, box, pbc, rc2, dt, mode, (float4 *)force, (float4 *)vel,
                                          (float4 *)pos_out, sell_kind == 7 ? 8 : 4, inbox, s);
            }
            return pat::lj_force_sell((const float4 *)p4, rowptr, sptr, scol, order_on ? perm : nullptr, N, box, pbc,
                                      rc2, dt, mode, (float4 *)force, (float4 *)vel, (float4 *)pos_out, sell_kind,
                                      inbox, s);
        }
        return pat::lj_force_csr((const float4 *)p4, rowptr, cols, N, box, pbc, rc2, dt, l, mode, (float4 *)force,
                                 (float4 *)vel, (float4 *)pos_out, inbox, s);
    }

    // Sets the inbox flag from positions p4 (stream-ordered before what
    // follows on s).
    int check_box(const float *p4, float box, int pbc, hipStream_t s)
    {
        return pat::box_check_dev((const float4 *)p4, N, box, pbc, inbox, s);
    }

    // Kick-drift of positions p4 in place, after checking them into the
    // inbox flag: the start of every run of steps.
    int begin_steps(float *p4, float box, int pbc, float dt, hipStream_t s)
    {
        PAT_TRY(check_box(p4, box, pbc, s));
        return pat::vv_kick_drift_dev((float4 *)p4, (float4 *)vel, (const float4 *)force, N, box, pbc, dt, inbox, s);
    }

    int csr_force(const float *p4, float box, int pbc, float rc2, float dt, int mode, float *pos_out, hipStream_t s)
    {
        return csr_force_l(klpr(), p4, box, pbc, rc2, dt, mode, pos_out, s);
    }

    int force_eval(const float *p4, float box, int pbc, float rc2)
    {
        PAT_TRY(check_box(p4, box, pbc, stream));
        if (variant == PAT_IR_ATOMIC)
            return pat::lj_force_atomic((const float4 *)p4, edges, E, N, box, pbc, rc2, (float4 *)force, stream);
        return csr_force(p4, box, pbc, rc2, 0.f, kForceOnly, nullptr, stream);
    }

    // Sliced layout in slice order slice_order (see perm), with or without
    // its halo copy; rebuilt from positions p4 where it differs.
    int set_layout(int slice_order, bool halo, const float *p4, float box, int pbc, float rc2)
    {
        want_order = pbc && sentinel() ? slice_order : 0;
        want_halo = halo && sentinel();
        if (!sell_ok) return PAT_OK;
        if (want_order != order_on) return build_sell(p4, box, pbc, rc2);  // with the halo when wanted
        if (want_halo && !halo_ok) return build_halo();
        return PAT_OK;
    }

    // Times the in-order force kernel for each lanes-per-row candidate, each
    // sliced form, and the sliced forms again over the class-ordered layout,
    // on this CSR, and keeps the fastest (cached by size class).  Overwrites
    // the force buffer: callers recompute the force afterwards.
    // moved: the positions have moved away from their initial lattice (the
    // re-tune of a binned handle at its first rebuild): the choice comes
    // from, and goes to, a cache of its own.
    int tune(const float *p4, float box, int pbc, float rc2, bool moved = false)
    {
        auto &cache = moved ? g_lpr_cache_moved : g_lpr_cache;
        const int mdeg = N > 0 ? (int)((2 * E + N / 2) / N) : 0;
        const int k0 = sentinel() ? 3 : 0;  // sliced kinds for this layout
        choice_tuned = true;
        const auto key = std::make_tuple((int64_t)(63 - __builtin_clzll((unsigned long long)N | 1)), mdeg,
                                         order + 2 * k0 + 8 * std::min(rebuild_every, 1 << 20));
        layout_gen++;
        if (!pat::tuning_enabled()) {
            lpr = sell_ok ? PAT_IR_SLICED : 1;
            sell_kind = k0;
            return set_layout(0, false, p4, box, pbc, rc2);
        }
        {
            std::lock_guard<std::mutex> g(g_lpr_mutex);
            auto it = cache.find(key);
            if (it != cache.end() && (it->second.lanes != PAT_IR_SLICED || sell_ok)) {
                lpr = it->second.lanes;
                sell_kind = it->second.kind;
                return set_layout(it->second.slice_order, it->second.kind >= 6, p4, box, pbc, rc2);
            }
        }
        PAT_TRY(check_box(p4, box, pbc, stream));
        PAT_TRY(set_layout(0, true, p4, box, pbc, rc2));
        pat::EventPair ev;
        PAT_HIP(ev.create());
        hipEvent_t a = ev.a, b = ev.b;
        struct Cand {
            int lanes, kind, slice_order;
        };
        std::vector<Cand> cands = {{1, k0, 0}, {4, k0, 0}, {8, k0, 0}, {16, k0, 0}, {32, k0, 0}};
        if (sell_ok)
            for (int k = k0; k < k0 + 3; k++) cands.push_back({PAT_IR_SLICED, k, 0});
        if (halo_ok)
            for (int k = 6; k < 8; k++) cands.push_back({PAT_IR_SLICED, k, 0});
        // a layout rebuilt with the neighbour list costs its build time over
        // rebuild_every force evaluations
        float halo_share = 0.0f;  // per kReps launches
        if (halo_ok && rebuild_every > 0) {
            PAT_HIP(hipStreamSynchronize(stream));
            auto h0 = std::chrono::steady_clock::now();
            PAT_TRY(build_halo());
            PAT_HIP(hipStreamSynchronize(stream));
            const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - h0).count();
            halo_share = 3.0f * ms / (float)rebuild_every;
        }
        std::vector<float> t(cands.size(), 1e30f);
        auto launch = [&](size_t c) {
            sell_kind = cands[c].kind;
            return csr_force_l(cands[c].lanes, p4, box, pbc, rc2, 0.f, kForceOnly, nullptr, stream);
        };
        // times candidates [c0, c1): rounds that time every one in turn,
        // each keeping its best round
        // (a candidate over 2.5x the best so far after its first round is
        // not timed again)
        auto time_range = [&](size_t c0, size_t c1) -> int {
            constexpr int kRounds = 5, kReps = 3;
            for (int r = 0; r < kRounds; r++)
                for (size_t c = c0; c < c1; c++) {
                    if (r > 0 && t[c] > 2.5f * *std::min_element(t.begin(), t.end())) continue;
                    PAT_HIP(hipEventRecord(a, stream));
                    for (int i = 0; i < kReps; i++) PAT_TRY(launch(c));
                    PAT_HIP(hipEventRecord(b, stream));
                    PAT_HIP(hipEventSynchronize(b));
                    float ms = 0;
                    PAT_HIP(hipEventElapsedTime(&ms, a, b));
                    t[c] = std::min(t[c], ms + (cands[c].kind >= 6 ? halo_share : 0.0f));
                }
            return PAT_OK;
        };
        // warm-up of a few milliseconds first (clocks ramp up under load)
        auto w0 = std::chrono::steady_clock::now();
        for (int i = 0; i < 1000; i++) {
            PAT_TRY(launch(cands.size() - 1));
            if (i % 16 == 15) {
                PAT_HIP(hipStreamSynchronize(stream));
                if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(5)) break;
            }
        }
        PAT_TRY(time_range(0, cands.size()));
        // the sliced forms over the class-ordered and the spatially ordered
        // layouts (the halo copy rebuilt for each; its share re-measured)
        for (int so = 1; so <= 2 && sell_ok && pbc && sentinel(); so++) {
            PAT_TRY(set_layout(so, true, p4, box, pbc, rc2));
            const size_t c0 = cands.size();
            for (int k = k0; k < k0 + 3; k++) cands.push_back({PAT_IR_SLICED, k, so});
            if (halo_ok)
                for (int k = 6; k < 8; k++) cands.push_back({PAT_IR_SLICED, k, so});
            t.resize(cands.size(), 1e30f);
            PAT_TRY(time_range(c0, cands.size()));
        }
        size_t bi = 0;
        for (size_t c = 1; c < cands.size(); c++)
            if (t[c] < t[bi]) bi = c;
        if (pat::tune_log() && halo_share > 0.0f)
            fprintf(stderr, "[tune lj N=%lld] kinds 6-7 include %.2f us of halo build per step (rebuild every %d)\n",
                    (long long)N, 1000.0f * halo_share / 3, rebuild_every);
        if (pat::tune_log())
            for (size_t c = 0; c < cands.size(); c++)
                fprintf(stderr, "[tune lj N=%lld] lanes %d kind %d slice-order %d: %.2f us%s\n", (long long)N,
                        cands[c].lanes, cands[c].kind, cands[c].slice_order, 1000.0f * t[c] / 3,
                        c == bi ? "  <- chosen" : "");
        lpr = cands[bi].lanes;
        sell_kind = lpr == PAT_IR_SLICED ? cands[bi].kind : k0;
        PAT_TRY(set_layout(cands[bi].slice_order, lpr == PAT_IR_SLICED && sell_kind >= 6, p4, box, pbc, rc2));
        std::lock_guard<std::mutex> g(g_lpr_mutex);
        cache[key] = KernelChoice{lpr, sell_kind, cands[bi].slice_order};
        return PAT_OK;
    }
};

int check_variant(int variant, int lanes_per_row)
{
    PAT_REQUIRE(variant == PAT_IR_CSR || variant == PAT_IR_ATOMIC || variant == PAT_IR_CSR_TREE, PAT_ERR_UNSUPPORTED,
                "irregular: unknown variant %d", variant);
    if (variant == PAT_IR_CSR)
        PAT_REQUIRE(lanes_per_row == 0 || lanes_per_row == 1 || lanes_per_row == 4 || lanes_per_row == 8 ||
                        lanes_per_row == 16 || lanes_per_row == 32 || lanes_per_row == PAT_IR_SLICED,
                    PAT_ERR_UNSUPPORTED,
                    "irregular: in-order lanes per row %d not in {0 (tuned), 1, 4, 8, 16, 32, %d (sliced)}",
                    lanes_per_row, PAT_IR_SLICED);
    if (variant == PAT_IR_CSR_TREE)
        PAT_REQUIRE(lanes_per_row == 0 || lanes_per_row == 4 || lanes_per_row == 8 || lanes_per_row == 16 ||
                        lanes_per_row == 32 || lanes_per_row == 64,
                    PAT_ERR_UNSUPPORTED, "irregular: tree lanes per row %d not in {0, 4, 8, 16, 32, 64}",
                    lanes_per_row);
    return PAT_OK;
}

int check_sliced(const IrState &st, int variant, int lanes_per_row)
{
    PAT_REQUIRE(!(variant == PAT_IR_CSR && lanes_per_row == PAT_IR_SLICED && !st.sell_ok), PAT_ERR_UNSUPPORTED,
                "irregular: the sliced layout is not available for this CSR (padding above %.1fx)", kSellMaxPad);
    return PAT_OK;
}

int check_sizes(int64_t num_edges, int64_t num_nodes, const int32_t *edges)
{
    PAT_REQUIRE(num_nodes > 0 && num_nodes < (1LL << 31) - 1, PAT_ERR_INVALID, "irregular: node count %lld out of range",
                (long long)num_nodes);
    PAT_REQUIRE(num_edges >= 0 && 2 * num_edges < (1LL << 31), PAT_ERR_INVALID,
                "irregular: edge count %lld out of range", (long long)num_edges);
    PAT_REQUIRE(num_edges == 0 || edges, PAT_ERR_INVALID, "irregular: null edge list");
    for (int64_t e = 0; e < 2 * num_edges; e++)
        PAT_REQUIRE(edges[e] >= 0 && edges[e] < num_nodes, PAT_ERR_INVALID,
                    "irregular: edge %lld references node %d outside [0, %lld)", (long long)(e / 2), edges[e],
                    (long long)num_nodes);
    for (int64_t e = 0; e < num_edges; e++)
        PAT_REQUIRE(edges[2 * e] != edges[2 * e + 1], PAT_ERR_INVALID, "irregular: edge %lld joins node %d to itself",
                    (long long)e, edges[2 * e]);
    return PAT_OK;
}

}  // namespace

// The irregular runtime handle.  map_idx is the functor index the caller
// selected: a library functor (PAT_IR_LJ_*), or with an application table
// registered an index of that table; builtin is the library functor that runs
// it (-1: the table's own per-node kernel runs it).
struct pat_ir {
    IrState st;
    int map_idx = PAT_IR_LJ_MINIMAGE, reduce_idx = 0;
    int builtin = PAT_IR_LJ_MINIMAGE;
    float box = 0.f;
    int pbc = 1;
    float rc2 = 0.f;
    bool have_nodes = false;
    bool have_result = false;
    // application table path
    const pat_ir_table *table = nullptr;
    bool table_kernel = false;  // run the table's kernel even where a library functor has its bits
    std::vector<unsigned char> param;
    void *param_dev = nullptr, *node_raw = nullptr, *out_raw = nullptr;
    int64_t node_raw_bytes = 0;
    int32_t *eids = nullptr, *ident = nullptr, *foreign = nullptr;
    bool last_table = false;  // the last start ran the table's kernel

    int resolve()
    {
        if (!table) {
            builtin = map_idx;
        } else {
            const int same = table->same_as ? table->same_as[map_idx] : -1;
            builtin = table_kernel ? -1 : same;
        }
        return PAT_OK;
    }

    // Device copy of the parameter bytes for the table functors.
    int upload_param()
    {
        if (param_dev) (void)pat::dev_free(param_dev);
        param_dev = nullptr;
        if (param.empty()) return PAT_OK;
        if (pat::dev_malloc(&param_dev, param.size()) != hipSuccess) {
            (void)hipGetLastError();
            param_dev = nullptr;
            pat::set_error("irregular: hipMalloc of %zu parameter bytes failed", param.size());
            return PAT_ERR_NOMEM;
        }
        PAT_HIP(hipMemcpyAsync(param_dev, param.data(), param.size(), hipMemcpyHostToDevice, st.stream));
        PAT_HIP(hipStreamSynchronize(st.stream));
        return PAT_OK;
    }

    // The table kernel's inputs: edge ids per row, identity part index,
    // output values and the refusal flag (built once).
    int table_buffers()
    {
        if (eids) return PAT_OK;
        auto am = [](void **p, size_t b) -> int {
            if (pat::dev_malloc(p, b ? b : 16) != hipSuccess) {
                (void)hipGetLastError();
                *p = nullptr;
                pat::set_error("irregular: hipMalloc of %zu bytes failed", b);
                return PAT_ERR_NOMEM;
            }
            return PAT_OK;
        };
        const int64_t N = st.N, E = st.E;
        int32_t *rp = nullptr;
        PAT_TRY(am((void **)&eids, sizeof(int32_t) * 2 * (size_t)E));
        PAT_TRY(am((void **)&ident, sizeof(int32_t) * (size_t)N));
        PAT_TRY(am((void **)&foreign, sizeof(int32_t)));
        PAT_TRY(am(&out_raw, (size_t)table->value_size * (size_t)N));
        PAT_TRY(am((void **)&rp, sizeof(int32_t) * (size_t)(N + 1)));
        const int64_t wb = pat::csr_workspace_bytes(N, E);
        void *work = nullptr;
        int rc = am(&work, (size_t)wb);
        if (rc == PAT_OK) rc = pat::csr_edge_ids(st.edges, E, N, rp, eids, work, wb, st.stream);
        std::vector<int32_t> iota((size_t)N);
        for (int64_t i = 0; i < N; i++) iota[(size_t)i] = (int32_t)i;
        if (rc == PAT_OK &&
            hipMemcpyAsync(ident, iota.data(), sizeof(int32_t) * (size_t)N, hipMemcpyHostToDevice, st.stream) !=
                hipSuccess)
            rc = PAT_ERR_HIP;
        if (hipStreamSynchronize(st.stream) != hipSuccess && rc == PAT_OK) rc = PAT_ERR_HIP;
        if (work) (void)pat::dev_free(work);
        (void)pat::dev_free(rp);  // equal to st.rowptr: a row's length does not depend on its order
        if (rc == PAT_ERR_HIP) pat::set_error("irregular: building the edge-id rows failed");
        return rc;
    }

    void release_table()
    {
        for (void *p : {param_dev, node_raw, out_raw, (void *)eids, (void *)ident, (void *)foreign})
            if (p) (void)pat::dev_free(p);
        param_dev = node_raw = out_raw = nullptr;
        eids = ident = foreign = nullptr;
    }
};

struct pat_md {
    IrState st;
    float box = 0.f;
    int pbc = 1;
    float rc2 = 0.f;
    float dt = 0.f;
    int cur = 0;  // position buffer holding the current coordinates
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    // binned neighbour rebuild (0 = fixed list)
    int every = 0;
    double cutneigh = 0.0;
    int64_t cap = 0;    // capacity of st.cols
    int64_t steps = 0;  // steps completed
    // a run failed after its first launch was queued: positions, velocities
    // and the step count no longer agree, so later calls refuse
    bool failed = false;
    int64_t rebuilds = 0;
    float rebuild_ms = 0.0f;  // device time of all rebuilds (settled ones)
    pat::EventPair rb_ev;     // around the last rebuild's launches
    bool rb_pending = false;  // rb_ev not yet added to rebuild_ms

    // Adds the last rebuild's device time once its end event has passed.
    int settle_rebuild_time()
    {
        if (!rb_pending) return PAT_OK;
        float ms = 0.0f;
        PAT_HIP(hipEventSynchronize(rb_ev.b));
        PAT_HIP(hipEventElapsedTime(&ms, rb_ev.a, rb_ev.b));
        rebuild_ms += ms;
        rb_pending = false;
        return PAT_OK;
    }
    int begin_rebuild_time()
    {
        PAT_TRY(settle_rebuild_time());
        if (!rb_ev.a) PAT_HIP(rb_ev.create());
        PAT_HIP(hipEventRecord(rb_ev.a, st.stream));
        return PAT_OK;
    }
    int end_rebuild_time()
    {
        PAT_HIP(hipEventRecord(rb_ev.b, st.stream));
        rb_pending = true;
        return PAT_OK;
    }
    pat::NbWork *nbw = nullptr;

    // Rebuilds the CSR from the positions in buffer c.
    int rebuild(int c)
    {
        int64_t total = 0;
        PAT_TRY(begin_rebuild_time());
        PAT_TRY(pat::neighbor_build((const float4 *)st.pos[c], st.N, box, cutneigh, nbw, st.rowptr, &st.cols, &cap,
                                    &total, st.stream));
        PAT_HIP(hipStreamSynchronize(st.stream));
        st.E = total / 2;
        PAT_TRY(st.build_sell(st.pos[c], box, pbc, rc2));
        PAT_TRY(end_rebuild_time());
        if (rebuilds++ >= 1 && st.choice_tuned && !st.retuned) {
            st.retuned = true;
            PAT_TRY(st.tune(st.pos[c], box, pbc, rc2, true));
        }
        return PAT_OK;
    }

    // Binned mode: steps launched eagerly, with a rebuild before the force
    // evaluation of every step whose index is a multiple of 'every'.
    int run_binned(int nsteps, float *kernel_ms)
    {
        hipStream_t s = st.stream;
        pat::EventVec ev;
        if (kernel_ms) PAT_HIP(ev.create(2 * (size_t)nsteps));
        int c = cur;
        int rc = st.begin_steps(st.pos[c], box, pbc, dt, s);
        failed = rc != PAT_OK;
        for (int k = 0; k < nsteps && rc == PAT_OK; k++) {
            const bool last = k == nsteps - 1;
            if ((steps + k + 1) % every == 0) rc = rebuild(c);
            if (rc != PAT_OK) break;
            if (kernel_ms) (void)hipEventRecord(ev[2 * k], s);
            rc = st.csr_force(st.pos[c], box, pbc, rc2, dt, last ? kKick : kKickDrift, st.pos[1 - c], s);
            if (kernel_ms) (void)hipEventRecord(ev[2 * k + 1], s);
            if (!last) c = 1 - c;
        }
        if (rc == PAT_OK && hipStreamSynchronize(s) != hipSuccess) {
            pat::set_error("md: stream synchronisation failed");
            rc = PAT_ERR_HIP;
        }
        if (rc == PAT_OK) {
            if (kernel_ms)
                for (int k = 0; k < nsteps; k++) (void)hipEventElapsedTime(&kernel_ms[k], ev[2 * k], ev[2 * k + 1]);
            cur = c;
            steps += nsteps;
        } else {
            failed = true;
        }
        return rc;
    }

    void drop_graphs()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }

    // kick-drift from the stored force, then nsteps force evaluations; the
    // last step ends with the closing half kick.
    int enqueue(int nsteps, int start)
    {
        hipStream_t s = st.stream;
        if (st.variant == PAT_IR_ATOMIC) {
            for (int k = 0; k < nsteps; k++) {
                PAT_TRY(pat::vv_kick_drift_dev((float4 *)st.pos[start], (float4 *)st.vel, (const float4 *)st.force,
                                               st.N, box, pbc, dt, nullptr, s));
                PAT_TRY(pat::lj_force_atomic((const float4 *)st.pos[start], st.edges, st.E, st.N, box, pbc, rc2,
                                             (float4 *)st.force, s));
                PAT_TRY(pat::vv_kick_dev((float4 *)st.vel, (const float4 *)st.force, st.N, dt, s));
            }
            return PAT_OK;
        }
        PAT_TRY(st.begin_steps(st.pos[start], box, pbc, dt, s));
        int c = start;
        for (int k = 0; k < nsteps; k++) {
            const bool last = k == nsteps - 1;
            PAT_TRY(st.csr_force(st.pos[c], box, pbc, rc2, dt, last ? kKick : kKickDrift, st.pos[1 - c], s));
            if (!last) c = 1 - c;
        }
        return PAT_OK;
    }
};

extern "C" {

// ------------------------------------------------------------ device level

int pat_ir_csr_workspace(int64_t num_nodes, int64_t num_edges, int64_t *bytes)
{
    PAT_REQUIRE(bytes && num_nodes > 0 && num_edges >= 0, PAT_ERR_INVALID, "pat_ir_csr_workspace: bad arguments");
    *bytes = pat::csr_workspace_bytes(num_nodes, num_edges);
    return PAT_OK;
}

int pat_ir_csr_build_dev(const int32_t *edges, int64_t num_edges, int64_t num_nodes, int order, int32_t *rowptr,
                         int32_t *cols, void *work, int64_t work_bytes, void *stream)
{
    PAT_REQUIRE(rowptr && (num_edges == 0 || (edges && cols)) && work, PAT_ERR_INVALID,
                "pat_ir_csr_build_dev: null argument");
    PAT_REQUIRE(order == PAT_IR_ORDER_EDGE || order == PAT_IR_ORDER_NODE, PAT_ERR_INVALID,
                "pat_ir_csr_build_dev: row order %d not in {PAT_IR_ORDER_EDGE, PAT_IR_ORDER_NODE}", order);
    return pat::csr_build(edges, num_edges, num_nodes, order, rowptr, cols, work, work_bytes, (hipStream_t)stream);
}

int pat_lj_force_csr_dev(const float *pos4, const int32_t *rowptr, const int32_t *cols, int64_t num_nodes, float box,
                         int pbc, float cutoff, int lanes_per_row, float *force4, void *stream)
{
    PAT_REQUIRE(pos4 && rowptr && cols && force4, PAT_ERR_INVALID, "pat_lj_force_csr_dev: null argument");
    PAT_REQUIRE(!pbc || box > 2 * cutoff, PAT_ERR_INVALID, "pat_lj_force_csr_dev: box must exceed twice the cutoff");
    return pat::lj_force_csr((const float4 *)pos4, rowptr, cols, num_nodes, box, pbc ? 1 : 0, cutoff * cutoff, 0.f,
                             lanes_per_row, kForceOnly, (float4 *)force4, nullptr, nullptr, nullptr,
                             (hipStream_t)stream);
}

int pat_ir_sell_build_dev(const int32_t *rowptr, const int32_t *cols, int64_t num_nodes, int32_t *sptr,
                          int32_t *scol, int64_t scol_capacity, int64_t *num_entries, void *stream)
{
    PAT_REQUIRE(rowptr && sptr && num_entries && (scol || scol_capacity == 0), PAT_ERR_INVALID,
                "pat_ir_sell_build_dev: null argument");
    return pat::sell_build_into(rowptr, cols, num_nodes, sptr, scol, scol_capacity, num_entries, (hipStream_t)stream);
}

int pat_lj_force_sell_dev(const float *pos4, const int32_t *rowptr, const int32_t *sptr, const int32_t *scol,
                          int64_t num_nodes, float box, int pbc, float cutoff, int kind, float *force4, void *stream)
{
    PAT_REQUIRE(pos4 && rowptr && sptr && scol && force4, PAT_ERR_INVALID, "pat_lj_force_sell_dev: null argument");
    PAT_REQUIRE(kind >= 0 && kind < 3, PAT_ERR_INVALID, "pat_lj_force_sell_dev: kind %d not in 0..2", kind);
    PAT_REQUIRE(!pbc || box > 2 * cutoff, PAT_ERR_INVALID, "pat_lj_force_sell_dev: box must exceed twice the cutoff");
    return pat::lj_force_sell((const float4 *)pos4, rowptr, sptr, scol, nullptr, num_nodes, box, pbc ? 1 : 0,
                              cutoff * cutoff, 0.f, kForceOnly, (float4 *)force4, nullptr, nullptr, kind, nullptr,
                              (hipStream_t)stream);
}

int pat_lj_rcp_selftest(float lo, float hi, int64_t *bad_one_step, int64_t *bad_two_step)
{
    PAT_REQUIRE(bad_one_step && bad_two_step && lo > 0 && hi >= lo && std::isfinite(hi), PAT_ERR_INVALID,
                "pat_lj_rcp_selftest: need 0 < lo <= hi < inf");
    uint32_t a, b;
    memcpy(&a, &lo, 4);
    memcpy(&b, &hi, 4);
    return pat::rcp_selftest(a, b, bad_one_step, bad_two_step);
}

int pat_lj_force_atomic_dev(const float *pos4, const int32_t *edges, int64_t num_edges, int64_t num_nodes, float box,
                            int pbc, float cutoff, float *force4, void *stream)
{
    PAT_REQUIRE(pos4 && force4 && (num_edges == 0 || edges), PAT_ERR_INVALID, "pat_lj_force_atomic_dev: null argument");
    PAT_REQUIRE(num_nodes > 0 && num_nodes < (1LL << 31), PAT_ERR_INVALID, "pat_lj_force_atomic_dev: bad node count");
    PAT_REQUIRE(!pbc || box > 2 * cutoff, PAT_ERR_INVALID, "pat_lj_force_atomic_dev: box must exceed twice the cutoff");
    return pat::lj_force_atomic((const float4 *)pos4, edges, num_edges, num_nodes, box, pbc, cutoff * cutoff,
                                (float4 *)force4, (hipStream_t)stream);
}

// ------------------------------------------------------- IrregularRuntime

int pat_ir_create(int64_t num_edges, int64_t num_nodes, const int32_t *edges, int map_idx, const void *param,
                  int64_t param_size, pat_ir **out)
{
    PAT_REQUIRE(out, PAT_ERR_INVALID, "pat_ir_create: null output");
    PAT_TRY(check_sizes(num_edges, num_nodes, edges));
    auto *h = new pat_ir();
    int rc = pat_ir_set_map_idx(h, map_idx);
    if (rc == PAT_OK) rc = pat_ir_set_param(h, param, param_size);
    if (rc == PAT_OK) rc = h->st.alloc(num_nodes, num_edges, edges, false);
    if (rc != PAT_OK) {
        h->st.release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

int pat_ir_set_map_idx(pat_ir *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_set_map_idx: null handle");
    if (idx == PAT_IR_NO_FUNCTOR && !h->table) {  // an application table follows
        h->map_idx = h->builtin = idx;
        return PAT_OK;
    }
    if (h->table)
        PAT_REQUIRE(idx >= 0 && idx < h->table->map_count, PAT_ERR_UNSUPPORTED,
                    "pat_ir_set_map_idx: functor index %d not in the application's table", idx);
    else
        PAT_REQUIRE(idx == PAT_IR_LJ_MINIMAGE || idx == PAT_IR_LJ_OPEN, PAT_ERR_UNSUPPORTED,
                    "pat_ir_set_map_idx: functor index %d not in the library's table (register the application's "
                    "with pat_ir_set_table)", idx);
    h->map_idx = idx;
    PAT_TRY(h->resolve());
    if (h->builtin >= 0) h->pbc = h->builtin == PAT_IR_LJ_MINIMAGE;
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_set_reduce_idx(pat_ir *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_set_reduce_idx: null handle");
    const int n = h->table ? h->table->reduce_count : 1;
    PAT_REQUIRE(idx >= 0 && idx < n, PAT_ERR_UNSUPPORTED, "pat_ir_set_reduce_idx: reduce index %d not in the table",
                idx);
    h->reduce_idx = idx;
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_set_param(pat_ir *h, const void *param, int64_t param_size)
{
    PAT_REQUIRE(h && (param || param_size == 0) && param_size >= 0, PAT_ERR_INVALID,
                "pat_ir_set_param: bad argument");
    if (h->builtin == PAT_IR_LJ_MINIMAGE) {
        // struct para {double side; double cutoffRadius;} (kmeans/moldyn.h:39-43)
        PAT_REQUIRE(param_size == 2 * (int64_t)sizeof(double), PAT_ERR_INVALID,
                    "pat_ir_set_param: moldyn parameter is {double side, double cutoffRadius}");
        double side, rc;
        memcpy(&side, param, sizeof(double));
        memcpy(&rc, (const char *)param + sizeof(double), sizeof(double));
        PAT_REQUIRE(side > 2 * rc && rc > 0, PAT_ERR_INVALID, "pat_ir_set_param: box must exceed twice the cutoff");
        h->box = (float)side;
        h->rc2 = (float)rc * (float)rc;
    } else if (h->builtin == PAT_IR_LJ_OPEN) {
        // double cutforcesq (miniMD/irregular.h:111)
        PAT_REQUIRE(param_size == (int64_t)sizeof(double), PAT_ERR_INVALID,
                    "pat_ir_set_param: miniMD parameter is double cutforcesq");
        double c2;
        memcpy(&c2, param, sizeof(double));
        PAT_REQUIRE(c2 > 0, PAT_ERR_INVALID, "pat_ir_set_param: cutoff must be positive");
        h->box = 1.0f;
        h->rc2 = (float)c2;
    }
    h->param.assign((const unsigned char *)param, (const unsigned char *)param + param_size);
    if (h->table) PAT_TRY(h->upload_param());
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_set_table(pat_ir *h, const pat_ir_table *table)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_set_table: null handle");
    PAT_REQUIRE(!table || (table->map_count > 0 && table->reduce_count > 0 && table->reduce_rows &&
                           table->value_size > 0 && table->value_size <= 1024),
                PAT_ERR_INVALID, "pat_ir_set_table: bad table");
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    h->release_table();
    h->table = table;
    h->map_idx = 0;
    h->reduce_idx = 0;
    h->table_kernel = false;
    PAT_TRY(h->resolve());
    if (h->builtin >= 0) h->pbc = h->builtin == PAT_IR_LJ_MINIMAGE;
    if (table) PAT_TRY(h->upload_param());
    h->have_nodes = false;
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_set_variant(pat_ir *h, int variant, int lanes_per_row)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_set_variant: null handle");
    if (variant == PAT_IR_TABLE_KERNEL) {
        PAT_REQUIRE(h->table, PAT_ERR_UNSUPPORTED, "pat_ir_set_variant: no application table registered");
        h->table_kernel = true;
        PAT_TRY(h->resolve());
        h->have_result = false;
        return PAT_OK;
    }
    if (h->table_kernel) {
        h->table_kernel = false;
        PAT_TRY(h->resolve());
    }
    PAT_TRY(check_variant(variant, lanes_per_row));
    PAT_TRY(check_sliced(h->st, variant, lanes_per_row));
    h->st.variant = variant;
    h->st.lpr = lanes_per_row ? lanes_per_row : (variant == PAT_IR_CSR_TREE ? default_lpr(h->st.N, h->st.E) : 1);
    h->st.layout_gen++;
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_set_row_order(pat_ir *h, int order)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_set_row_order: null handle");
    PAT_REQUIRE(order == PAT_IR_ORDER_EDGE || order == PAT_IR_ORDER_NODE, PAT_ERR_INVALID,
                "pat_ir_set_row_order: row order %d not in {PAT_IR_ORDER_EDGE, PAT_IR_ORDER_NODE}", order);
    if (order != h->st.order) PAT_TRY(h->st.build_csr(order));
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_reset_node_data(pat_ir *h, const void *node_data, int node_elm_size)
{
    PAT_REQUIRE(h && node_data && node_elm_size > 0, PAT_ERR_INVALID, "pat_ir_reset_node_data: bad argument");
    if (h->table) {
        // the bytes as given, for the application's functors
        const int64_t bytes = (int64_t)node_elm_size * h->st.N;
        if (bytes != h->node_raw_bytes) {
            if (h->node_raw) (void)pat::dev_free(h->node_raw);
            h->node_raw = nullptr;
            h->node_raw_bytes = 0;
            if (pat::dev_malloc(&h->node_raw, (size_t)bytes) != hipSuccess) {
                (void)hipGetLastError();
                h->node_raw = nullptr;
                pat::set_error("pat_ir_reset_node_data: hipMalloc of %lld bytes failed", (long long)bytes);
                return PAT_ERR_NOMEM;
            }
            h->node_raw_bytes = bytes;
        }
        PAT_HIP(hipMemcpyAsync(h->node_raw, node_data, (size_t)bytes, hipMemcpyHostToDevice, h->st.stream));
        PAT_HIP(hipStreamSynchronize(h->st.stream));
    }
    if (node_elm_size == 12 || node_elm_size == 24) {
        PAT_TRY(h->st.upload3(h->st.pos[0], node_data, node_elm_size));
    } else {
        PAT_REQUIRE(h->table, PAT_ERR_UNSUPPORTED,
                    "pat_ir_reset_node_data: node element size %d needs an application table (the library's LJ "
                    "functors read float3 or double3)", node_elm_size);
    }
    h->have_nodes = true;
    h->have_result = false;
    return PAT_OK;
}

int pat_ir_start(pat_ir *h)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_ir_start: null handle");
    PAT_REQUIRE(h->have_nodes, PAT_ERR_STATE, "pat_ir_start: no node data (call pat_ir_reset_node_data)");
    PAT_REQUIRE(h->table || h->builtin >= 0, PAT_ERR_STATE, "pat_ir_start: no functor (register a table)");
    if (h->builtin >= 0) {
        PAT_TRY(h->st.force_eval(h->st.pos[0], h->box, h->pbc, h->rc2));
        PAT_HIP(hipStreamSynchronize(h->st.stream));
        h->last_table = false;
        h->have_result = true;
        return PAT_OK;
    }
    PAT_REQUIRE(h->node_raw, PAT_ERR_STATE, "pat_ir_start: node data was set before the table");
    PAT_TRY(h->table_buffers());
    int32_t refused = 0;
    PAT_HIP(hipMemsetAsync(h->foreign, 0, sizeof(int32_t), h->st.stream));
    const int rc = h->table->reduce_rows(h->map_idx, h->reduce_idx, h->st.rowptr, h->eids, h->st.N, h->node_raw,
                                         h->ident, h->st.edges, h->param_dev, h->out_raw, h->foreign, h->st.stream);
    PAT_REQUIRE(rc == PAT_OK, rc, "pat_ir_start: the table's kernel for map %d / reduce %d failed (%d)", h->map_idx,
                h->reduce_idx, rc);
    PAT_HIP(hipMemcpyAsync(&refused, h->foreign, sizeof(int32_t), hipMemcpyDeviceToHost, h->st.stream));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    PAT_REQUIRE(!refused, PAT_ERR_UNSUPPORTED,
                "pat_ir_start: map %d inserted a value for a node other than the partition's own (only an edge's "
                "endpoints may be keys, each filtered by part_index[node] == part_id)", h->map_idx);
    h->last_table = true;
    h->have_result = true;
    return PAT_OK;
}

int pat_ir_get_result(pat_ir *h, float *out)
{
    PAT_REQUIRE(h && out, PAT_ERR_INVALID, "pat_ir_get_result: null argument");
    PAT_REQUIRE(h->have_result, PAT_ERR_STATE, "pat_ir_get_result: no result (call pat_ir_start)");
    if (h->last_table) {
        PAT_HIP(hipMemcpyAsync(out, h->out_raw, (size_t)h->table->value_size * (size_t)h->st.N, hipMemcpyDeviceToHost,
                               h->st.stream));
        PAT_HIP(hipStreamSynchronize(h->st.stream));
        return PAT_OK;
    }
    return h->st.download3(out, h->st.force);
}

int pat_ir_destroy(pat_ir *h)
{
    if (!h) return PAT_OK;
    if (h->st.stream) (void)hipStreamSynchronize(h->st.stream);
    h->release_table();
    h->st.release();
    delete h;
    return PAT_OK;
}

// ------------------------------------------------------------- MD loop

int pat_md_create(int64_t num_nodes, const float *pos, const float *vel, int64_t num_edges, const int32_t *edges,
                  double box, int pbc, double cutoff, double dt, pat_md **out)
{
    PAT_REQUIRE(out && pos, PAT_ERR_INVALID, "pat_md_create: null argument");
    PAT_REQUIRE(cutoff > 0 && dt > 0 && (!pbc || box > 2 * cutoff), PAT_ERR_INVALID,
                "pat_md_create: need cutoff > 0, dt > 0 and box > 2*cutoff");
    PAT_TRY(check_sizes(num_edges, num_nodes, edges));
    auto *h = new pat_md();
    h->box = (float)box;
    h->pbc = pbc ? 1 : 0;
    h->rc2 = (float)cutoff * (float)cutoff;
    h->dt = (float)dt;
    int rc = h->st.alloc(num_nodes, num_edges, edges, true);
    if (rc == PAT_OK) rc = h->st.upload3(h->st.pos[0], pos, 12);
    if (rc == PAT_OK) {
        std::vector<float> zero(3 * (size_t)num_nodes, 0.0f);
        rc = h->st.upload3(h->st.vel, vel ? vel : zero.data(), 12);
    }
    if (rc == PAT_OK) rc = h->st.tune(h->st.pos[0], h->box, h->pbc, h->rc2);
    if (rc == PAT_OK) rc = h->st.force_eval(h->st.pos[0], h->box, h->pbc, h->rc2);
    if (rc == PAT_OK && hipStreamSynchronize(h->st.stream) != hipSuccess) {
        pat::set_error("pat_md_create: initial force evaluation failed");
        rc = PAT_ERR_HIP;
    }
    if (rc != PAT_OK) {
        h->st.release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

namespace {

// Binned MD handle whose positions/velocities come from host float3 arrays
// (dev = false) or device float4 arrays (dev = true, vel may be null).
int md_create_binned(int64_t num_nodes, const float *pos, const float *vel, bool dev, double box, double cutoff,
                     double skin, int every, double dt, pat_md **out)
{
    PAT_REQUIRE(out && pos, PAT_ERR_INVALID, "pat_md_create_binned: null argument");
    PAT_REQUIRE(num_nodes > 1 && num_nodes < (1LL << 31) - 1, PAT_ERR_INVALID, "pat_md_create_binned: bad node count");
    PAT_REQUIRE(cutoff > 0 && skin >= 0 && dt > 0 && every > 0, PAT_ERR_INVALID,
                "pat_md_create_binned: need cutoff > 0, skin >= 0, dt > 0, every > 0");
    // three bins of cutoff + skin per side, or two (every pair then has one
    // image within the cutoff) for at most 65,536 atoms, whose rows are
    // built by testing every pair
    PAT_REQUIRE(box >= 3 * (cutoff + skin) * 1.0001 || (box >= 2 * (cutoff + skin) * 1.0001 && num_nodes <= 65536),
                PAT_ERR_UNSUPPORTED,
                "pat_md_create_binned: box %.3f must hold 3 bins of cutoff+skin %.3f (2 for at most 65,536 atoms)", box,
                cutoff + skin);
    auto *h = new pat_md();
    h->box = (float)box;
    h->pbc = 1;
    h->rc2 = (float)cutoff * (float)cutoff;
    h->dt = (float)dt;
    h->every = every;
    h->cutneigh = cutoff + skin;
    h->nbw = pat::nb_create();
    int rc = h->st.alloc(num_nodes, 0, nullptr, true, false);
    if (rc == PAT_OK) {
        (void)pat::dev_free(h->st.cols);
        h->st.cols = nullptr;
        if (dev) {
            const size_t b4 = sizeof(float) * 4 * (size_t)num_nodes;
            if (hipMemcpyAsync(h->st.pos[0], pos, b4, hipMemcpyDeviceToDevice, h->st.stream) != hipSuccess ||
                (vel ? hipMemcpyAsync(h->st.vel, vel, b4, hipMemcpyDeviceToDevice, h->st.stream)
                     : hipMemsetAsync(h->st.vel, 0, b4, h->st.stream)) != hipSuccess ||
                hipStreamSynchronize(h->st.stream) != hipSuccess) {
                pat::set_error("pat_md_create_binned_dev: copying the device arrays failed");
                rc = PAT_ERR_HIP;
            }
        } else {
            rc = h->st.upload3(h->st.pos[0], pos, 12);
        }
    }
    if (rc == PAT_OK && !dev) {
        std::vector<float> zero(3 * (size_t)num_nodes, 0.0f);
        rc = h->st.upload3(h->st.vel, vel ? vel : zero.data(), 12);
    }
    h->st.rebuild_every = every;
    if (rc == PAT_OK) rc = h->rebuild(0);
    if (rc == PAT_OK) rc = h->st.tune(h->st.pos[0], h->box, h->pbc, h->rc2);
    if (rc == PAT_OK) rc = h->st.force_eval(h->st.pos[0], h->box, h->pbc, h->rc2);
    if (rc == PAT_OK && hipStreamSynchronize(h->st.stream) != hipSuccess) {
        pat::set_error("pat_md_create_binned: initial force evaluation failed");
        rc = PAT_ERR_HIP;
    }
    if (rc != PAT_OK) {
        pat::nb_destroy(h->nbw);
        h->st.release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

}  // namespace

int pat_md_create_binned(int64_t num_nodes, const float *pos, const float *vel, double box, double cutoff,
                         double skin, int every, double dt, pat_md **out)
{
    return md_create_binned(num_nodes, pos, vel, false, box, cutoff, skin, every, dt, out);
}

int pat_md_create_binned_dev(int64_t num_nodes, const float *pos4, const float *vel4, double box, double cutoff,
                             double skin, int every, double dt, pat_md **out)
{
    return md_create_binned(num_nodes, pos4, vel4, true, box, cutoff, skin, every, dt, out);
}

int pat_md_stats(pat_md *h, int64_t *num_pairs, int64_t *steps, int64_t *rebuilds, float *rebuild_ms)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_stats: null handle");
    PAT_TRY(h->settle_rebuild_time());
    if (num_pairs) *num_pairs = h->st.E;
    if (steps) *steps = h->steps;
    if (rebuilds) *rebuilds = h->rebuilds;
    if (rebuild_ms) *rebuild_ms = h->rebuild_ms;
    return PAT_OK;
}

int pat_md_csr(pat_md *h, int32_t *rowptr, int32_t *cols)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_csr: null handle");
    if (rowptr)
        PAT_HIP(hipMemcpyAsync(rowptr, h->st.rowptr, sizeof(int32_t) * (h->st.N + 1), hipMemcpyDeviceToHost,
                               h->st.stream));
    if (cols)
        PAT_HIP(hipMemcpyAsync(cols, h->st.cols, sizeof(int32_t) * 2 * h->st.E, hipMemcpyDeviceToHost, h->st.stream));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    return PAT_OK;
}

int pat_md_set_variant(pat_md *h, int variant, int lanes_per_row)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_set_variant: null handle");
    PAT_TRY(check_variant(variant, lanes_per_row));
    PAT_REQUIRE(!(variant == PAT_IR_ATOMIC && h->every > 0), PAT_ERR_UNSUPPORTED,
                "pat_md_set_variant: the atomic kernel needs a half edge list (binned handles have a CSR only)");
    PAT_TRY(check_sliced(h->st, variant, lanes_per_row));
    h->st.variant = variant;
    h->st.layout_gen++;
    h->st.choice_tuned = false;
    if (lanes_per_row) h->st.lpr = lanes_per_row;
    else if (variant == PAT_IR_CSR_TREE) h->st.lpr = default_lpr(h->st.N, h->st.E);
    else PAT_TRY(h->st.tune(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    h->drop_graphs();
    // the stored force seeds the next half kick: recompute it with this variant
    PAT_TRY(h->st.force_eval(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    return PAT_OK;
}

int pat_md_set_sliced_form(pat_md *h, int form)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_set_sliced_form: null handle");
    PAT_REQUIRE(form >= 0 && form < 5, PAT_ERR_INVALID, "pat_md_set_sliced_form: form %d not in 0..4", form);
    PAT_TRY(check_sliced(h->st, PAT_IR_CSR, PAT_IR_SLICED));
    if (form >= 3) {
        PAT_REQUIRE(h->st.sentinel(), PAT_ERR_UNSUPPORTED, "pat_md_set_sliced_form: the halo forms need N < 2^28 - 1");
        PAT_TRY(h->st.set_layout(h->st.want_order, true, h->st.pos[h->cur], h->box, h->pbc, h->rc2));
        PAT_REQUIRE(h->st.halo_ok, PAT_ERR_UNSUPPORTED, "pat_md_set_sliced_form: the list does not fit the halo form");
    }
    h->st.variant = PAT_IR_CSR;
    h->st.lpr = PAT_IR_SLICED;
    h->st.choice_tuned = false;
    h->st.sell_kind = form >= 3 ? form + 3 : (h->st.sentinel() ? 3 : 0) + form;
    h->st.layout_gen++;
    h->drop_graphs();
    PAT_TRY(h->st.force_eval(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    return PAT_OK;
}

int pat_md_set_slice_order(pat_md *h, int order)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_set_slice_order: null handle");
    PAT_REQUIRE(order >= 0 && order <= 2, PAT_ERR_INVALID, "pat_md_set_slice_order: order %d not in 0..2", order);
    PAT_REQUIRE(h->st.sell_ok && h->st.sentinel() && h->pbc, PAT_ERR_UNSUPPORTED,
                "pat_md_set_slice_order: needs the sliced sentinel layout of a periodic box");
    h->st.choice_tuned = false;
    PAT_TRY(h->st.set_layout(order, h->st.want_halo, h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    h->st.layout_gen++;
    h->drop_graphs();
    PAT_TRY(h->st.force_eval(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    return PAT_OK;
}

int pat_md_get_slice_order(pat_md *h, int *order)
{
    PAT_REQUIRE(h && order, PAT_ERR_INVALID, "pat_md_get_slice_order: null argument");
    *order = h->st.order_on;
    return PAT_OK;
}

int pat_md_get_sliced_form(pat_md *h, int *form)
{
    PAT_REQUIRE(h && form, PAT_ERR_INVALID, "pat_md_get_sliced_form: null argument");
    const IrState &st = h->st;
    if (st.variant != PAT_IR_CSR || st.lpr != PAT_IR_SLICED) *form = -1;
    else if (st.sell_kind >= 6) *form = st.sell_kind - 3;
    else *form = st.sell_kind - (st.sentinel() ? 3 : 0);
    return PAT_OK;
}

int pat_md_get_variant(pat_md *h, int *variant, int *lanes_per_row, int *order)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_get_variant: null handle");
    if (variant) *variant = h->st.variant;
    if (lanes_per_row) *lanes_per_row = h->st.lpr;
    if (order) *order = h->st.order;
    return PAT_OK;
}

int pat_md_set_row_order(pat_md *h, int order)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_set_row_order: null handle");
    PAT_REQUIRE(order == PAT_IR_ORDER_EDGE || order == PAT_IR_ORDER_NODE, PAT_ERR_INVALID,
                "pat_md_set_row_order: row order %d not in {PAT_IR_ORDER_EDGE, PAT_IR_ORDER_NODE}", order);
    if (h->every > 0) {
        PAT_REQUIRE(order == PAT_IR_ORDER_NODE, PAT_ERR_UNSUPPORTED,
                    "pat_md_set_row_order: binned handles have no edge list; their rows are in node order");
        return PAT_OK;
    }
    if (order == h->st.order) return PAT_OK;
    PAT_TRY(h->st.build_csr(order));
    if (h->st.variant == PAT_IR_CSR) PAT_TRY(h->st.tune(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    h->drop_graphs();
    // the stored force seeds the next half kick: recompute it in the new order
    PAT_TRY(h->st.force_eval(h->st.pos[h->cur], h->box, h->pbc, h->rc2));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    return PAT_OK;
}

// nsteps steps (no rebuilds) replayed from the handle's captured graph for
// that step count and start buffer, captured on first use; device_ms, when
// non-NULL, receives the replay's time between HIP events recorded on the
// handle's stream around it.  Blocking.
static int md_run_graph(pat_md *h, int nsteps, float *device_ms)
{
    // the atomic path updates positions in place; the CSR path ping-pongs
    const int start = h->cur;
    auto key = std::make_pair(nsteps, start);
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
        hipGraph_t g;
        PAT_HIP(hipStreamBeginCapture(h->st.stream, hipStreamCaptureModeThreadLocal));
        int rc = h->enqueue(nsteps, start);
        hipError_t ec = hipStreamEndCapture(h->st.stream, &g);
        if (rc != PAT_OK) return rc;
        PAT_HIP(ec);
        hipGraphExec_t ge;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PAT_HIP(ei);
        it = h->graphs.emplace(key, ge).first;
    }
    pat::EventPair ev;
    if (device_ms) {
        PAT_HIP(ev.create());
        PAT_HIP(hipEventRecord(ev.a, h->st.stream));
    }
    h->failed = true;  // until the replay has completed
    PAT_HIP(hipGraphLaunch(it->second, h->st.stream));
    if (device_ms) PAT_HIP(hipEventRecord(ev.b, h->st.stream));
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    h->failed = false;
    if (device_ms) PAT_HIP(hipEventElapsedTime(device_ms, ev.a, ev.b));
    if (h->st.variant != PAT_IR_ATOMIC && (nsteps - 1) % 2 == 1) h->cur = 1 - h->cur;
    h->steps += nsteps;
    return PAT_OK;
}

int pat_md_run(pat_md *h, int nsteps)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_run: null handle");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_run: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(nsteps >= 0, PAT_ERR_INVALID, "pat_md_run: negative step count");
    if (nsteps == 0) return PAT_OK;
    if (h->every > 0) return h->run_binned(nsteps, nullptr);
    return md_run_graph(h, nsteps, nullptr);
}

int pat_md_run_graph_timed(pat_md *h, int nsteps, float *device_ms)
{
    PAT_REQUIRE(h && device_ms, PAT_ERR_INVALID, "pat_md_run_graph_timed: null argument");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_run_graph_timed: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(nsteps > 0, PAT_ERR_INVALID, "pat_md_run_graph_timed: need at least one step");
    PAT_REQUIRE(h->every == 0, PAT_ERR_UNSUPPORTED, "pat_md_run_graph_timed: binned rebuilds run outside the graph");
    return md_run_graph(h, nsteps, device_ms);
}

int pat_md_run_async(pat_md *h, int nsteps, void *stream)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_run_async: null handle");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_run_async: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(h->every == 0, PAT_ERR_UNSUPPORTED, "pat_md_run_async: binned rebuilds need blocking runs");
    PAT_REQUIRE(nsteps >= 0, PAT_ERR_INVALID, "pat_md_run_async: negative step count");
    if (nsteps == 0) return PAT_OK;
    hipStream_t saved = h->st.stream;
    h->st.stream = stream ? (hipStream_t)stream : saved;
    int rc = h->enqueue(nsteps, h->cur);
    h->st.stream = saved;
    if (rc != PAT_OK) {
        h->failed = true;
        return rc;
    }
    if (h->st.variant != PAT_IR_ATOMIC && (nsteps - 1) % 2 == 1) h->cur = 1 - h->cur;
    h->steps += nsteps;
    return PAT_OK;
}

int pat_md_run_timed(pat_md *h, int nsteps, float *total_ms, float *kernel_ms)
{
    PAT_REQUIRE(h && total_ms, PAT_ERR_INVALID, "pat_md_run_timed: null argument");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_run_timed: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(nsteps > 0, PAT_ERR_INVALID, "pat_md_run_timed: need at least one step");
    if (h->every > 0) {
        auto t0 = std::chrono::steady_clock::now();
        int r = h->run_binned(nsteps, kernel_ms);
        *total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return r;
    }
    hipStream_t s = h->st.stream;
    pat::EventVec ev;
    PAT_HIP(ev.create(2 * (size_t)nsteps + 2));
    PAT_HIP(hipEventRecord(ev[0], s));
    int c = h->cur;
    int rc = PAT_OK;
    if (h->st.variant == PAT_IR_ATOMIC) {
        for (int k = 0; k < nsteps && rc == PAT_OK; k++) {
            rc = pat::vv_kick_drift_dev((float4 *)h->st.pos[c], (float4 *)h->st.vel, (const float4 *)h->st.force,
                                        h->st.N, h->box, h->pbc, h->dt, nullptr, s);
            if (rc == PAT_OK) rc = hipEventRecord(ev[1 + 2 * k], s) == hipSuccess ? PAT_OK : PAT_ERR_HIP;
            if (rc == PAT_OK)
                rc = pat::lj_force_atomic((const float4 *)h->st.pos[c], h->st.edges, h->st.E, h->st.N, h->box, h->pbc,
                                          h->rc2, (float4 *)h->st.force, s);
            if (rc == PAT_OK) rc = hipEventRecord(ev[2 + 2 * k], s) == hipSuccess ? PAT_OK : PAT_ERR_HIP;
            if (rc == PAT_OK) rc = pat::vv_kick_dev((float4 *)h->st.vel, (const float4 *)h->st.force, h->st.N, h->dt, s);
        }
    } else {
        rc = h->st.begin_steps(h->st.pos[c], h->box, h->pbc, h->dt, s);
        for (int k = 0; k < nsteps && rc == PAT_OK; k++) {
            const bool last = k == nsteps - 1;
            rc = hipEventRecord(ev[1 + 2 * k], s) == hipSuccess ? PAT_OK : PAT_ERR_HIP;
            if (rc == PAT_OK)
                rc = h->st.csr_force(h->st.pos[c], h->box, h->pbc, h->rc2, h->dt, last ? kKick : kKickDrift,
                                     h->st.pos[1 - c], s);
            if (rc == PAT_OK) rc = hipEventRecord(ev[2 + 2 * k], s) == hipSuccess ? PAT_OK : PAT_ERR_HIP;
            if (!last) c = 1 - c;
        }
    }
    if (rc == PAT_OK && hipEventRecord(ev[2 * (size_t)nsteps + 1], s) != hipSuccess) rc = PAT_ERR_HIP;
    if (rc == PAT_OK && hipStreamSynchronize(s) != hipSuccess) rc = PAT_ERR_HIP;
    if (rc == PAT_OK) {
        (void)hipEventElapsedTime(total_ms, ev[0], ev[2 * (size_t)nsteps + 1]);
        if (kernel_ms)
            for (int k = 0; k < nsteps; k++) (void)hipEventElapsedTime(&kernel_ms[k], ev[1 + 2 * k], ev[2 + 2 * k]);
        h->cur = c;
        h->steps += nsteps;
    } else {
        h->failed = true;
    }
    if (rc == PAT_ERR_HIP) {
        pat::set_error("pat_md_run_timed: HIP call failed: %s", hipGetErrorString(hipGetLastError()));
    }
    return rc;
}

int pat_md_get(pat_md *h, float *pos, float *vel, float *force)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_get: null handle");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_get: an earlier run of this handle failed part-way; its state is undefined");
    PAT_HIP(hipStreamSynchronize(h->st.stream));
    if (pos) PAT_TRY(h->st.download3(pos, h->st.pos[h->cur]));
    if (vel) PAT_TRY(h->st.download3(vel, h->st.vel));
    if (force) PAT_TRY(h->st.download3(force, h->st.force));
    return PAT_OK;
}

int pat_md_energy(pat_md *h, double *epot, double *ekin)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_energy: null handle");
    PAT_REQUIRE(!h->failed, PAT_ERR_STATE,
                "pat_md_energy: an earlier run of this handle failed part-way; its state is undefined");
    if (epot) {
        PAT_TRY(pat::lj_energy_csr((const float4 *)h->st.pos[h->cur], h->st.rowptr, h->st.cols, h->st.N, h->box,
                                   h->pbc, h->rc2, h->st.node_e, h->st.stream));
        std::vector<double> e((size_t)h->st.N);
        PAT_HIP(hipMemcpyAsync(e.data(), h->st.node_e, sizeof(double) * h->st.N, hipMemcpyDeviceToHost, h->st.stream));
        PAT_HIP(hipStreamSynchronize(h->st.stream));
        double s = 0;
        for (double v : e) s += v;
        *epot = s;
    }
    if (ekin) {
        std::vector<float> v(3 * (size_t)h->st.N);
        PAT_TRY(h->st.download3(v.data(), h->st.vel));
        double s = 0;
        for (float x : v) s += (double)x * x;
        *ekin = 0.5 * s;
    }
    return PAT_OK;
}

int pat_md_device_buffers(pat_md *h, void **pos4, void **vel4, void **force4)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_md_device_buffers: null handle");
    if (pos4) *pos4 = h->st.pos[h->cur];
    if (vel4) *vel4 = h->st.vel;
    if (force4) *force4 = h->st.force;
    return PAT_OK;
}

int pat_md_destroy(pat_md *h)
{
    if (!h) return PAT_OK;
    if (h->nbw) pat::nb_destroy(h->nbw);
    h->drop_graphs();
    h->st.release();
    delete h;
    return PAT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------
// Batched independent systems: nodes of every system concatenated, each node
// carrying its system's box side in the float4 w lane, so one fused CSR
// kernel launch advances all systems (uneven sizes included) by one step,
// and the steps between neighbour rebuilds replay as one captured graph.

struct pat_mdb {
    pat_md md;                      // the concatenated system (pbc mode 2)
    std::vector<int64_t> off;       // node offsets, nsys + 1
    std::vector<double> box;        // per system

    // One binned build over every system (global atom indices, so the CSR
    // is the concatenated one), then the sliced layout.
    int rebuild_all(int c)
    {
        PAT_TRY(md.begin_rebuild_time());
        int64_t entries = 0;
        PAT_TRY(pat::neighbor_build_multi((const float4 *)md.st.pos[c], (int)box.size(), box.data(), off.data(),
                                          md.cutneigh, md.nbw, md.st.rowptr, &md.st.cols, &md.cap, &entries,
                                          md.st.stream));
        PAT_HIP(hipStreamSynchronize(md.st.stream));
        md.st.E = entries / 2;
        PAT_TRY(md.st.build_sell(md.st.pos[c], 1.0f, 2, md.rc2));
        PAT_TRY(md.end_rebuild_time());
        if (md.rebuilds++ >= 1 && md.st.choice_tuned && !md.st.retuned) {
            md.st.retuned = true;
            PAT_TRY(md.st.tune(md.st.pos[c], 1.0f, 2, md.rc2, true));
        }
        return PAT_OK;
    }

    std::map<std::tuple<int, int, int>, hipGraphExec_t> seg_graphs;
    int64_t graph_gen = -1;  // layout generation the cached graphs were captured with

    void drop_seg_graphs()
    {
        for (auto &g : seg_graphs) (void)hipGraphExecDestroy(g.second);
        seg_graphs.clear();
    }

    // seg fused force steps starting from buffer c; the run's final step
    // (ends_run) closes with a half kick only.
    int enqueue_segment(int seg, int c, bool ends_run, hipStream_t s, float *kernel_ms)
    {
        for (int j = 0; j < seg; j++) {
            const bool final_step = ends_run && j == seg - 1;
            pat::EventPair ev;
            if (kernel_ms) {
                PAT_HIP(ev.create());
                PAT_HIP(hipEventRecord(ev.a, s));
            }
            PAT_TRY(md.st.csr_force(md.st.pos[c], 1.0f, 2, md.rc2, md.dt, final_step ? kKick : kKickDrift,
                                    md.st.pos[1 - c], s));
            if (kernel_ms) {
                PAT_HIP(hipEventRecord(ev.b, s));
                PAT_HIP(hipEventSynchronize(ev.b));
                (void)hipEventElapsedTime(&kernel_ms[j], ev.a, ev.b);
            }
            if (!final_step) c = 1 - c;
        }
        return PAT_OK;
    }

    // Step k (1-based over the handle's life) rebuilds before its force when
    // k % every == 0.  The force steps between rebuilds replay a captured
    // graph (eager launches with events when kernel times are requested).
    int run(int nsteps, float *kernel_ms)
    {
        hipStream_t s = md.st.stream;
        int c = md.cur;
        PAT_TRY(md.st.begin_steps(md.st.pos[c], 1.0f, 2, md.dt, s));
        int k = 0;
        while (k < nsteps) {
            if ((md.steps + k + 1) % md.every == 0) PAT_TRY(rebuild_all(c));
            int k_end = k + 1;
            while (k_end < nsteps && (md.steps + k_end + 1) % md.every != 0) k_end++;
            const int seg = k_end - k;
            const bool ends_run = k_end == nsteps;
            if (kernel_ms) {
                PAT_TRY(enqueue_segment(seg, c, ends_run, s, kernel_ms + k));
            } else {
                if (graph_gen != md.st.layout_gen) {
                    drop_seg_graphs();
                    graph_gen = md.st.layout_gen;
                }
                auto key = std::make_tuple(seg, c, (int)ends_run);
                auto it = seg_graphs.find(key);
                if (it == seg_graphs.end()) {
                    hipGraph_t g;
                    PAT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    int rc = enqueue_segment(seg, c, ends_run, s, nullptr);
                    hipError_t ec = hipStreamEndCapture(s, &g);
                    if (rc != PAT_OK) return rc;
                    PAT_HIP(ec);
                    hipGraphExec_t ge;
                    hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    PAT_HIP(ei);
                    it = seg_graphs.emplace(key, ge).first;
                }
                PAT_HIP(hipGraphLaunch(it->second, s));
            }
            if ((seg - (ends_run ? 1 : 0)) % 2) c = 1 - c;
            k = k_end;
        }
        PAT_HIP(hipStreamSynchronize(s));
        md.cur = c;
        md.steps += nsteps;
        return PAT_OK;
    }
};

extern "C" {

} // extern "C"

namespace {

// pat_mdb_create for host positions (pos: N x 3) or, with dev, device
// positions (pos: N float4 with each node's box side in w, copied as they
// are); velocities from the host, or zero.
int mdb_create(int nsys, const int64_t *offsets, const float *pos, bool dev, const float *vel, const double *box,
               double cutoff, double skin, int every, double dt, pat_mdb **out)
{
    PAT_REQUIRE(out && offsets && pos && box && nsys > 0, PAT_ERR_INVALID, "pat_mdb_create: null argument");
    PAT_REQUIRE(cutoff > 0 && skin >= 0 && dt > 0 && every > 0, PAT_ERR_INVALID, "pat_mdb_create: bad parameters");
    PAT_REQUIRE(offsets[0] == 0, PAT_ERR_INVALID, "pat_mdb_create: offsets must start at 0");
    for (int s = 0; s < nsys; s++) {
        PAT_REQUIRE(offsets[s + 1] - offsets[s] > 1, PAT_ERR_INVALID, "pat_mdb_create: system %d has < 2 nodes", s);
        PAT_REQUIRE(box[s] >= 3 * (cutoff + skin) * 1.0001, PAT_ERR_UNSUPPORTED,
                    "pat_mdb_create: system %d box %.3f must hold 3 bins of %.3f", s, box[s], cutoff + skin);
    }
    const int64_t N = offsets[nsys];
    PAT_REQUIRE(N < (1LL << 31) - 1, PAT_ERR_INVALID, "pat_mdb_create: too many nodes");
    auto *h = new pat_mdb();
    h->off.assign(offsets, offsets + nsys + 1);
    h->box.assign(box, box + nsys);
    pat_md &m = h->md;
    m.box = 1.0f;
    m.pbc = 2;
    m.rc2 = (float)cutoff * (float)cutoff;
    m.dt = (float)dt;
    m.every = every;
    m.cutneigh = cutoff + skin;
    m.nbw = pat::nb_create();
    int rc = m.st.alloc(N, 0, nullptr, true, false);
    if (rc == PAT_OK && dev) {
        (void)pat::dev_free(m.st.cols);
        m.st.cols = nullptr;
        if (hipMemcpyAsync(m.st.pos[0], pos, sizeof(float) * 4 * N, hipMemcpyDeviceToDevice, m.st.stream) !=
                hipSuccess ||
            hipStreamSynchronize(m.st.stream) != hipSuccess) {
            pat::set_error("pat_mdb_create: position copy failed");
            rc = PAT_ERR_HIP;
        }
    } else if (rc == PAT_OK) {
        (void)pat::dev_free(m.st.cols);
        m.st.cols = nullptr;
        // positions with each node's box side in w
        std::vector<float> h4(4 * (size_t)N);
        for (int s = 0; s < nsys; s++)
            for (int64_t i = offsets[s]; i < offsets[s + 1]; i++) {
                h4[4 * i] = pos[3 * i];
                h4[4 * i + 1] = pos[3 * i + 1];
                h4[4 * i + 2] = pos[3 * i + 2];
                h4[4 * i + 3] = (float)box[s];
            }
        if (hipMemcpy(m.st.pos[0], h4.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice) != hipSuccess) {
            pat::set_error("pat_mdb_create: position upload failed");
            rc = PAT_ERR_HIP;
        }
    }
    if (rc == PAT_OK) {
        std::vector<float> zero(3 * (size_t)N, 0.0f);
        rc = m.st.upload3(m.st.vel, vel ? vel : zero.data(), 12);
    }
    m.st.rebuild_every = every;
    m.st.sysoff_h.assign(offsets, offsets + nsys + 1);
    if (rc == PAT_OK) rc = h->rebuild_all(0);
    if (rc == PAT_OK) rc = m.st.tune(m.st.pos[0], 1.0f, 2, m.rc2);
    if (rc == PAT_OK)
        rc = m.st.force_eval(m.st.pos[0], 1.0f, 2, m.rc2);
    if (rc == PAT_OK && hipStreamSynchronize(m.st.stream) != hipSuccess) rc = PAT_ERR_HIP;
    if (rc != PAT_OK) {
        pat::nb_destroy(m.nbw);
        m.nbw = nullptr;
        m.st.release();
        m.st = IrState{};
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

}  // namespace

namespace pat {

int mdb_create_dev(int nsys, const int64_t *offsets, const float *pos4, const double *box, double cutoff,
                   double skin, int every, double dt, pat_mdb **out)
{
    return mdb_create(nsys, offsets, pos4, true, nullptr, box, cutoff, skin, every, dt, out);
}

int mdb_positions(pat_mdb *h, int sys, const float **pos4, int64_t *n)
{
    PAT_REQUIRE(h && pos4 && n, PAT_ERR_INVALID, "mdb_positions: null argument");
    PAT_REQUIRE(!h->md.failed, PAT_ERR_STATE, "mdb_positions: an earlier run of this handle failed part-way");
    PAT_REQUIRE(sys >= 0 && sys < (int)h->box.size(), PAT_ERR_INVALID, "mdb_positions: system %d out of range", sys);
    *pos4 = (const float *)h->md.st.pos[h->md.cur] + 4 * h->off[sys];
    *n = h->off[sys + 1] - h->off[sys];
    return PAT_OK;
}

int mdb_lanes(pat_mdb *h) { return h ? h->md.st.lpr : 0; }

}  // namespace pat

extern "C" {

int pat_mdb_create(int nsys, const int64_t *offsets, const float *pos, const float *vel, const double *box,
                   double cutoff, double skin, int every, double dt, pat_mdb **out)
{
    return mdb_create(nsys, offsets, pos, false, vel, box, cutoff, skin, every, dt, out);
}

int pat_mdb_run(pat_mdb *h, int nsteps, float *total_ms, float *kernel_ms)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_mdb_run: null handle");
    PAT_REQUIRE(!h->md.failed, PAT_ERR_STATE,
                "pat_mdb_run: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(nsteps >= 0, PAT_ERR_INVALID, "pat_mdb_run: negative step count");
    if (nsteps == 0) return PAT_OK;
    auto t0 = std::chrono::steady_clock::now();
    int rc = h->run(nsteps, kernel_ms);
    if (rc != PAT_OK) h->md.failed = true;
    if (total_ms) *total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int pat_mdb_set_lanes(pat_mdb *h, int lanes_per_row)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_mdb_set_lanes: null handle");
    PAT_TRY(check_variant(PAT_IR_CSR, lanes_per_row));
    PAT_TRY(check_sliced(h->md.st, PAT_IR_CSR, lanes_per_row));
    h->md.st.choice_tuned = false;
    if (lanes_per_row) h->md.st.lpr = lanes_per_row;
    else PAT_TRY(h->md.st.tune(h->md.st.pos[h->md.cur], 1.0f, 2, h->md.rc2));
    h->md.st.layout_gen++;
    h->drop_seg_graphs();
    // the stored force must come from the same summation order
    PAT_TRY(h->md.st.check_box(h->md.st.pos[h->md.cur], 1.0f, 2, h->md.st.stream));
    PAT_TRY(h->md.st.csr_force(h->md.st.pos[h->md.cur], 1.0f, 2, h->md.rc2, 0.f, kForceOnly, nullptr,
                               h->md.st.stream));
    PAT_HIP(hipStreamSynchronize(h->md.st.stream));
    return PAT_OK;
}

int pat_mdb_set_slice_order(pat_mdb *h, int order)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_mdb_set_slice_order: null handle");
    PAT_REQUIRE(order >= 0 && order <= 2, PAT_ERR_INVALID, "pat_mdb_set_slice_order: order %d not in 0..2", order);
    IrState &st = h->md.st;
    PAT_REQUIRE(st.sell_ok && st.sentinel(), PAT_ERR_UNSUPPORTED,
                "pat_mdb_set_slice_order: needs the sliced sentinel layout");
    st.choice_tuned = false;
    PAT_TRY(st.set_layout(order, st.want_halo, st.pos[h->md.cur], 1.0f, 2, h->md.rc2));
    st.layout_gen++;
    h->drop_seg_graphs();
    PAT_TRY(st.check_box(st.pos[h->md.cur], 1.0f, 2, st.stream));
    PAT_TRY(st.csr_force(st.pos[h->md.cur], 1.0f, 2, h->md.rc2, 0.f, kForceOnly, nullptr, st.stream));
    PAT_HIP(hipStreamSynchronize(st.stream));
    return PAT_OK;
}

int pat_mdb_get(pat_mdb *h, int sys, float *pos, float *vel, float *force)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_mdb_get: null handle");
    PAT_REQUIRE(!h->md.failed, PAT_ERR_STATE,
                "pat_mdb_get: an earlier run of this handle failed part-way; its state is undefined");
    PAT_REQUIRE(sys >= 0 && sys < (int)h->box.size(), PAT_ERR_INVALID, "pat_mdb_get: system %d out of range", sys);
    const int64_t o = h->off[sys], n = h->off[sys + 1] - o;
    std::vector<float> buf(4 * (size_t)n);
    auto get = [&](const float *src4, float *dst) -> int {
        PAT_HIP(hipMemcpy(buf.data(), src4 + 4 * o, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) dst[3 * i + c] = buf[4 * i + c];
        return PAT_OK;
    };
    PAT_HIP(hipStreamSynchronize(h->md.st.stream));
    if (pos) PAT_TRY(get(h->md.st.pos[h->md.cur], pos));
    if (vel) PAT_TRY(get(h->md.st.vel, vel));
    if (force) PAT_TRY(get(h->md.st.force, force));
    return PAT_OK;
}

int pat_mdb_stats(pat_mdb *h, int64_t *num_pairs, int64_t *steps, int64_t *rebuilds, float *rebuild_ms)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_mdb_stats: null handle");
    return pat_md_stats(&h->md, num_pairs, steps, rebuilds, rebuild_ms);
}

int pat_mdb_destroy(pat_mdb *h)
{
    if (!h) return PAT_OK;
    if (h->md.nbw) pat::nb_destroy(h->md.nbw);
    h->md.nbw = nullptr;
    h->drop_seg_graphs();
    h->md.drop_graphs();
    h->md.st.release();
    h->md.st = IrState{};
    delete h;
    return PAT_OK;
}

}  // extern "C"
