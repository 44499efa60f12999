// Handle-level generalized reduction with additive functors: the
// RegularRuntime work beside kmeans (kmeans/regular/regular_runtime.h:79-105
// with map/reduce index 0, kmeans/regular.h:30-73; miniMD/regular.h:139-181
// map2/reduce2).  Units and parameter are copied to the device once; a run
// is one map+reduce pass over every unit.
#include <cstring>
#include <vector>

#include "pat_common.hpp"

namespace pat {
int gr_run(int functor, const void *units, int64_t n, const void *param, const double *aux, double *part,
           unsigned long long *cnt, double *out, unsigned long long *outn, hipStream_t st);
}

namespace {
constexpr int kMaxBlocks = 4096;
constexpr int kMaxKV = 8;

struct Shape {
    int64_t unit_bytes;  // bytes per unit
    int keys, values;    // K, V of the reduction object
};

bool shape_of(int map_idx, Shape *s)
{
    if (map_idx == PAT_GR_VSTATS) *s = Shape{24, 1, 2};
    else if (map_idx == PAT_GR_LJ_ENERGY) *s = Shape{8, 1, 1};
    else return false;
    return true;
}
}  // namespace

struct pat_gr {
    int map_idx = -1, reduce_idx = -1;
    int64_t n = 0, unit_bytes = 0;
    void *units = nullptr;
    std::vector<unsigned char> param;  // the functor's scalar parameter (host copy)
    double *aux = nullptr;             // array part of the parameter (pair energy: positions)
    int64_t aux_len = 0;
    double *part = nullptr, *out = nullptr;
    unsigned long long *cnt = nullptr, *outn = nullptr;
    hipStream_t stream = nullptr;
    bool have_result = false;
    std::vector<int32_t> host_edges;  // pair energy: kept to validate node indices against a new parameter

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {units, (void *)aux, (void *)part, (void *)out, (void *)cnt, (void *)outn})
            if (p) (void)pat::dev_free(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Checks and uploads the parameter for the current map index.
    int set_param(const void *p, int64_t bytes)
    {
        PAT_REQUIRE(p || bytes == 0, PAT_ERR_INVALID, "pat_gr: null parameter");
        if (map_idx == PAT_GR_VSTATS) {
            PAT_REQUIRE(bytes == (int64_t)sizeof(float), PAT_ERR_INVALID,
                        "pat_gr: the velocity statistics take one float (vaverh), got %lld bytes", (long long)bytes);
            param.assign((const unsigned char *)p, (const unsigned char *)p + bytes);
        } else if (map_idx == PAT_GR_LJ_ENERGY) {
            PAT_REQUIRE(bytes >= (int64_t)sizeof(double) && (bytes - (int64_t)sizeof(double)) % 24 == 0,
                        PAT_ERR_INVALID, "pat_gr: the pair energy takes {double cutsq, double x[3N]}, got %lld bytes",
                        (long long)bytes);
            const int64_t nodes = (bytes - (int64_t)sizeof(double)) / 24;
            for (size_t e = 0; e < host_edges.size(); e++)
                PAT_REQUIRE(host_edges[e] >= 0 && host_edges[e] < nodes, PAT_ERR_INVALID,
                            "pat_gr: edge %zu references node %d outside the %lld parameter positions", e / 2,
                            host_edges[e], (long long)nodes);
            param.assign((const unsigned char *)p, (const unsigned char *)p + sizeof(double));
            if (3 * nodes > aux_len) {
                if (aux) (void)pat::dev_free(aux);
                aux = nullptr;
                aux_len = 0;
                if (pat::dev_malloc(&aux, sizeof(double) * (size_t)std::max<int64_t>(3 * nodes, 1)) != hipSuccess) {
                    (void)hipGetLastError();
                    pat::set_error("pat_gr: hipMalloc of the positions failed");
                    return PAT_ERR_NOMEM;
                }
                aux_len = 3 * nodes;
            }
            if (nodes > 0)
                PAT_HIP(hipMemcpyAsync(aux, (const char *)p + sizeof(double), sizeof(double) * 3 * nodes,
                                       hipMemcpyHostToDevice, stream));
            PAT_HIP(hipStreamSynchronize(stream));
        } else {
            PAT_REQUIRE(false, PAT_ERR_STATE, "pat_gr: set the map index before the parameter");
        }
        have_result = false;
        return PAT_OK;
    }
};

extern "C" {

int pat_gr_create(const void *units, int64_t num_units, int64_t unit_bytes, int map_idx, const void *param,
                  int64_t param_bytes, pat_gr **out)
{
    PAT_REQUIRE(out && units, PAT_ERR_INVALID, "pat_gr_create: null argument");
    PAT_REQUIRE(num_units > 0 && num_units < (1LL << 40), PAT_ERR_INVALID, "pat_gr_create: bad unit count");
    Shape sh;
    PAT_REQUIRE(shape_of(map_idx, &sh), PAT_ERR_UNSUPPORTED,
                "pat_gr_create: map index %d has no additive engine (0 = velocity statistics, 2 = pair energy; "
                "kmeans map1 runs on pat_kmeans_*)",
                map_idx);
    PAT_REQUIRE(unit_bytes == sh.unit_bytes, PAT_ERR_INVALID, "pat_gr_create: map %d takes %lld-byte units, got %lld",
                map_idx, (long long)sh.unit_bytes, (long long)unit_bytes);
    auto *h = new pat_gr();
    h->map_idx = h->reduce_idx = map_idx;
    h->n = num_units;
    h->unit_bytes = unit_bytes;
    int rc = PAT_OK;
    if (map_idx == PAT_GR_LJ_ENERGY) {
        const int32_t *e = (const int32_t *)units;
        h->host_edges.assign(e, e + 2 * num_units);
        for (int64_t k = 0; k < num_units && rc == PAT_OK; k++)
            if (e[2 * k] < 0 || e[2 * k + 1] < 0) {
                pat::set_error("pat_gr_create: edge %lld has a negative node index", (long long)k);
                rc = PAT_ERR_INVALID;
            }
    }
    auto am = [&](void **p, size_t b) {
        if (rc != PAT_OK) return;
        if (pat::dev_malloc(p, b) != hipSuccess) {
            (void)hipGetLastError();
            pat::set_error("pat_gr_create: hipMalloc of %zu bytes failed", b);
            rc = PAT_ERR_NOMEM;
        }
    };
    if (rc == PAT_OK && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) rc = PAT_ERR_HIP;
    am(&h->units, (size_t)(num_units * unit_bytes));
    am((void **)&h->part, sizeof(double) * kMaxBlocks * kMaxKV);
    am((void **)&h->cnt, sizeof(unsigned long long) * kMaxBlocks * kMaxKV);
    am((void **)&h->out, sizeof(double) * kMaxKV);
    am((void **)&h->outn, sizeof(unsigned long long) * kMaxKV);
    if (rc == PAT_OK &&
        hipMemcpyAsync(h->units, units, (size_t)(num_units * unit_bytes), hipMemcpyHostToDevice, h->stream) !=
            hipSuccess)
        rc = PAT_ERR_HIP;
    if (rc == PAT_OK) rc = h->set_param(param, param_bytes);
    if (rc != PAT_OK) {
        if (rc == PAT_ERR_HIP) pat::set_error("pat_gr_create: HIP call failed");
        h->release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

int pat_gr_set_map_idx(pat_gr *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_gr_set_map_idx: null handle");
    PAT_REQUIRE(idx == h->map_idx, PAT_ERR_UNSUPPORTED,
                "pat_gr_set_map_idx: the handle's units are laid out for map %d (create a handle per map)",
                h->map_idx);
    return PAT_OK;
}

int pat_gr_set_reduce_idx(pat_gr *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_gr_set_reduce_idx: null handle");
    PAT_REQUIRE(idx == h->map_idx, PAT_ERR_UNSUPPORTED,
                "pat_gr_set_reduce_idx: map %d pairs with reduce %d (adds its values), not %d", h->map_idx,
                h->map_idx, idx);
    h->reduce_idx = idx;
    return PAT_OK;
}

int pat_gr_set_param(pat_gr *h, const void *param, int64_t param_bytes)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_gr_set_param: null handle");
    return h->set_param(param, param_bytes);
}

int pat_gr_run(pat_gr *h)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_gr_run: null handle");
    PAT_TRY(pat::gr_run(h->map_idx, h->units, h->n, h->param.data(), h->aux, h->part, h->cnt, h->out, h->outn,
                        h->stream));
    PAT_HIP(hipStreamSynchronize(h->stream));
    h->have_result = true;
    return PAT_OK;
}

int pat_gr_get_output(pat_gr *h, int *num_keys, int *num_values, double *values, int64_t *emits)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_gr_get_output: null handle");
    PAT_REQUIRE(h->have_result, PAT_ERR_STATE, "pat_gr_get_output: no result (call pat_gr_run)");
    Shape sh{0, 0, 0};
    PAT_REQUIRE(shape_of(h->map_idx, &sh), PAT_ERR_STATE, "pat_gr_get_output: no engine for map %d", h->map_idx);
    if (num_keys) *num_keys = sh.keys;
    if (num_values) *num_values = sh.values;
    if (values)
        PAT_HIP(hipMemcpy(values, h->out, sizeof(double) * sh.keys * sh.values, hipMemcpyDeviceToHost));
    if (emits) {
        std::vector<unsigned long long> c((size_t)sh.keys);
        PAT_HIP(hipMemcpy(c.data(), h->outn, sizeof(unsigned long long) * sh.keys, hipMemcpyDeviceToHost));
        for (int k = 0; k < sh.keys; k++) emits[k] = (int64_t)c[k];
    }
    return PAT_OK;
}

int pat_gr_destroy(pat_gr *h)
{
    if (!h) return PAT_OK;
    h->release();
    delete h;
    return PAT_OK;
}

}  // extern "C"
