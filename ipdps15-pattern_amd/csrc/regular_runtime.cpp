// Handle-level generalized reduction with an application's own functors and
// keyed reduction objects: RegularRuntime (kmeans/regular/regular_runtime.h:
// 79-105, regular_runtime.cu:31-291) for any map_fp / reduce_fp pair.  The
// input and its unit offsets are copied to the device once; each start
// sizes the wavefront tables and the final table, runs the application
// table's kernels (api/gpu/regular_table.hpp) and reads the final table back,
// its records sorted by key bytes.
//
// The reference's RegularRuntime splits units over the CPU threads and GPUs
// by a speed profile, feeds the GPU from mapped host memory and merges the
// device objects on the host (regular_runtime.cu:98-286; its merge_device is
// never called, regular_runtime.cu:285); with one GPU holding the input,
// one start is one device pass and the object comes back whole.
#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "pat_common.hpp"

namespace {

constexpr int64_t kDefaultKeys = 4096;
constexpr int64_t kDefaultRecordBytes = 256;
constexpr int kMaxWaves = 1024;
constexpr int64_t kUnitsPerWave = 4096;

int64_t pow2_at_least(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

struct pat_regular {
    void *input = nullptr;
    int64_t input_bytes = 0, num_units = 0, unit_bytes = 0;
    int64_t *offsets = nullptr;
    const pat_gr_table *table = nullptr;
    int map_idx = 0, reduce_idx = 0;
    void *param = nullptr;
    int64_t param_bytes = 0;
    int64_t max_keys = kDefaultKeys, max_record_bytes = kDefaultRecordBytes;
    int lds_keys = 0;
    uint32_t *tables = nullptr;  // wave tables then the final table
    int64_t tables_words = 0;
    int32_t *status = nullptr;
    hipStream_t stream = nullptr;
    // last start
    bool have_result = false;
    std::vector<std::pair<std::string, std::string>> out;
    int64_t merges = 0, reruns = 0;
    int waves = 0;

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {input, (void *)offsets, param, (void *)tables, (void *)status})
            if (p) (void)pat::dev_free(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Table geometry: buckets (a power of two, twice the keys) and pool words
    // (a header, key and value per record).
    int64_t buckets() const { return pow2_at_least(2 * max_keys); }
    int64_t pool_words() const { return max_keys * (1 + (max_record_bytes + 3) / 4 + 1); }

    int start()
    {
        PAT_REQUIRE(table, PAT_ERR_STATE, "pat_regular_start: no application table (pat_regular_set_table)");
        const int64_t nb = buckets(), pw = pool_words();
        PAT_REQUIRE(nb <= (1 << 24) && pw < (1LL << 31), PAT_ERR_INVALID, "pat_regular_start: object too large");
        waves = (int)std::min<int64_t>(kMaxWaves, std::max<int64_t>(1, (num_units + kUnitsPerWave - 1) / kUnitsPerWave));
        const int64_t words = (nb + 1 + pw) * (waves + 1);
        if (words > tables_words) {
            if (tables) (void)pat::dev_free(tables);
            tables = nullptr;
            tables_words = 0;
            if (pat::dev_malloc(&tables, sizeof(uint32_t) * (size_t)words) != hipSuccess) {
                (void)hipGetLastError();
                tables = nullptr;
                pat::set_error("pat_regular_start: hipMalloc of %lld table bytes failed", (long long)(4 * words));
                return PAT_ERR_NOMEM;
            }
            tables_words = words;
        }
        // LDS table: buckets a power of two covering lds_keys (or the
        // kernels' largest), pool words for that many records
        int lnb = table->max_lds_buckets, lpw = table->max_lds_pool_words;
        if (lds_keys > 0) {
            lnb = (int)std::min<int64_t>(lnb, pow2_at_least(lds_keys));
            lpw = (int)std::min<int64_t>(lpw, (int64_t)lds_keys * (2 + (max_record_bytes + 3) / 4));
        }
        pat_gr_object o;
        o.data = input;
        o.offsets = offsets;
        o.unit_bytes = unit_bytes;
        o.num_units = num_units;
        o.param = param;
        o.wave_tables = tables;
        o.final_table = tables + (nb + 1 + pw) * waves;
        o.status = status;
        o.waves = waves;
        o.table_buckets = (int32_t)nb;
        o.table_pool_words = (int32_t)pw;
        o.lds_buckets = lnb;
        o.lds_pool_words = lpw;
        PAT_HIP(hipMemsetAsync(status, 0, 4 * sizeof(int32_t), stream));
        const int rc = table->map_reduce(map_idx, reduce_idx, &o, stream);
        PAT_REQUIRE(rc == PAT_OK, rc, "pat_regular_start: the table's kernels for map %d / reduce %d failed (%d)",
                    map_idx, reduce_idx, rc);
        int32_t st[4] = {0, 0, 0, 0};
        PAT_HIP(hipMemcpyAsync(st, status, sizeof(st), hipMemcpyDeviceToHost, stream));
        std::vector<uint32_t> fin((size_t)(nb + 1 + pw));
        PAT_HIP(hipMemcpyAsync(fin.data(), o.final_table, sizeof(uint32_t) * fin.size(), hipMemcpyDeviceToHost,
                               stream));
        PAT_HIP(hipStreamSynchronize(stream));
        PAT_REQUIRE(!(st[0] & 1), PAT_ERR_UNSUPPORTED,
                    "pat_regular_start: a wavefront met more keys than the object holds (%lld keys of %lld bytes; "
                    "pat_regular_set_capacity)", (long long)max_keys, (long long)max_record_bytes);
        PAT_REQUIRE(!(st[0] & 2), PAT_ERR_UNSUPPORTED,
                    "pat_regular_start: one unit's values do not fit an empty LDS table");
        PAT_REQUIRE(!(st[0] & 4), PAT_ERR_UNSUPPORTED,
                    "pat_regular_start: a map's insert was refused after another of its inserts was taken; it "
                    "cannot run again without inserting twice (a larger LDS table avoids the refusal)");
        PAT_REQUIRE(!(st[0] & 8), PAT_ERR_UNSUPPORTED,
                    "pat_regular_start: the final object is full (%lld keys of %lld bytes; pat_regular_set_capacity)",
                    (long long)max_keys, (long long)max_record_bytes);
        merges = st[1];
        reruns = st[2];
        // decode the final table's records, sorted by key bytes
        out.clear();
        for (int64_t b = 0; b < nb; b++) {
            const uint32_t s = fin[(size_t)b];
            if (s == 0u) continue;
            PAT_REQUIRE(s != 0xffffffffu && (int64_t)s - 1 < pw, PAT_ERR_INVALID,
                        "pat_regular_start: damaged final table");
            const uint32_t *r = fin.data() + nb + 1 + (s - 1u);
            const int ks = (int)(r[0] >> 16), vs = (int)(r[0] & 0xffffu);
            PAT_REQUIRE((int64_t)s - 1 + 1 + (ks + 3) / 4 + (vs + 3) / 4 <= pw, PAT_ERR_INVALID,
                        "pat_regular_start: damaged final table record");
            const char *kp = (const char *)(r + 1);
            const char *vp = (const char *)(r + 1 + (ks + 3) / 4);
            out.emplace_back(std::string(kp, (size_t)ks), std::string(vp, (size_t)vs));
        }
        std::sort(out.begin(), out.end());
        have_result = true;
        return PAT_OK;
    }
};

extern "C" {

int pat_regular_create(const void *input, int64_t input_bytes, int64_t num_units, int64_t unit_bytes,
                       const int64_t *offsets, pat_regular **out)
{
    PAT_REQUIRE(out && input && input_bytes > 0 && num_units > 0 && unit_bytes > 0, PAT_ERR_INVALID,
                "pat_regular_create: bad arguments");
    PAT_REQUIRE(unit_bytes <= input_bytes, PAT_ERR_INVALID, "pat_regular_create: units larger than the input");
    if (offsets) {
        for (int64_t i = 0; i < num_units; i++)
            PAT_REQUIRE(offsets[i] >= 0 && offsets[i] <= input_bytes - unit_bytes, PAT_ERR_INVALID,
                        "pat_regular_create: unit %lld at %lld + %lld bytes lies outside the %lld-byte input",
                        (long long)i, (long long)offsets[i], (long long)unit_bytes, (long long)input_bytes);
    } else {
        PAT_REQUIRE(num_units <= input_bytes / unit_bytes, PAT_ERR_INVALID,
                    "pat_regular_create: %lld units of %lld bytes exceed the %lld-byte input", (long long)num_units,
                    (long long)unit_bytes, (long long)input_bytes);
    }
    auto *h = new pat_regular();
    h->input_bytes = input_bytes;
    h->num_units = num_units;
    h->unit_bytes = unit_bytes;
    int rc = PAT_OK;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->stream = nullptr;
        pat::set_error("pat_regular_create: stream creation failed");
        rc = PAT_ERR_HIP;
    }
    auto am = [&](void **p, size_t b) {
        if (rc == PAT_OK && pat::dev_malloc(p, b) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            pat::set_error("pat_regular_create: hipMalloc of %zu bytes failed", b);
            rc = PAT_ERR_NOMEM;
        }
    };
    am(&h->input, (size_t)input_bytes);
    am((void **)&h->status, 4 * sizeof(int32_t));
    if (offsets) am((void **)&h->offsets, sizeof(int64_t) * (size_t)num_units);
    if (rc == PAT_OK &&
        (hipMemcpyAsync(h->input, input, (size_t)input_bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
         (offsets && hipMemcpyAsync(h->offsets, offsets, sizeof(int64_t) * (size_t)num_units, hipMemcpyHostToDevice,
                                    h->stream) != hipSuccess) ||
         hipStreamSynchronize(h->stream) != hipSuccess)) {
        pat::set_error("pat_regular_create: copying the input failed");
        rc = PAT_ERR_HIP;
    }
    if (rc != PAT_OK) {
        h->release();
        delete h;
        return rc;
    }
    *out = h;
    return PAT_OK;
}

int pat_regular_set_table(pat_regular *h, const pat_gr_table *table)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_set_table: null handle");
    PAT_REQUIRE(table && table->map_count > 0 && table->reduce_count > 0 && table->map_reduce &&
                    table->max_lds_buckets > 0 && table->max_lds_pool_words > 0,
                PAT_ERR_INVALID, "pat_regular_set_table: bad table");
    h->table = table;
    h->map_idx = h->reduce_idx = 0;
    h->have_result = false;
    return PAT_OK;
}

int pat_regular_set_capacity(pat_regular *h, int64_t max_keys, int64_t max_record_bytes, int lds_keys)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_set_capacity: null handle");
    PAT_REQUIRE(max_keys > 0 && max_keys <= (1 << 22) && max_record_bytes > 0 && max_record_bytes <= 65535 &&
                    lds_keys >= 0,
                PAT_ERR_INVALID, "pat_regular_set_capacity: out of range");
    h->max_keys = max_keys;
    h->max_record_bytes = max_record_bytes;
    h->lds_keys = lds_keys;
    return PAT_OK;
}

int pat_regular_set_map_idx(pat_regular *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_set_map_idx: null handle");
    PAT_REQUIRE(!h->table || (idx >= 0 && idx < h->table->map_count), PAT_ERR_UNSUPPORTED,
                "pat_regular_set_map_idx: map index %d not in the table", idx);
    h->map_idx = idx;
    return PAT_OK;
}

int pat_regular_set_reduce_idx(pat_regular *h, int idx)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_set_reduce_idx: null handle");
    PAT_REQUIRE(!h->table || (idx >= 0 && idx < h->table->reduce_count), PAT_ERR_UNSUPPORTED,
                "pat_regular_set_reduce_idx: reduce index %d not in the table", idx);
    h->reduce_idx = idx;
    return PAT_OK;
}

int pat_regular_set_param(pat_regular *h, const void *param, int64_t param_bytes)
{
    PAT_REQUIRE(h && param_bytes >= 0 && (param || param_bytes == 0), PAT_ERR_INVALID,
                "pat_regular_set_param: bad argument");
    if (param_bytes != h->param_bytes) {
        if (h->param) (void)pat::dev_free(h->param);
        h->param = nullptr;
        h->param_bytes = 0;
        if (param_bytes > 0) {
            if (pat::dev_malloc(&h->param, (size_t)param_bytes) != hipSuccess) {
                (void)hipGetLastError();
                h->param = nullptr;
                pat::set_error("pat_regular_set_param: hipMalloc of %lld bytes failed", (long long)param_bytes);
                return PAT_ERR_NOMEM;
            }
        }
        h->param_bytes = param_bytes;
    }
    if (param_bytes > 0) {
        PAT_HIP(hipMemcpyAsync(h->param, param, (size_t)param_bytes, hipMemcpyHostToDevice, h->stream));
        PAT_HIP(hipStreamSynchronize(h->stream));
    }
    h->have_result = false;
    return PAT_OK;
}

int pat_regular_start(pat_regular *h)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_start: null handle");
    h->have_result = false;
    return h->start();
}

int pat_regular_get_output(pat_regular *h, int64_t *num_keys, int64_t *key_bytes, int64_t *val_bytes, void *keys,
                           uint32_t *key_index, uint32_t *key_size, void *vals, uint32_t *val_index,
                           uint32_t *val_size)
{
    PAT_REQUIRE(h && num_keys, PAT_ERR_INVALID, "pat_regular_get_output: null argument");
    PAT_REQUIRE(h->have_result, PAT_ERR_STATE, "pat_regular_get_output: no result (pat_regular_start)");
    int64_t kb = 0, vb = 0;
    for (const auto &e : h->out) {
        kb += (int64_t)e.first.size();
        vb += (int64_t)e.second.size();
    }
    *num_keys = (int64_t)h->out.size();
    if (key_bytes) *key_bytes = kb;
    if (val_bytes) *val_bytes = vb;
    int64_t ko = 0, vo = 0, i = 0;
    for (const auto &e : h->out) {
        if (keys) memcpy((char *)keys + ko, e.first.data(), e.first.size());
        if (vals) memcpy((char *)vals + vo, e.second.data(), e.second.size());
        if (key_index) key_index[i] = (uint32_t)ko;
        if (val_index) val_index[i] = (uint32_t)vo;
        if (key_size) key_size[i] = (uint32_t)e.first.size();
        if (val_size) val_size[i] = (uint32_t)e.second.size();
        ko += (int64_t)e.first.size();
        vo += (int64_t)e.second.size();
        i++;
    }
    return PAT_OK;
}

int pat_regular_stats(pat_regular *h, int64_t *merges, int64_t *reruns, int *waves)
{
    PAT_REQUIRE(h, PAT_ERR_INVALID, "pat_regular_stats: null handle");
    if (merges) *merges = h->merges;
    if (reruns) *reruns = h->reruns;
    if (waves) *waves = h->waves;
    return PAT_OK;
}

int pat_regular_destroy(pat_regular *h)
{
    if (!h) return PAT_OK;
    h->release();
    delete h;
    return PAT_OK;
}

}  // extern "C"
