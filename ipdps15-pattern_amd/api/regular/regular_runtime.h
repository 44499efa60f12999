// RegularRuntime (kmeans/regular/regular_runtime.h:11-106): the
// generalized-reduction runtime.  Same constructor and calls; the map/reduce
// functor pair selected by set_map_idx/set_reduce_idx runs on the GPU through
// libpattern.  Functor table (kmeans/regular.h:143-145, miniMD/regular.h):
//   index 0 = map/reduce (kmeans/regular.h:30-73): units are velocities
//     (double[3]), the parameter is float vaverh; output: one key, the value
//     {sum |v|, count of |v| > vaverh} as float[2] (reduce's layout);
//   index 1 = map1/reduce1 (nearest centre, emit {x, 1, min_dist}): the
//     parameter is the K x D centre array; output per key present
//     {sum x_0..x_{D-1}, count, sum min_dist} as floats (kmeans/regular.h:
//     121-136);
//   index 2 = map2/reduce2 (miniMD/regular.h:139-181): units are edges
//     (int[2]), the parameter is {double cutsq, double x[3N]}; output: one
//     key, the value {sum phi} as double[1].
// The reference spreads indices 0 and 2 over ANY_KEY buckets; here they are
// one key (the buckets' sum).  RegularStart() is one map+reduce pass over
// every unit.
//
// An application that links its functor tables compiled for the device
// (api/gpu/regular_table.hpp, which defines pat_app_gr_table) runs its own
// map/reduce pairs instead, any index, with keyed reduction objects
// (pat_regular_*): the units are the input bytes at the Offset array's
// offsets as given, and RegularGetOutput returns every key the object holds
// with its value, sorted by key bytes.
#ifndef REGULAR_RUNTIME_H_
#define REGULAR_RUNTIME_H_

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../pat_check.h"
#include "data_type.h"

// Defined by an application's device functor tables when they are linked in.
extern "C" const pat_gr_table *pat_app_gr_table(void) __attribute__((weak));

class RegularRuntime {
  public:
    int num_procs_;
    size_t num_offsets_;
    size_t input_size_;
    int parameter_size_;
    int map_idx_;
    int reduce_idx_;
    void *input_;
    Offset *offsets_;
    void *parameter_;

    RegularRuntime(int num_procs, void *input, size_t input_size, Offset *offsets, size_t offset_number,
                   void *parameters, int parameter_size)
        : num_procs_(num_procs), num_offsets_(offset_number), input_size_(input_size), parameter_size_(parameter_size),
          map_idx_(0), reduce_idx_(0), input_(input), offsets_(offsets), parameter_(parameters), handle_(nullptr),
          dim_(0), k_(0)
    {
        out_ = output{};
    }

    ~RegularRuntime()
    {
        if (handle_) pat_kmeans_destroy(handle_);
        if (gr_) pat_gr_destroy(gr_);
        if (reg_) pat_regular_destroy(reg_);
    }

    // Validates the units and uploads them.  Units of equal size laid out
    // back to back are used in place; any other layout is gathered first.
    void RegularInit()
    {
        if (num_offsets_ == 0 || offsets_ == nullptr) fail("RegularInit: no units");
        if (pat_app_gr_table) {
            // the units as the Offset array places them, on the device
            std::vector<int64_t> off(num_offsets_);
            int64_t usz = 0;
            for (size_t i = 0; i < num_offsets_; i++) {
                off[i] = (int64_t)offsets_[i].offset;
                usz = std::max<int64_t>(usz, (int64_t)offsets_[i].size);
            }
            PAT_API_CALL(pat_regular_create(input_, (int64_t)input_size_, (int64_t)num_offsets_, usz, off.data(),
                                            &reg_));
            PAT_API_CALL(pat_regular_set_table(reg_, pat_app_gr_table()));
            return;
        }
        const size_t usz = offsets_[0].size;
        if (usz == 0 || usz % sizeof(float)) fail("RegularInit: unit size must be a whole number of floats");
        for (size_t i = 0; i < num_offsets_; i++)
            if (offsets_[i].size != usz || offsets_[i].offset + usz > input_size_)
                fail("RegularInit: units must have equal sizes inside the input");
        dim_ = (int)(usz / sizeof(float));
        bool packed = true;
        for (size_t i = 0; i < num_offsets_ && packed; i++) packed = offsets_[i].offset == i * usz;
        if (!packed) {
            packed_.resize(num_offsets_ * dim_);
            for (size_t i = 0; i < num_offsets_; i++)
                memcpy(&packed_[i * dim_], (const char *)input_ + offsets_[i].offset, usz);
        }
        src_ = packed ? (const float *)input_ : packed_.data();
    }

    void set_map_idx(int idx) { map_idx_ = idx; }
    void set_reduce_idx(int idx) { reduce_idx_ = idx; }

    void RegularStart()
    {
        if (reg_) {
            PAT_API_CALL(pat_regular_set_map_idx(reg_, map_idx_));
            PAT_API_CALL(pat_regular_set_reduce_idx(reg_, reduce_idx_));
            PAT_API_CALL(pat_regular_set_param(reg_, parameter_, parameter_size_));
            PAT_API_CALL(pat_regular_start(reg_));
            return;
        }
        if (map_idx_ != reduce_idx_ || map_idx_ < 0 || map_idx_ > 2)
            fail("RegularStart: functor pairs 0/0, 1/1 and 2/2 have GPU engines");
        if (map_idx_ != 1) {
            start_additive();
            return;
        }
        if (parameter_size_ <= 0 || parameter_size_ % (int)(sizeof(float) * dim_))
            fail("RegularStart: the parameter must be K x D floats");
        const int k = parameter_size_ / (int)(sizeof(float) * dim_);
        if (!handle_ || k != k_) {
            if (handle_) pat_kmeans_destroy(handle_);
            const char *m = getenv("PAT_KMEANS_MODE");
            const int mode = m && atoi(m) == 0 ? PAT_KM_PLAIN : PAT_KM_MFMA;
            PAT_API_CALL(pat_kmeans_create(src_, (int64_t)num_offsets_, dim_, k, mode, 0, &handle_));
            k_ = k;
        }
        PAT_API_CALL(pat_kmeans_set_centroids(handle_, (const float *)parameter_));
        PAT_API_CALL(pat_kmeans_run(handle_, 1));
    }

    struct output RegularGetOutput()
    {
        if (reg_) return keyed_output();
        if (gr_) return additive_output();
        if (!handle_) fail("RegularGetOutput: RegularStart has not run");
        std::vector<int64_t> sums((size_t)k_ * dim_), counts(k_), dsums(k_);
        int s = 0, ds = 0;
        PAT_API_CALL(pat_kmeans_get_object(handle_, sums.data(), counts.data(), dsums.data(), &s, &ds));
        keys_.clear();
        vals_.clear();
        for (int key = 0; key < k_; key++) {
            if (counts[key] == 0) continue;
            keys_.push_back((unsigned int)key);
            for (int d = 0; d < dim_; d++) vals_.push_back((float)std::ldexp((double)sums[(size_t)key * dim_ + d], -s));
            vals_.push_back((float)counts[key]);
            vals_.push_back((float)std::ldexp((double)dsums[key], -ds));
        }
        kidx_.resize(keys_.size());
        vidx_.resize(keys_.size());
        for (size_t i = 0; i < keys_.size(); i++) {
            kidx_[i] = (unsigned int)(i * sizeof(unsigned int));
            vidx_[i] = (unsigned int)(i * (dim_ + 2) * sizeof(float));
        }
        out_.output_keys = (char *)keys_.data();
        out_.output_vals = (char *)vals_.data();
        out_.key_index = kidx_.data();
        out_.val_index = vidx_.data();
        out_.key_num = (unsigned int)keys_.size();
        out_.key_size = sizeof(unsigned int);
        out_.value_size = (unsigned int)((dim_ + 2) * sizeof(float));
        return out_;
    }

    // Capacity of the keyed reduction objects (with device tables; not in the
    // reference API, whose objects have compile-time sizes): distinct keys,
    // bytes of key + value, keys an LDS object takes before it is merged.
    void set_object_capacity(int64_t max_keys, int64_t record_bytes, int lds_keys)
    {
        if (!reg_) fail("set_object_capacity: RegularInit with a device table first");
        PAT_API_CALL(pat_regular_set_capacity(reg_, max_keys, record_bytes, lds_keys));
    }

    // Assignment of every unit from the last pass (not in the reference API;
    // the kmeans driver uses it to report cluster sizes).
    void get_keys(int *keys)
    {
        if (!handle_) fail("get_keys: RegularStart has not run");
        PAT_API_CALL(pat_kmeans_get(handle_, nullptr, keys, nullptr));
    }

  private:
    // Indices 0 and 2: the units back to back (unit = the map's record), the
    // parameter as given; one handle per runtime, the parameter refreshed on
    // every start.
    void start_additive()
    {
        const size_t usz = map_idx_ == 0 ? 3 * sizeof(double) : 2 * sizeof(int);
        for (size_t i = 0; i < num_offsets_; i++)
            if (offsets_[i].size != usz || offsets_[i].offset != i * usz)
                fail("RegularStart: map 0 / 2 units must be back-to-back double[3] / int[2] records");
        if (!gr_)
            PAT_API_CALL(pat_gr_create(input_, (int64_t)num_offsets_, (int64_t)usz, map_idx_, parameter_,
                                       parameter_size_, &gr_));
        else
            PAT_API_CALL(pat_gr_set_param(gr_, parameter_, parameter_size_));
        PAT_API_CALL(pat_gr_run(gr_));
    }

    struct output keyed_output()
    {
        int64_t nk = 0, kb = 0, vb = 0;
        PAT_API_CALL(pat_regular_get_output(reg_, &nk, &kb, &vb, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
        kbytes_.resize((size_t)std::max<int64_t>(kb, 1));
        vbytes_.resize((size_t)std::max<int64_t>(vb, 1));
        kidx_.resize((size_t)nk);
        vidx_.resize((size_t)nk);
        std::vector<uint32_t> ksz((size_t)nk), vsz((size_t)nk);
        PAT_API_CALL(pat_regular_get_output(reg_, &nk, &kb, &vb, kbytes_.data(), kidx_.data(), ksz.data(),
                                            vbytes_.data(), vidx_.data(), vsz.data()));
        out_.output_keys = kbytes_.data();
        out_.output_vals = vbytes_.data();
        out_.key_index = kidx_.data();
        out_.val_index = vidx_.data();
        out_.key_num = (unsigned int)nk;
        out_.key_size = nk ? ksz[0] : 0;
        out_.value_size = nk ? vsz[0] : 0;
        return out_;
    }

    struct output additive_output()
    {
        double vals[2] = {0.0, 0.0};
        int nk = 0, nv = 0;
        PAT_API_CALL(pat_gr_get_output(gr_, &nk, &nv, vals, nullptr));
        keys_.assign(1, 0u);
        kidx_.assign(1, 0u);
        vidx_.assign(1, 0u);
        if (map_idx_ == 0) {
            vals_.assign({(float)vals[0], (float)vals[1]});
            out_.output_vals = (char *)vals_.data();
            out_.value_size = 2 * sizeof(float);
        } else {
            dvals_.assign(1, vals[0]);
            out_.output_vals = (char *)dvals_.data();
            out_.value_size = sizeof(double);
        }
        out_.output_keys = (char *)keys_.data();
        out_.key_index = kidx_.data();
        out_.val_index = vidx_.data();
        out_.key_num = 1;
        out_.key_size = sizeof(unsigned int);
        return out_;
    }

    static void fail(const char *msg)
    {
        fprintf(stderr, "pattern error: %s\n", msg);
        exit(EXIT_FAILURE);
    }
    pat_kmeans *handle_;
    int dim_, k_;
    const float *src_ = nullptr;
    std::vector<float> packed_;
    std::vector<unsigned int> keys_, kidx_, vidx_;
    std::vector<float> vals_;
    std::vector<double> dvals_;
    pat_gr *gr_ = nullptr;
    pat_regular *reg_ = nullptr;
    std::vector<char> kbytes_, vbytes_;
    struct output out_;
};

#endif
