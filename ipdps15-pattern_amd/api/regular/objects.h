// The reduction objects a generalized-reduction functor inserts into, with
// the insert(key, key_size, value, value_size, reduce_ptr) call of the
// reference (kmeans/regular/ros.h:33 on the GPU, roc.h:39 on the CPU): the
// first value for a key is copied, later ones are folded into it by the
// reduce functor; insert returns false when the object has no room.
//
// CO (host): a map from key bytes to value bytes, never full.
// SO (device, api/gpu/regular_table.hpp): a wavefront's own open-addressing
// hash table in LDS.  The lanes that call insert together take turns in lane
// order, so the table needs no locks and every value is folded in a fixed
// order.  When it is full, insert returns false; the engine then merges the
// table into the wavefront's table in device memory, clears it and runs the
// unit's map again (the reference's merge-and-retry,
// kmeans/regular/gpu_kernel.cu:143-150).
//
// Also: ANY_KEY() (kmeans/regular/kernel_tools.h:15 -- 0 on the host, the
// thread index mod 41 on the device, spreading one logical key over 41
// buckets as the reference does) and copy_val.
#ifndef REGULAR_OBJECTS_H_
#define REGULAR_OBJECTS_H_

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "data_type.h"

class CO {
  public:
    typedef void (*reduce_fn)(void *, unsigned short, void *, unsigned short);
    CO() = default;
    // kept for drivers that size a dense object: the bound is not needed
    CO(int max_keys, int value_size) { (void)max_keys, (void)value_size; }
    bool insert(void *key, unsigned short ksize, void *value, unsigned short vsize, reduce_fn reduce)
    {
        std::string k((const char *)key, ksize);
        auto it = vals_.find(k);
        if (it == vals_.end()) {
            Val v(vsize);
            memcpy(v.p, value, vsize);
            vals_.emplace(std::move(k), std::move(v));
        } else {
            reduce(it->second.p, (unsigned short)it->second.n, value, vsize);
        }
        return true;
    }
    // key bytes -> value bytes
    std::map<std::string, std::string> entries() const
    {
        std::map<std::string, std::string> r;
        for (const auto &e : vals_) r.emplace(e.first, std::string((const char *)e.second.p, e.second.n));
        return r;
    }
    // by an unsigned int key
    bool used(unsigned int k) const { return vals_.count(ukey(k)) != 0; }
    const void *value(unsigned int k) const { return vals_.at(ukey(k)).p; }

  private:
    // a value's bytes in malloc'd storage, which the functors read and write
    // as their own types
    struct Val {
        void *p = nullptr;
        size_t n = 0;
        explicit Val(size_t bytes) : p(std::malloc(bytes ? bytes : 1)), n(bytes) {}
        Val(Val &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; }
        Val(const Val &) = delete;
        Val &operator=(const Val &) = delete;
        ~Val() { std::free(p); }
    };
    static std::string ukey(unsigned int k) { return std::string((const char *)&k, sizeof(k)); }
    std::map<std::string, Val> vals_;
};

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

namespace pat_gpu {

// Words of the tables: keys and values of any type are stored into and read
// out of them, so every access goes through a type that may alias the
// functors' own (float, double, struct) accesses to the same bytes.
typedef uint32_t __attribute__((__may_alias__)) ro_word;

// FNV-1a over the key bytes.
__device__ __forceinline__ uint32_t ro_hash(const unsigned char *k, int n)
{
    uint32_t h = 2166136261u;
    for (int i = 0; i < n; i++) h = (h ^ k[i]) * 16777619u;
    return h;
}

__device__ __forceinline__ void ro_copy(void *dst, const void *src, int n)
{
    if ((((uintptr_t)dst | (uintptr_t)src | (uintptr_t)n) & 3) == 0) {
        for (int i = 0; i < n / 4; i++) ((ro_word *)dst)[i] = ((const ro_word *)src)[i];
    } else {
        for (int i = 0; i < n; i++) ((unsigned char *)dst)[i] = ((const unsigned char *)src)[i];
    }
}

__device__ __forceinline__ bool ro_equal(const void *a, const void *b, int n)
{
    for (int i = 0; i < n; i++)
        if (((const unsigned char *)a)[i] != ((const unsigned char *)b)[i]) return false;
    return true;
}

// One hash table: nb buckets (a power of two) of 32-bit states (0 empty,
// else 1 + the record's word offset in the pool), then the pool's fill
// counter, then pw pool words.  Record: a header word (key_size << 16 |
// value_size), the key and the value, each padded to whole words.
struct RoTable {
    ro_word *bk, *top, *pool;
    int nb, pw;

    __device__ static RoTable at(ro_word *base, int nb, int pw)
    {
        return RoTable{base, base + nb, base + nb + 1, nb, pw};
    }
    __device__ static int words(int nb, int pw) { return nb + 1 + pw; }
    __device__ void clear(int lane, int nlanes)
    {
        for (int b = lane; b < nb; b += nlanes) bk[b] = 0u;
        if (lane == 0) *top = 0u;
    }
    __device__ static int key_words(int ks) { return (ks + 3) >> 2; }

    // Single writer: finds key or adds it (copy), else folds value into it.
    // false when neither a bucket nor pool room is left.
    template <typename R>
    __device__ bool insert(const void *key, int ks, const void *val, int vs, R reduce)
    {
        const uint32_t h = ro_hash((const unsigned char *)key, ks);
        const uint32_t hdr = ((uint32_t)ks << 16) | (uint32_t)vs;
        for (int p = 0; p < nb; p++) {
            const int b = (int)((h + (uint32_t)p) & (uint32_t)(nb - 1));
            const uint32_t st = bk[b];
            if (st == 0u) {
                const int need = 1 + key_words(ks) + ((vs + 3) >> 2);
                const uint32_t rec = *top;
                if ((int)rec + need > pw) return false;
                *top = rec + (uint32_t)need;
                ro_word *r = pool + rec;
                r[0] = hdr;
                for (int w = 0; w < key_words(ks); w++) r[1 + w] = 0u;
                ro_copy(r + 1, key, ks);
                ro_copy(r + 1 + key_words(ks), val, vs);
                bk[b] = rec + 1u;
                return true;
            }
            ro_word *r = pool + (st - 1u);
            if (r[0] == hdr && ro_equal(r + 1, key, ks)) {
                reduce((void *)(r + 1 + key_words(ks)), (unsigned short)vs, const_cast<void *>(val),
                       (unsigned short)vs);
                return true;
            }
        }
        return false;
    }
};

}  // namespace pat_gpu

// A wavefront's LDS reduction object.  Lanes that call insert together take
// turns in lane order.  Per lane, over one map call: full = an insert was
// refused, accepted = inserts taken.
class SO {
  public:
    pat_gpu::RoTable t;
    bool full = false;
    int accepted = 0;

    template <typename R>
    __device__ bool insert(void *key, unsigned short ksize, void *value, unsigned short vsize, R reduce_ptr)
    {
        unsigned long long turn = __ballot(1);
        const int lane = (int)__lane_id();
        bool ok = false;
        while (turn) {
            const int leader = __ffsll((long long)turn) - 1;
            if (lane == leader) ok = t.insert(key, ksize, value, vsize, reduce_ptr);
            turn &= turn - 1;
            __builtin_amdgcn_wave_barrier();
        }
        if (ok)
            accepted++;
        else
            full = true;
        return ok;
    }
};

__device__ __forceinline__ unsigned int ANY_KEY() { return threadIdx.x % 41; }
__device__ __forceinline__ void copy_val(void *dst, const void *src, unsigned size) { pat_gpu::ro_copy(dst, src, (int)size); }
#else
inline unsigned int ANY_KEY() { return 0; }
inline void copy_val(void *dst, const void *src, unsigned size) { memcpy(dst, src, size); }
#endif

#endif
