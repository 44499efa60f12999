// The user-functor interface of the pattern API, kept as the reference
// defines it so application functor headers compile unchanged:
//   map_fp / reduce_fp          generalized reduction (kmeans/regular.h:25-27)
//   irmap_fp / irreduce_fp      irregular reduction   (kmeans/irregular.h:22-24)
//   stencil_fp                  stencil with parameter (heat_3D/stencil.h:32)
// and the grid accessors GET_FLOAT3 & co (kmeans/lib/macro.h:70-76), which
// read an implicit 'size' array (halo-inclusive extents, dimension 0 fastest).
// CPU_KERNEL selects host functors (namespaces FFCPU/FCPU, as the reference);
// the GPU engines bind the library's compiled functors by table index.
#ifndef PATTERN_API_H_
#define PATTERN_API_H_

#include <cstring>
#include <vector>

#define GET_FLOAT3(input, i, j, k) *((float *)input + size[0] * (j) + (i)*size[0] * size[1] + (k))
#define GET_DOUBLE3(input, i, j, k) *((double *)input + size[0] * (j) + (i)*size[0] * size[1] + (k))
#define GET_INT3(input, i, j, k) *((int *)input + size[0] * (j) + (i)*size[0] * size[1] + (k))
#define GET_FLOAT2(input, i, j) *((float *)input + size[0] * (i) + j)
#define GET_INT2(input, i, j) *((int *)input + size[0] * (i) + j)

// Host reduction object for the generalized-reduction functors: a dense
// key -> value table with the insert(key, ksize, value, vsize, reduce) call
// the functors use (kmeans/regular/roc.h:39).
class CO {
  public:
    typedef void (*reduce_fn)(void *, unsigned short, void *, unsigned short);
    CO(int max_keys, int value_size) : nkeys_(max_keys), vsize_(value_size), used_(max_keys, 0),
                                       vals_((size_t)max_keys * value_size, 0) {}
    bool insert(void *key, unsigned short ksize, void *value, unsigned short vsize, reduce_fn reduce)
    {
        unsigned int k = 0;
        memcpy(&k, key, ksize < sizeof(k) ? ksize : sizeof(k));
        if ((int)k >= nkeys_ || vsize != vsize_) return false;
        char *slot = vals_.data() + (size_t)k * vsize_;
        if (!used_[k]) {
            memcpy(slot, value, vsize);
            used_[k] = 1;
        } else {
            reduce(slot, vsize, value, vsize);
        }
        return true;
    }
    bool used(int k) const { return used_[k] != 0; }
    const void *value(int k) const { return vals_.data() + (size_t)k * vsize_; }

  private:
    int nkeys_;
    unsigned short vsize_;
    std::vector<char> used_;
    std::vector<char> vals_;
};

static inline void copy_val(void *dst, const void *src, int size) { memcpy(dst, src, size); }

// Host reduction object for the irregular functors: a dense node -> VALUE
// array with the insert(&key, &value, reduce) call of
// kmeans/irregular/roc.h:19 (the reference's is a hash whose key is the node).
template <typename KEY_T, typename VALUE_T>
class DenseObject {
  public:
    typedef void (*reduce_fn)(VALUE_T *, VALUE_T *);
    explicit DenseObject(long n) : vals_(n), used_(n, 0) {}
    bool insert(KEY_T *key, VALUE_T *value, reduce_fn reduce)
    {
        const long k = (long)*key;
        if (k < 0 || k >= (long)vals_.size()) return false;
        if (!used_[k]) {
            vals_[k] = *value;
            used_[k] = 1;
        } else {
            reduce(&vals_[k], value);
        }
        return true;
    }
    const VALUE_T &value(long k) const { return vals_[k]; }
    bool used(long k) const { return used_[k] != 0; }

  private:
    std::vector<VALUE_T> vals_;
    std::vector<char> used_;
};

#endif
