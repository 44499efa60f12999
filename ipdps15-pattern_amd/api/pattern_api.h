// The user-functor interface of the pattern API, kept as the reference
// defines it so application functor headers compile unchanged:
//   map_fp / reduce_fp          generalized reduction (kmeans/regular.h:25-27)
//   irmap_fp / irreduce_fp      irregular reduction   (kmeans/irregular.h:22-24)
//   stencil_fp                  stencil with parameter (heat_3D/stencil.h:32)
// and the grid accessors GET_FLOAT3 & co (kmeans/lib/macro.h:70-76), which
// read an implicit 'size' array (halo-inclusive extents, dimension 0 fastest;
// here with 64-bit offsets, so grids past 2^31 cells address correctly).
// An application header selects host functors with CPU_KERNEL (namespaces
// FFCPU/FCPU, as the reference) or device functors with GPU_KERNEL (FFGPU/
// FGPU, compiled by hipcc with the templates of api/gpu/, whose
// PAT_*_TABLE macros register the table with the library's runtime); its
// function tables are declared CONSTANT = __device__ constexpr there, so
// every index dispatches at compile time.
#ifndef PATTERN_API_H_
#define PATTERN_API_H_

#include <cstring>
#include <vector>

#define PAT_GRID_OFFSET3(i, j, k) \
    ((long long)size[0] * (j) + (long long)(i) * size[0] * size[1] + (k))
#define GET_FLOAT3(input, i, j, k) *((float *)input + PAT_GRID_OFFSET3(i, j, k))
#define GET_DOUBLE3(input, i, j, k) *((double *)input + PAT_GRID_OFFSET3(i, j, k))
#define GET_INT3(input, i, j, k) *((int *)input + PAT_GRID_OFFSET3(i, j, k))
#define GET_FLOAT2(input, i, j) *((float *)input + (long long)size[0] * (i) + (j))
#define GET_INT2(input, i, j) *((int *)input + (long long)size[0] * (i) + (j))

// The generalized-reduction objects (CO on the host, SO on the device) and
// ANY_KEY / copy_val are in regular/objects.h; the irregular ones in
// irregular/objects.h.

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
