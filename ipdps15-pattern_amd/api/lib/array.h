// IndexArray / IntArray: the small fixed-size index vectors of the stencil
// API (kmeans/lib/array.h:10-60, there built on boost::array).
#ifndef LIB_ARRAY_H_
#define LIB_ARRAY_H_

#include <cstddef>
#include <cstdint>

template <typename T>
class PatArray3 {
  public:
    T v[3];
    PatArray3() : v{0, 0, 0} {}
    explicit PatArray3(T a) : v{a, 0, 0} {}
    PatArray3(T a, T b) : v{a, b, 0} {}
    PatArray3(T a, T b, T c) : v{a, b, c} {}
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
    T accumulate(int n) const
    {
        T p = 1;
        for (int i = 0; i < n; i++) p *= v[i];
        return p;
    }
};

typedef PatArray3<int64_t> IndexArray;
typedef PatArray3<int> IntArray;

#endif
