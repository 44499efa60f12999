// kmeans functor table in the generalized-reduction API's shape
// (kmeans/regular.h:25-27, 75-145), generalised to D dimensions and K
// centres: map1 emits key = nearest centre (FP32 difference form in
// dimension order, sqrt, strict <, lowest index wins) with value
// {x_0..x_{D-1}, 1, min_dist}; reduce1 adds values.  Table index 1 is the
// GPU engine's kmeans pass.  Host versions, used by the program's host check.
#ifndef REDUCTION
#define REDUCTION

#include <cmath>

#include "pattern_api.h"
#include "regular/objects.h"

namespace FFCPU {

struct KmParam {
    const float *centres;
    int dim, k;
};

typedef void (*reduce_fp)(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size);
typedef bool (*map_fp)(CO *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                       reduce_fp reduce_ptr);

inline bool map1(CO *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                 reduce_fp reduce_ptr)
{
    (void)device_type;
    const KmParam *p = (const KmParam *)parameter;
    const float *x = (const float *)((char *)global_data + index);
    unsigned int key = 0;
    float min_dist = 65536 * 65;
    for (int i = 0; i < p->k; i++) {
        float dist = 0;
        for (int d = 0; d < p->dim; d++) {
            const float t = p->centres[i * p->dim + d] - x[d];
            dist = dist + t * t;
        }
        dist = std::sqrt(dist);
        if (dist < min_dist) {
            min_dist = dist;
            key = i;
        }
    }
    float value[130];
    for (int d = 0; d < p->dim; d++) value[d] = x[d];
    value[p->dim] = 1;
    value[p->dim + 1] = min_dist;
    return object->insert(&key, sizeof(key), value, (unsigned short)((p->dim + 2) * sizeof(float)), reduce_ptr);
}

inline void reduce1(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value2_size;
    const int n = value1_size / sizeof(float);
    for (int i = 0; i < n; i++) ((float *)value1)[i] += ((float *)value2)[i];
}

static const map_fp map_function_table[] = {nullptr, map1};
static const reduce_fp reduce_function_table[] = {nullptr, reduce1};

}  // namespace FFCPU
#endif
