// Generalized-reduction functor tables in the API's shape (map_fp /
// reduce_fp, kmeans/regular.h:25-27), compiled for the host (CPU_KERNEL:
// namespace FFCPU, object CO) or for the device (GPU_KERNEL: FFGPU, object
// SO; see gpu_table.hip):
//   0  map / reduce    moldyn's velocity statistics (kmeans/regular.h:30-73):
//                      unit = a velocity (double[3]), parameter float vaverh;
//                      key ANY_KEY(), value float[2] {|v|, |v| > vaverh};
//                      reduce adds both;
//   1  map1 / reduce1  kmeans' nearest centre (kmeans/regular.h:75-136) at the
//                      reference's sizes (KMEANS_DIM 3, KMEANS_K 40,
//                      kmeans/kmeans.h:4-5): unit = a point (float[3]),
//                      parameter the K x 3 centres; key the nearest centre
//                      (FP32 distances in dimension order, sqrt, strict <
//                      from 65536*65), value float[5] {x, y, z, 1, min_dist};
//                      reduce1 adds all five;
//   2  map2 / reduce2  miniMD's pair energy (miniMD/regular.h:139-181): unit =
//                      an edge (int[2]), parameter {double cutsq,
//                      double x[3N]}; key ANY_KEY(), value double
//                      phi = sr6 (sr6 - 1) for pairs within the cutoff;
//                      reduce2 adds;
//   3  map3 / reduce3  a many-key histogram: unit = an unsigned value v, key
//                      v mod 5003, value {long long count 1, double v/1000};
//                      reduce3 adds both;
//   4  map4 / reduce4  records of odd sizes: unit = an unsigned value v, key
//                      the 6 bytes {v % 7, v / 7 % 11, v / 77 % 3} (231
//                      keys), value the 3 bytes {1, v & 255, v >> 8 & 255};
//                      reduce4 adds them byte by byte modulo 256.
// The reference spreads one logical key over ANY_KEY's 41 buckets on the
// device and uses key 0 on the host; a caller adds the buckets.
#ifndef REDUCTION
#define REDUCTION

#include <math.h>

#include "regular/objects.h"

#define KMEANS_DIM 3
#define KMEANS_K 40
#define HIST_KEYS 5003

#if defined(GPU_KERNEL)
#define DEVICE __device__
#define OBJECT SO
#define CONSTANT __device__ constexpr
namespace FFGPU {
#elif defined(CPU_KERNEL)
#define DEVICE inline
#define OBJECT CO
#define CONSTANT static constexpr
namespace FFCPU {
#endif

typedef void (*reduce_fp)(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size);
typedef bool (*map_fp)(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                       reduce_fp reduce_ptr);

DEVICE bool map(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                reduce_fp reduce_ptr)
{
    (void)device_type;
    const float vaverh = *(float *)parameter;
    const double *v = (const double *)((char *)global_data + index);
    unsigned int k = ANY_KEY();
    float value[2];
    value[0] = (float)sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    value[1] = value[0] > vaverh ? 1.0f : 0.0f;
    object->insert(&k, sizeof(k), value, sizeof(value), reduce_ptr);
    return true;
}

DEVICE void reduce(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value1_size, (void)value2_size;
    ((float *)value1)[0] += ((float *)value2)[0];
    ((float *)value1)[1] += ((float *)value2)[1];
}

DEVICE bool map1(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                 reduce_fp reduce_ptr)
{
    (void)device_type;
    const float *x = (const float *)((char *)global_data + index);
    const float *c = (const float *)parameter;
    unsigned int key = 0;
    float min_dist = 65536 * 65;
    for (int i = 0; i < KMEANS_K; i++) {
        float dist = 0;
        for (int d = 0; d < KMEANS_DIM; d++) {
            const float t = c[KMEANS_DIM * i + d] - x[d];
            dist = dist + t * t;
        }
        dist = sqrtf(dist);
        if (dist < min_dist) {
            min_dist = dist;
            key = i;
        }
    }
    float value[KMEANS_DIM + 2];
    for (int d = 0; d < KMEANS_DIM; d++) value[d] = x[d];
    value[KMEANS_DIM] = 1;
    value[KMEANS_DIM + 1] = min_dist;
    return object->insert(&key, sizeof(key), value, sizeof(value), reduce_ptr);
}

DEVICE void reduce1(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value2_size;
    const int n = value1_size / (int)sizeof(float);
    float t[KMEANS_DIM + 2];
    for (int i = 0; i < n && i < KMEANS_DIM + 2; i++) t[i] = ((float *)value1)[i] + ((float *)value2)[i];
    copy_val(value1, t, sizeof(float) * (KMEANS_DIM + 2));
}

DEVICE bool map2(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                 reduce_fp reduce_ptr)
{
    (void)device_type;
    const double cutsq = ((const double *)parameter)[0];
    const double *v = (const double *)parameter + 1;
    const int *e = (const int *)((char *)global_data + index);
    const int i = e[0], j = e[1];
    const double dx = v[3 * i] - v[3 * j], dy = v[3 * i + 1] - v[3 * j + 1], dz = v[3 * i + 2] - v[3 * j + 2];
    const double rsq = dx * dx + dy * dy + dz * dz;
    int k = ANY_KEY();
    if (rsq < cutsq) {
        const double sr2 = 1.0 / rsq, sr6 = sr2 * sr2 * sr2;
        double phi = sr6 * (sr6 - 1.0);
        object->insert(&k, sizeof(int), &phi, sizeof(double), reduce_ptr);
    }
    return true;
}

DEVICE void reduce2(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value1_size, (void)value2_size;
    ((double *)value1)[0] += ((double *)value2)[0];
}

struct HistValue {
    long long count;
    double sum;
};

DEVICE bool map3(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                 reduce_fp reduce_ptr)
{
    (void)parameter, (void)device_type;
    const unsigned int v = *(const unsigned int *)((char *)global_data + index);
    unsigned int key = v % HIST_KEYS;
    HistValue h = {1, (double)v / 1000.0};
    return object->insert(&key, sizeof(key), &h, sizeof(h), reduce_ptr);
}

DEVICE void reduce3(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value1_size, (void)value2_size;
    ((HistValue *)value1)->count += ((HistValue *)value2)->count;
    ((HistValue *)value1)->sum += ((HistValue *)value2)->sum;
}

struct SmallKey {
    unsigned short a, b, c;
};

DEVICE bool map4(OBJECT *object, void *global_data, unsigned long long index, void *parameter, int device_type,
                 reduce_fp reduce_ptr)
{
    (void)parameter, (void)device_type;
    const unsigned int v = *(const unsigned int *)((char *)global_data + index);
    SmallKey key = {(unsigned short)(v % 7), (unsigned short)(v / 7 % 11), (unsigned short)(v / 77 % 3)};
    unsigned char value[3] = {1, (unsigned char)(v & 255), (unsigned char)(v >> 8 & 255)};
    return object->insert(&key, sizeof(key), value, sizeof(value), reduce_ptr);
}

DEVICE void reduce4(void *value1, unsigned short value1_size, void *value2, unsigned short value2_size)
{
    (void)value2_size;
    unsigned char *a = (unsigned char *)value1;
    const unsigned char *b = (const unsigned char *)value2;
    for (int i = 0; i < value1_size; i++) a[i] = (unsigned char)(a[i] + b[i]);
}

CONSTANT map_fp map_function_table[] = {map, map1, map2, map3, map4};
CONSTANT reduce_fp reduce_function_table[] = {reduce, reduce1, reduce2, reduce3, reduce4};

}  // namespace FFCPU / FFGPU
#endif
