// moldyn functor tables in the irregular-reduction API's shape
// (kmeans/irregular.h:22-118, moldyn.h:32-43), compiled for the host
// (CPU_KERNEL: namespace FCPU, object Cobject) or for the device
// (GPU_KERNEL: namespace FGPU, object Sobject; see gpu_table.hip).  Node data
// is moldyn's point_data, double[3] per molecule; the parameter is
// struct para {double side; double cutoffRadius;}.  Each map computes one
// edge's Lennard-Jones force with the minimum image in a cubic box and
// inserts +F for i and -F for j, each only where the partition owns the node
// (part_index[node] == part_id):
//   map 0  map      the workload's FP32 force: positions, side and cutoff
//                   rounded to float, F = 48 (r^-14 - r^-8 / 2) d -- the bits
//                   of the library's PAT_IR_LJ_MINIMAGE kernels;
//   map 1  map_ref  the reference functor's arithmetic: FP64 throughout,
//                   r148 d without the 48 eps prefactor (which moldyn folds
//                   into its integrator), rounded to float at the insert.
//   reduce 0        adds VALUEs.
#ifndef IRREGULAR_H_
#define IRREGULAR_H_

#include "irregular/objects.h"

struct point_data {
    double points[3];
};

struct para {
    double side;
    double cutoffRadius;
};

#if defined(GPU_KERNEL)
#define DEVICE __device__
#define OBJECT Sobject
#define CONSTANT __device__ constexpr
namespace FGPU {
#elif defined(CPU_KERNEL)
#define DEVICE inline
#define OBJECT Cobject
#define CONSTANT static constexpr
namespace FCPU {
#endif

typedef void (*irreduce_fp)(VALUE *value1, VALUE *value2);
typedef bool (*irmap_fp)(OBJECT *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                         unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr);

DEVICE bool map(OBJECT *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    const struct para *pa = (const struct para *)parameter;
    const double *x = ((const struct point_data *)point_data)->points;
    const int *edges = (const int *)edge_data;
    const float side = (float)pa->side, half = 0.5f * side;
    const float rc = (float)pa->cutoffRadius;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    float d[3];
    for (int c = 0; c < 3; c++) d[c] = (float)x[3 * i + c] - (float)x[3 * j + c];
    for (int c = 0; c < 3; c++)
        if (d[c] < -half) d[c] += side;
    for (int c = 0; c < 3; c++)
        if (d[c] > half) d[c] -= side;
    float rd = d[0] * d[0] + d[1] * d[1];
    rd = rd + d[2] * d[2];
    if (rd < rc * rc) {
        const float rrd = 1.0f / rd, rrd2 = rrd * rrd, rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
        const float F = 48.0f * (rrd7 - 0.5f * rrd4);
        VALUE f = {d[0] * F, d[1] * F, d[2] * F};
        if (part_index[i] == (int)part_id) object->insert(&i, &f, reduce_ptr);
        f.x = -f.x;
        f.y = -f.y;
        f.z = -f.z;
        if (part_index[j] == (int)part_id) object->insert(&j, &f, reduce_ptr);
    }
    return true;
}

DEVICE bool map_ref(OBJECT *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                    unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    const struct para *pa = (const struct para *)parameter;
    const double *x = ((const struct point_data *)point_data)->points;
    const int *edges = (const int *)edge_data;
    const double side = pa->side, half = 0.5 * side, cut2 = pa->cutoffRadius * pa->cutoffRadius;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    double d[3];
    for (int c = 0; c < 3; c++) d[c] = x[3 * i + c] - x[3 * j + c];
    for (int c = 0; c < 3; c++)
        if (d[c] < -half) d[c] += side;
    for (int c = 0; c < 3; c++)
        if (d[c] > half) d[c] -= side;
    const double rd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (rd < cut2) {
        const double rrd = 1.0 / rd, rrd2 = rrd * rrd, rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
        const double r148 = rrd7 - 0.5 * rrd4;
        VALUE f = {(float)(d[0] * r148), (float)(d[1] * r148), (float)(d[2] * r148)};
        if (part_index[i] == (int)part_id) object->insert(&i, &f, reduce_ptr);
        f.x = -f.x;
        f.y = -f.y;
        f.z = -f.z;
        if (part_index[j] == (int)part_id) object->insert(&j, &f, reduce_ptr);
    }
    return true;
}

DEVICE void reduce(VALUE *value1, VALUE *value2)
{
    value1->x += value2->x;
    value1->y += value2->y;
    value1->z += value2->z;
}

CONSTANT irmap_fp map_function_table[] = {map, map_ref};
CONSTANT irreduce_fp reduce_function_table[] = {reduce};

}  // namespace FCPU / FGPU

// Library functors with the same bits, for runs without the device table.
#define PAT_IR_MAP_TABLE {PAT_IR_LJ_MINIMAGE, -1}
#endif
