// miniMD functor tables in the irregular-reduction API's shape
// (miniMD/irregular.h:109-174), compiled for the host (CPU_KERNEL: FCPU,
// Cobject) or the device (GPU_KERNEL: FGPU, Sobject; see gpu_table.hip).
// Node data is double[3] per atom, the parameter double cutforcesq; no
// periodic images, as miniMD has none.  Each map inserts +F for i and -F for
// j where the partition owns the node:
//   map 0  map1      the workload's FP32 force: positions and cutoff rounded
//                    to float, F = 48 (r^-14 - r^-8 / 2) d -- the bits of the
//                    library's PAT_IR_LJ_OPEN kernels;
//   map 1  map1_ref  the reference functor's arithmetic: FP64,
//                    force = rrd3 (rrd3 - 0.5) rrd with rrd3 = r^-6 (no
//                    48 eps prefactor), rounded to float at the insert.
//   reduce 0         adds VALUEs.
#ifndef IRREGULAR_H_
#define IRREGULAR_H_

#include "irregular/objects.h"

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

DEVICE bool map1(OBJECT *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                 unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    const float cutforcesq = (float)*(double *)parameter;
    const double *x = (const double *)point_data;
    const int *edges = (const int *)edge_data;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    const float dx = (float)x[3 * i] - (float)x[3 * j], dy = (float)x[3 * i + 1] - (float)x[3 * j + 1],
                dz = (float)x[3 * i + 2] - (float)x[3 * j + 2];
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (rd < cutforcesq) {
        const float rrd = 1.0f / rd, rrd2 = rrd * rrd, rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
        const float F = 48.0f * (rrd7 - 0.5f * rrd4);
        VALUE f = {dx * F, dy * F, dz * F};
        if (part_index[i] == (int)part_id) object->insert(&i, &f, reduce_ptr);
        f.x = -f.x;
        f.y = -f.y;
        f.z = -f.z;
        if (part_index[j] == (int)part_id) object->insert(&j, &f, reduce_ptr);
    }
    return true;
}

DEVICE bool map1_ref(OBJECT *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                     unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    const double cutforcesq = *(double *)parameter;
    const double *x = (const double *)point_data;
    const int *edges = (const int *)edge_data;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    const double rd = dx * dx + dy * dy + dz * dz;
    if (rd < cutforcesq) {
        const double rrd = 1.0 / rd, rrd2 = rrd * rrd, rrd3 = rrd2 * rrd;
        const double force = rrd3 * (rrd3 - 0.5) * rrd;
        VALUE f = {(float)(dx * force), (float)(dy * force), (float)(dz * force)};
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

CONSTANT irmap_fp map_function_table[] = {map1, map1_ref};
CONSTANT irreduce_fp reduce_function_table[] = {reduce};

}  // namespace FCPU / FGPU

// Library functors with the same bits, for runs without the device table.
#define PAT_IR_MAP_TABLE {PAT_IR_LJ_OPEN, -1}
#endif
