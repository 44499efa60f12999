// miniMD functor table in the irregular-reduction API's shape
// (miniMD/irregular.h:109-174): map1 computes the LJ force of one neighbour
// pair without periodic images (parameter: double cutforcesq) and inserts
// +F for i and -F for j.  Written with the 48*eps prefactor.  Table index 0
// is the GPU engine's PAT_IR_LJ_OPEN.  Host version, used by the driver's
// check.
#ifndef IRREGULAR_H_
#define IRREGULAR_H_

#include "irregular/data_type.h"
#include "pattern_api.h"

namespace FCPU {

typedef DenseObject<KEY, VALUE> Cobject;
typedef void (*irreduce_fp)(VALUE *value1, VALUE *value2);
typedef bool (*irmap_fp)(Cobject *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                         unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr);

inline bool map1(Cobject *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                 unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    (void)part_index;
    (void)part_id;
    const float cutforcesq = (float)*(double *)parameter;
    const float *x = (const float *)point_data;
    const int *edges = (const int *)edge_data;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    const float dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (rd < cutforcesq) {
        const float rrd = 1.0f / rd, rrd2 = rrd * rrd, rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
        const float F = 48.0f * (rrd7 - 0.5f * rrd4);
        VALUE f = {dx * F, dy * F, dz * F};
        object->insert(&i, &f, reduce_ptr);
        f.x = -f.x;
        f.y = -f.y;
        f.z = -f.z;
        object->insert(&j, &f, reduce_ptr);
    }
    return true;
}

inline void reduce(VALUE *value1, VALUE *value2)
{
    value1->x += value2->x;
    value1->y += value2->y;
    value1->z += value2->z;
}

static const irmap_fp map_function_table[] = {map1};
static const irreduce_fp reduce_function_table[] = {reduce};

}  // namespace FCPU

#define PAT_IR_MAP_TABLE {PAT_IR_LJ_OPEN}
#endif
