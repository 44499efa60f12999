// moldyn functor table in the irregular-reduction API's shape
// (kmeans/irregular.h:22-118, moldyn.h:32-43): map computes the LJ force of
// one edge with the minimum image in a cubic box and inserts +F for i and
// -F for j; reduce adds VALUEs.  Written with the 48*eps prefactor the
// reference folds into its integrator.  Table index 0 is the GPU engine's
// PAT_IR_LJ_MINIMAGE.  Host versions, used by the driver's check.
#ifndef IRREGULAR_H_
#define IRREGULAR_H_

#include "irregular/data_type.h"
#include "pattern_api.h"

struct point_data {
    double points[3];
};

struct para {
    double side;
    double cutoffRadius;
};

namespace FCPU {

typedef DenseObject<KEY, VALUE> Cobject;
typedef void (*irreduce_fp)(VALUE *value1, VALUE *value2);
typedef bool (*irmap_fp)(Cobject *object, void *point_data, int *part_index, void *edge_data, void *parameter,
                         unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr);

inline bool map(Cobject *object, void *point_data_, int *part_index, void *edge_data, void *parameter,
                unsigned int task_id, unsigned int part_id, irreduce_fp reduce_ptr)
{
    (void)part_index;
    (void)part_id;
    const struct para *pa = (const struct para *)parameter;
    const float *pos = (const float *)point_data_;
    const int *edges = (const int *)edge_data;
    const float side = (float)pa->side, half = 0.5f * side;
    const float rc = (float)pa->cutoffRadius;
    int i = edges[task_id * 2], j = edges[task_id * 2 + 1];
    float d[3];
    for (int c = 0; c < 3; c++) {
        d[c] = pos[3 * i + c] - pos[3 * j + c];
        if (d[c] < -half) d[c] += side;
        if (d[c] > half) d[c] -= side;
    }
    float rd = d[0] * d[0] + d[1] * d[1];
    rd = rd + d[2] * d[2];
    if (rd < rc * rc) {
        const float rrd = 1.0f / rd, rrd2 = rrd * rrd, rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
        const float F = 48.0f * (rrd7 - 0.5f * rrd4);
        VALUE f = {d[0] * F, d[1] * F, d[2] * F};
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

static const irmap_fp map_function_table[] = {map};
static const irreduce_fp reduce_function_table[] = {reduce};

}  // namespace FCPU

#define PAT_IR_MAP_TABLE {PAT_IR_LJ_MINIMAGE}
#endif
