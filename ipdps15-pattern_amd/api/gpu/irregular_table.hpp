// Device execution of an application's irregular-reduction functor table.
//
// The reference runs an edge functor per edge of a partition and inserts each
// result into an LDS hash keyed by node, merged into a device hash afterwards
// (kmeans/irregular/gpu_kernel.cu:21-103, ros.cu, rog.cu); a functor only
// inserts for the endpoints its partition owns (part_index[node] == part_id,
// kmeans/irregular.h:80-90).  Here the edges are reordered once on the device
// into a CSR by destination whose rows list each node's edge ids in
// increasing order, and every node is a partition of its own: thread r runs
// the map functor over row r's edges with part_index = identity and
// part_id = r, so the functor inserts only r's values, into an object held in
// registers that copies the first value and folds later ones with the
// reduce functor -- in edge order, the serial scatter's order for node r.  No
// hash, no atomics, and the same bits every run.
//
// An application compiles its functor header with GPU_KERNEL defined and
// names its tables:
//
//     #define GPU_KERNEL
//     #include "irregular.h"
//     #include "gpu/irregular_table.hpp"
//     static const int kSameAs[] = {PAT_IR_LJ_MINIMAGE};  // or -1
//     PAT_IR_TABLE(FGPU::map_function_table, FGPU::reduce_function_table, kSameAs)
//
// The object type its functors take under GPU_KERNEL is Sobject
// (api/irregular/objects.h).
#ifndef PAT_GPU_IRREGULAR_TABLE_HPP_
#define PAT_GPU_IRREGULAR_TABLE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "irregular/objects.h"
#include "pattern.h"

namespace pat_gpu {

template <class Tab, int MI, int RI>
__global__ __launch_bounds__(256) void ir_reduce_rows(const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ eids, int N, void *point_data,
                                                      int *part_index, void *edge_data, void *param,
                                                      VALUE *__restrict__ out, int32_t *__restrict__ foreign)
{
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= N) return;
    Sobject obj(r);
    const int b = rowptr[r], e = rowptr[r + 1];
    for (int k = b; k < e; k++)
        Tab::template map<MI, RI>(&obj, point_data, part_index, edge_data, param, (unsigned)eids[k], (unsigned)r);
    VALUE v;
    if (obj.used()) {
        v = obj.value();
    } else {
        v.x = 0.f;
        v.y = 0.f;
        v.z = 0.f;
    }
    out[r] = v;
    if (obj.refused()) *foreign = 1;
}

template <class Tab, int MI, int RI>
int ir_launch(const int32_t *rowptr, const int32_t *eids, int64_t N, void *point_data, int32_t *part_index,
              void *edge_data, void *param, void *out, int32_t *foreign, hipStream_t st)
{
    if constexpr (!Tab::template valid<MI, RI>()) {
        return PAT_ERR_UNSUPPORTED;
    } else {
        hipLaunchKernelGGL((ir_reduce_rows<Tab, MI, RI>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                           rowptr, eids, (int)N, point_data, (int *)part_index, edge_data, param, (VALUE *)out,
                           foreign);
        return hipGetLastError() == hipSuccess ? PAT_OK : PAT_ERR_HIP;
    }
}

template <class Tab, int MI, std::size_t... RI>
int ir_dispatch_reduce(int ri, const int32_t *rowptr, const int32_t *eids, int64_t N, void *pd, int32_t *pi,
                       void *ed, void *prm, void *out, int32_t *foreign, hipStream_t st, std::index_sequence<RI...>)
{
    int rc = PAT_ERR_UNSUPPORTED;
    ((ri == (int)RI ? (void)(rc = ir_launch<Tab, MI, (int)RI>(rowptr, eids, N, pd, pi, ed, prm, out, foreign, st))
                    : (void)0),
     ...);
    return rc;
}

template <class Tab, std::size_t... MI>
int ir_dispatch(int mi, int ri, const int32_t *rowptr, const int32_t *eids, int64_t N, void *pd, int32_t *pi,
                void *ed, void *prm, void *out, int32_t *foreign, hipStream_t st, std::index_sequence<MI...>)
{
    int rc = PAT_ERR_UNSUPPORTED;
    ((mi == (int)MI ? (void)(rc = ir_dispatch_reduce<Tab, (int)MI>(
                                 ri, rowptr, eids, N, pd, pi, ed, prm, out, foreign, st,
                                 std::make_index_sequence<(std::size_t)Tab::rcount>{}))
                    : (void)0),
     ...);
    return rc;
}

// pat_ir_table::reduce_rows for the tables of Tab.
template <class Tab>
int ir_reduce_entry(int map_idx, int reduce_idx, const int32_t *rowptr, const int32_t *eids, int64_t num_nodes,
                    const void *point_data, const int32_t *part_index, const void *edge_data, const void *param,
                    void *out, int32_t *foreign, void *stream)
{
    if (!rowptr || !eids || !part_index || !out || !foreign || num_nodes <= 0 || num_nodes >= (1LL << 31))
        return PAT_ERR_INVALID;
    return ir_dispatch<Tab>(map_idx, reduce_idx, rowptr, eids, num_nodes, const_cast<void *>(point_data),
                            const_cast<int32_t *>(part_index), const_cast<void *>(edge_data),
                            const_cast<void *>(param), out, foreign, (hipStream_t)stream,
                            std::make_index_sequence<(std::size_t)Tab::mcount>{});
}

}  // namespace pat_gpu

// Defines pat_app_ir_table() for the application's map and reduce tables
// (__device__ constexpr arrays); SAME_AS[i] names the library functor that
// computes map i's bits (PAT_IR_LJ_MINIMAGE, PAT_IR_LJ_OPEN) or is -1.
#define PAT_IR_TABLE(MAP_TABLE, REDUCE_TABLE, SAME_AS)                                                          \
    namespace {                                                                                                 \
    struct PatIrTab {                                                                                           \
        static constexpr int mcount = (int)(sizeof(MAP_TABLE) / sizeof(MAP_TABLE[0]));                          \
        static constexpr int rcount = (int)(sizeof(REDUCE_TABLE) / sizeof(REDUCE_TABLE[0]));                    \
        template <int MI, int RI>                                                                               \
        static constexpr bool valid()                                                                           \
        {                                                                                                       \
            return MAP_TABLE[MI] != nullptr && REDUCE_TABLE[RI] != nullptr;                                     \
        }                                                                                                       \
        template <int MI, int RI>                                                                               \
        __device__ static void map(Sobject *o, void *pd, int *pi, void *ed, void *prm, unsigned t, unsigned p)  \
        {                                                                                                       \
            MAP_TABLE[MI](o, pd, pi, ed, prm, t, p, REDUCE_TABLE[RI]);                                          \
        }                                                                                                       \
    };                                                                                                          \
    static_assert(sizeof(SAME_AS) / sizeof(SAME_AS[0]) == PatIrTab::mcount, "one same-as entry per map");       \
    }                                                                                                           \
    extern "C" const pat_ir_table *pat_app_ir_table(void)                                                       \
    {                                                                                                           \
        static const pat_ir_table t = {PatIrTab::mcount, PatIrTab::rcount, (int)sizeof(VALUE), SAME_AS,         \
                                       &pat_gpu::ir_reduce_entry<PatIrTab>};                                    \
        return &t;                                                                                              \
    }

#endif
