// IrregularRuntime (kmeans/irregular/irregular_runtime.h:12-112): the
// irregular-reduction runtime.  Same constructor and calls; the edge functor
// selected by set_map_idx runs on the GPU through libpattern and results come
// back in the caller's node order (the caller's arrays are never permuted).
//
// Each application compiles its own functor tables (the reference builds one
// binary per app directory).  An application that links its tables compiled
// for the device (api/gpu/irregular_table.hpp, which defines
// pat_app_ir_table) has them registered at IrregularInit: its map and reduce
// indices then select its own functors, which run on the GPU.  Without
// them, PAT_IR_MAP_TABLE (defined before including this header) maps each map
// index to the library functor with the same bits (-1: none); the default is
// moldyn's {PAT_IR_LJ_MINIMAGE} (kmeans/irregular.h:116, minimum-image LJ,
// parameter struct para {double side; double cutoffRadius;}); miniMD's
// (miniMD/irregular.h:172, no images, parameter double cutforcesq) is
// {PAT_IR_LJ_OPEN}.
#ifndef IRREGULAR_IRREGULAR_RUNTIME_H_
#define IRREGULAR_IRREGULAR_RUNTIME_H_

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../lib/common.h"
#include "../pat_check.h"
#include "data_type.h"

#ifndef PAT_IR_MAP_TABLE
#define PAT_IR_MAP_TABLE {PAT_IR_LJ_MINIMAGE}
#endif

// Defined by an application's device functor tables when they are linked in.
extern "C" const pat_ir_table *pat_app_ir_table(void) __attribute__((weak));

class IrregularRuntime {
  public:
    IRIndex global_num_edges_;
    IRIndex global_num_nodes_;
    EDGE *global_edges_;
    void *global_edge_data_;
    void *global_node_data_;
    int edge_data_elm_size_;
    int node_data_elm_size_;
    int reduction_elm_size_;
    int node_num_dims_;
    void *node_coordinates_;
    int coordinate_size_;
    int map_idx_;
    int reduce_idx_;
    int num_procs_;
    int current_iter_;
    int num_iters_;
    int parameter_size_;
    void *parameter_;

    IrregularRuntime(IRIndex num_edges, IRIndex num_nodes, EDGE *edges, void *edge_data, void *node_data,
                     int edge_data_elm_size, int node_data_elm_size, int reduction_elm_size, int node_num_dims,
                     void *node_coordinates, int coordinate_size, void *parameter, int parameter_size, int num_procs,
                     int num_iters)
        : global_num_edges_(num_edges), global_num_nodes_(num_nodes), global_edges_(edges),
          global_edge_data_(edge_data), global_node_data_(node_data), edge_data_elm_size_(edge_data_elm_size),
          node_data_elm_size_(node_data_elm_size), reduction_elm_size_(reduction_elm_size),
          node_num_dims_(node_num_dims), node_coordinates_(node_coordinates), coordinate_size_(coordinate_size),
          map_idx_(0), reduce_idx_(0), num_procs_(num_procs), current_iter_(0), num_iters_(num_iters),
          parameter_size_(parameter_size), parameter_(parameter), handle_(nullptr)
    {
    }

    ~IrregularRuntime()
    {
        if (handle_) pat_ir_destroy(handle_);
    }

    // Uploads the edge list and builds the by-destination CSR on the GPU
    // (the reference's reorder + partition + halo set-up).
    void IrregularInit()
    {
        if (pat_app_ir_table) {
            PAT_API_CALL(pat_ir_create(global_num_edges_, global_num_nodes_, (const int32_t *)global_edges_,
                                       PAT_IR_NO_FUNCTOR, nullptr, 0, &handle_));
            PAT_API_CALL(pat_ir_set_table(handle_, pat_app_ir_table()));
            PAT_API_CALL(pat_ir_set_map_idx(handle_, map_idx_));
            PAT_API_CALL(pat_ir_set_reduce_idx(handle_, reduce_idx_));
            PAT_API_CALL(pat_ir_set_param(handle_, parameter_, parameter_size_));
        } else {
            PAT_API_CALL(pat_ir_create(global_num_edges_, global_num_nodes_, (const int32_t *)global_edges_,
                                       library_functor(map_idx_), parameter_, parameter_size_, &handle_));
        }
        if (global_node_data_) reset_node_data(global_node_data_);
    }

    void set_map_idx(int idx)
    {
        map_idx_ = idx;
        if (handle_) {
            PAT_API_CALL(pat_ir_set_map_idx(handle_, pat_app_ir_table ? idx : library_functor(idx)));
            PAT_API_CALL(pat_ir_set_param(handle_, parameter_, parameter_size_));
        }
    }
    void set_reduce_idx(int idx)
    {
        reduce_idx_ = idx;
        if (handle_ && pat_app_ir_table) PAT_API_CALL(pat_ir_set_reduce_idx(handle_, idx));
    }

    // One pass of the edge functor over every edge, reducing into nodes.
    void IrregularStart()
    {
        if (!handle_) fail("IrregularStart before IrregularInit");
        PAT_API_CALL(pat_ir_start(handle_));
        current_iter_++;
    }

    void IrregularBarrier() { PAT_API_CALL(pat_device_sync()); }

    // VALUE per node in the caller's order.
    void get_reduction_result(void *buffer) { PAT_API_CALL(pat_ir_get_result(handle_, (float *)buffer)); }

    // New node coordinates (node_data_elm_size bytes per node: 12 = float3,
    // 24 = double3 as moldyn's point_data).
    void reset_node_data(void *node_data)
    {
        PAT_API_CALL(pat_ir_reset_node_data(handle_, node_data, node_data_elm_size_));
    }

  private:
    static int library_functor(int idx)
    {
        const int table[] = PAT_IR_MAP_TABLE;
        if (idx < 0 || idx >= (int)(sizeof(table) / sizeof(table[0]))) fail("map index not in the table");
        if (table[idx] < 0) fail("this map has no library kernel: link the application's device table");
        return table[idx];
    }
    static void fail(const char *msg)
    {
        fprintf(stderr, "pattern error: IrregularRuntime: %s\n", msg);
        exit(EXIT_FAILURE);
    }
    pat_ir *handle_;
};

#endif
