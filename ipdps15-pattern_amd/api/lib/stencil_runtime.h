// StencilRuntime (kmeans/lib/stencil_runtime.h:13-93,
// heat_3D/lib/stencil_runtime.h:16): same constructor and calls, running
// the stencil functor of the table index on the GPU through libpattern.
// StencilBegin() runs num_iters ping-pong sweeps and leaves the result in
// grid()->data().
//
// An application that links its functor table compiled for the device
// (api/gpu/stencil_table.hpp, which defines pat_app_stencil_table) has it
// registered at StencilInit, so its own functors run on the GPU on grids of
// any halo width and element type; without one the library's heat engine
// runs (3-D FP32, width 1).
#ifndef LIB_STENCIL_RUNTIME_H_
#define LIB_STENCIL_RUNTIME_H_

#include <cstring>
#include <vector>

#include "../pat_check.h"
#include "array.h"
#include "grid_mpi.h"

// Defined by an application's device functor table when it is linked in.
extern "C" const pat_stencil_table *pat_app_stencil_table(void) __attribute__((weak));

class StencilRuntime {
  public:
    StencilRuntime(int num_dims, int unit_size, const IndexArray &global_size, int proc_num_dims, IntArray proc_size,
                   int stencil_width, int num_iters, void *parameter, size_t parameter_size)
        : num_dims_(num_dims), unit_size_(unit_size), global_size_(global_size), proc_num_dims_(proc_num_dims),
          proc_size_(proc_size), stencil_width_(stencil_width), num_iters_(num_iters), parameter_(parameter),
          parameter_size_(parameter_size), grid_(nullptr), handle_(nullptr), stencil_idx_(0)
    {
    }
    ~StencilRuntime()
    {
        if (handle_) pat_stencil_destroy(handle_);
        delete grid_;
    }

    // The reference builds the MPI process grid here; one process owns the
    // whole domain.
    void StencilInit()
    {
        grid_ = new GridMPI(num_dims_, unit_size_, global_size_);
        int64_t interior[3] = {global_size_[0], num_dims_ > 1 ? global_size_[1] : 1,
                               num_dims_ > 2 ? global_size_[2] : 1};
        PAT_API_CALL(pat_stencil_create(num_dims_ == 2 ? 2 : 3, interior, stencil_width_, unit_size_, parameter_,
                                        (int64_t)parameter_size_, &handle_));
        if (pat_app_stencil_table) PAT_API_CALL(pat_stencil_set_table(handle_, pat_app_stencil_table()));
        PAT_API_CALL(pat_stencil_set_idx(handle_, stencil_idx_));
    }

    // The interior (grid()->data(), dimension 0 fastest) goes into a
    // halo-inclusive grid with a zero halo of the stencil width, and back.
    void StencilBegin()
    {
        const int64_t w = stencil_width_, u = unit_size_;
        const int64_t nx = global_size_[0], ny = num_dims_ > 1 ? global_size_[1] : 1,
                      nz = num_dims_ > 2 ? global_size_[2] : 1;
        const int64_t sx = nx + 2 * w, sy = ny + 2 * w, sz = num_dims_ > 2 ? nz + 2 * w : 1, zw = num_dims_ > 2 ? w : 0;
        std::vector<unsigned char> halo((size_t)(sx * sy * sz * u), 0);
        const unsigned char *in = (const unsigned char *)grid_->data();
        for (int64_t z = 0; z < nz; z++)
            for (int64_t y = 0; y < ny; y++)
                memcpy(&halo[(size_t)((((z + zw) * sy + (y + w)) * sx + w) * u)], in + (z * ny + y) * nx * u,
                       (size_t)(nx * u));
        PAT_API_CALL(pat_stencil_upload(handle_, halo.data()));
        PAT_API_CALL(pat_stencil_run(handle_, num_iters_));
        PAT_API_CALL(pat_stencil_download(handle_, halo.data()));
        unsigned char *out = (unsigned char *)grid_->data();
        for (int64_t z = 0; z < nz; z++)
            for (int64_t y = 0; y < ny; y++)
                memcpy(out + (z * ny + y) * nx * u, &halo[(size_t)((((z + zw) * sy + (y + w)) * sx + w) * u)],
                       (size_t)(nx * u));
    }

    void StencilFinalize() { PAT_API_CALL(pat_device_sync()); }

    int num_dims() { return num_dims_; }
    int unit_size() { return unit_size_; }
    int stencil_width() { return stencil_width_; }
    GridMPI *grid() { return grid_; }
    int my_rank() { return 0; }
    int along() { return num_dims_ - 1; }
    void set_stencil_idx(int idx)
    {
        stencil_idx_ = idx;
        if (handle_) PAT_API_CALL(pat_stencil_set_idx(handle_, idx));
    }
    int get_stencil_idx() { return stencil_idx_; }
    void *get_parameter() { return parameter_; }

  protected:
    int num_dims_;
    int unit_size_;
    IndexArray global_size_;
    int proc_num_dims_;
    IntArray proc_size_;
    int stencil_width_;
    int num_iters_;
    void *parameter_;
    size_t parameter_size_;
    GridMPI *grid_;
    pat_stencil *handle_;
    int stencil_idx_;
};

#endif
