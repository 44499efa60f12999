// GridMPI, the per-process grid the application fills before StencilBegin
// (kmeans/lib/grid_mpi.h:148-155; heat_3D/main.cpp:19-49 writes data()).
// With one GPU holding the whole domain the process grid is the global grid:
// data() is the interior, dimension 0 fastest (the GET_FLOAT3 convention,
// kmeans/lib/macro.h:70); the width-w halo is the fixed Dirichlet zero.
#ifndef LIB_GRID_MPI_H_
#define LIB_GRID_MPI_H_

#include <vector>

#include "array.h"

class GridMPI {
  public:
    GridMPI(int num_dims, int unit_size, const IndexArray &size)
        : num_dims_(num_dims), unit_size_(unit_size), size_(size),
          data_((size_t)size.accumulate(num_dims) * (size_t)unit_size, 0)
    {
    }
    int num_dims() const { return num_dims_; }
    int unit_size() const { return unit_size_; }
    const IndexArray &my_size() const { return size_; }
    const IndexArray &global_size() const { return size_; }
    void *data() { return data_.data(); }
    size_t bytes() const { return data_.size(); }

  private:
    int num_dims_;
    int unit_size_;
    IndexArray size_;
    std::vector<unsigned char> data_;
};

#endif
