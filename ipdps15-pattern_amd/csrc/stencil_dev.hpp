// Stencil kernel launchers shared by the device-level entry point and the
// handle runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pat {

struct HeatTile {
    int fused;           // 0: one sweep per launch; 1: two (fused window); 2: ry sweeps on LDS cubes
    int bx, by, ry, zc;  // plain: x threads, thread rows, rows per thread, z-chunk; fused 1: 64, 4, ry, zc;
                         // fused 2: cube shape index, 0, sweeps per launch, 0
};

// One grid of a fused batch: an entry of the device table that the batched
// two-sweep kernel reads (one table per direction, a -> b and b -> a).
struct HeatBatchGrid {
    const float *in;
    float *out;
    int nx, ny, nz, zc;
    int tiles_x, tiles_y;
    int first;  // first workgroup of this grid in the launch
    int pad;
};

// Fused sweeps of n grids.  heat3d_batch_table writes the two tables (2n
// entries at 'table', device memory) for grids a[i] <-> b[i] and returns
// the workgroups of one launch; heat3d_batch_enqueue then queues nsteps
// sweeps of every grid from a[i] -- two per launch over all of them, an odd
// last sweep one launch per grid -- and sets *in_b when the results end in
// b (the same for every grid).  The enqueue makes no allocation and no
// memset, so it may be captured into a graph.
int heat3d_batch_table(int n, float *const *a, float *const *b, const int64_t *nx, const int64_t *ny,
                       const int64_t *nz, HeatBatchGrid *table, hipStream_t st, int *nblocks);
int heat3d_batch_enqueue(int n, float *const *a, float *const *b, const int64_t *nx, const int64_t *ny,
                         const int64_t *nz, float alpha, int nsteps, const HeatBatchGrid *table, int nblocks,
                         hipStream_t st, int *in_b);

// Loads the stencil kernels' code object now rather than at a first launch.
int stencil_preload();

// Sweeps that one launch of tile t advances.
inline int heat3d_steps_per_launch(const HeatTile &t) { return t.fused == 0 ? 1 : (t.fused == 1 ? 2 : t.ry); }

int heat3d_sweep(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int variant, int bx,
                 int by, int ry, int zc, hipStream_t stream);
int heat3d_sweep2(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int ry, int zc,
                  hipStream_t stream);
int heat3d_launch(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, const HeatTile &t,
                  int variant, hipStream_t stream);
int heat3d_steps(float *const buf[2], int start, int64_t nx, int64_t ny, int64_t nz, float alpha, const HeatTile &t,
                 int variant, int nsteps, hipStream_t stream, int *end);
int heat3d_tuned(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, hipStream_t st,
                 bool allow_fused, HeatTile *t);
void heat3d_default_tile(int64_t nx, int64_t ny, int64_t nz, int *bx, int *by, int *ry, int *zc);
int heat3d_shapes(const int **triples);
// The tile of LDS cube shape 'shape' (PAT_STENCIL_CUBES); PAT_ERR_UNSUPPORTED
// for an unknown shape.
int heat3d_cube_tile(int shape, HeatTile *t);

}  // namespace pat
