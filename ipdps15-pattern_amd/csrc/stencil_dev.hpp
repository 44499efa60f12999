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
