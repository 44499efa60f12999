// Stencil kernel launchers shared by the device-level entry point and the
// handle runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pat {

struct HeatTile {
    int fused;           // 1: two sweeps per launch (temporal blocking)
    int bx, by, ry, zc;  // plain: x threads, thread rows, rows per thread, z-chunk; fused: 64, 4, ry, zc
};

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

}  // namespace pat
