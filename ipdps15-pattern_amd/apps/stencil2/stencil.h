// The W = 3 star of the reference's kmeans directory (kmeans/stencil.h:3,
// 93-163, its stencil_function_table at :251), in the 4-argument stencil
// shape (kmeans/stencil.h:32), compiled for the host (CPU_KERNEL, FCPU) or
// the device (GPU_KERNEL, FGPU; see gpu_table.hip):
//   0  star     FP32: the sum of the cells 1..W-1 away along each axis,
//               taken x+, x-, y+, y-, z+, z- (each side by increasing
//               distance) into a sum started at 0, divided by 36.  The
//               halo is W cells wide; the functor reads W-1 of them.
//   1  cross2d  FP32 on a 2-D grid with a halo of 1 (not in the reference:
//               the runtime's 2-D path): ((x+ + x-) + (y+ + y-)) * 0.25.
#ifndef STENCIL
#define STENCIL
#define W 3

#include "pattern_api.h"

#if defined(GPU_KERNEL)
#define DEVICE __device__
#define CONSTANT __device__ constexpr
namespace FGPU {
#elif defined(CPU_KERNEL)
#define DEVICE inline
#define CONSTANT static constexpr
namespace FCPU {
#endif

typedef void (*stencil_fp)(void *input, void *output, int *offset, int *size);

DEVICE void star(void *input, void *output, int *offset, int *size)
{
    const int k = offset[0], j = offset[1], i = offset[2];
    float t = 0.0f;
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i, j, k + m);
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i, j, k - m);
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i, j + m, k);
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i, j - m, k);
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i + m, j, k);
    for (int m = 1; m < W; m++) t += GET_FLOAT3(input, i - m, j, k);
    t /= 36;
    GET_FLOAT3(output, i, j, k) = t;
}

DEVICE void cross2d(void *input, void *output, int *offset, int *size)
{
    const int k = offset[0], j = offset[1];
    const float a = GET_FLOAT3(input, 0, j, k + 1) + GET_FLOAT3(input, 0, j, k - 1);
    const float b = GET_FLOAT3(input, 0, j + 1, k) + GET_FLOAT3(input, 0, j - 1, k);
    GET_FLOAT3(output, 0, j, k) = (a + b) * 0.25f;
}

CONSTANT stencil_fp stencil_function_table[] = {star, cross2d};

}  // namespace FCPU / FGPU
#endif
